"""Mie series for a homogeneous sphere (Bohren & Huffman, "Absorption and
Scattering of Light by Small Particles", ch. 4): the checker for bistatic
radar cross sections computed through TFSF + ADE + NF2FF.

Test infrastructure only (scipy spherical Bessel functions)."""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _psi(n, z, derivative=False):
    j = spherical_jn(n, z)
    if not derivative:
        return z * j
    return j + z * spherical_jn(n, z, derivative=True)


def _xi(n, z, derivative=False):
    h = spherical_jn(n, z) + 1j * spherical_yn(n, z)
    if not derivative:
        return z * h
    dh = spherical_jn(n, z, derivative=True) + 1j * spherical_yn(n, z, derivative=True)
    return h + z * dh


def coefficients(m, x, nmax=None):
    """a_n, b_n for relative index m (exp(-i w t) convention: Im m >= 0 for
    loss) and size parameter x = k a."""
    if nmax is None:
        nmax = int(np.ceil(x + 4 * x ** (1 / 3) + 2)) + 4
    n = np.arange(1, nmax + 1)
    mx = m * x
    px, dpx = _psi(n, x), _psi(n, x, True)
    pm, dpm = _psi(n, mx), _psi(n, mx, True)
    xx, dxx = _xi(n, x), _xi(n, x, True)
    a = (m * pm * dpx - px * dpm) / (m * pm * dxx - xx * dpm)
    b = (pm * dpx - m * px * dpm) / (pm * dxx - m * xx * dpm)
    return n, a, b


def amplitudes(m, x, theta):
    """Scattering amplitudes S1 (E perpendicular to the scattering plane)
    and S2 (parallel) at scattering angles theta."""
    n, a, b = coefficients(m, x)
    mu = np.cos(np.atleast_1d(theta))
    S1 = np.zeros(mu.shape, complex)
    S2 = np.zeros(mu.shape, complex)
    pi_prev, pi_cur = np.zeros_like(mu), np.ones_like(mu)
    for k, nn in enumerate(n):
        tau = nn * mu * pi_cur - (nn + 1) * pi_prev
        f = (2 * nn + 1) / (nn * (nn + 1))
        S1 += f * (a[k] * pi_cur + b[k] * tau)
        S2 += f * (a[k] * tau + b[k] * pi_cur)
        pi_prev, pi_cur = pi_cur, ((2 * nn + 1) * mu * pi_cur - (nn + 1) * pi_prev) / nn
    return S1, S2


def bistatic_rcs(eps, radius, freq, theta, perpendicular=True):
    """sigma(theta) = 4 pi |S|^2 / k^2 (m^2) of a sphere of complex relative
    permittivity eps (exp(-i w t) convention) in vacuum."""
    c0 = 299792458.0
    k = 2 * np.pi * freq / c0
    m = np.sqrt(complex(eps))
    if m.imag < 0:
        m = -m
    S1, S2 = amplitudes(m, k * radius, theta)
    S = S1 if perpendicular else S2
    return 4 * np.pi * np.abs(S) ** 2 / k ** 2


def drude_eps(freq, wp, gamma):
    """Drude permittivity in the exp(-i w t) convention."""
    w = 2 * np.pi * freq
    return 1.0 - wp * wp / (w * w + 1j * gamma * w)


def lorentz_eps(freq, eps_inf, wp, w0, gamma):
    """Single-pole Lorentz permittivity (exp(-i w t) convention), delta eps
    = wp^2 / w0^2."""
    w = 2 * np.pi * freq
    return eps_inf + wp * wp / (w0 * w0 - w * w - 1j * gamma * w)


def material_eps(freq, m):
    """Permittivity of a material record {eps_inf, wp, w0, gamma}."""
    if m["w0"] > 0:
        return lorentz_eps(freq, m["eps_inf"], m["wp"], m["w0"], m["gamma"])
    return m["eps_inf"] - m["wp"] ** 2 / ((2 * np.pi * freq) ** 2 + 1j * m["gamma"] * 2 * np.pi * freq)
