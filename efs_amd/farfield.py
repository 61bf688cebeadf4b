"""Far-field reporting from the engine's NF2FF accumulation (FP64 numpy).

Restates the reference worker's reporting step:

  rad_power      surface Poynting flux of the accumulated fields,
                 P = sum 0.5 * area * sign * (E_nP H_nPP* - E_nPP H_nP*)
                 over the surface points (app/sim-worker.js:650-673)
  directivity    Dmax = 4 pi r^2 max(U) / P from the far field's radiation
                 intensity U per angle (the first of the 5 values per angle
                 of computeNF2FFfarField), the |E| pattern and its dBi map
                 (app/sim-worker.js:996-1048)
  cuts           the xz (phi = 0) and yz (phi = 90 deg) cuts in dBi over
                 theta in [-180, 180) with linear interpolation on the theta
                 grid (app/sim-worker.js:1020-1040)

`sphere_grid` makes the reference's 2-degree theta/phi grid.

`circular` gives the circular-polarisation components of
src/nf2ff.mjs:554-573.

For scatterers under a plane wave (no counterpart in the reference's
worker): `incident_spectrum` and `rcs` turn the far field of the
scattered-field NF2FF surface into a bistatic radar cross section.
"""
import numpy as np


def sphere_grid(theta_step=2, phi_step=2):
    """theta 0..180, phi 0..360 degrees (inclusive) in radians."""
    theta = np.deg2rad(np.arange(0, 180 + theta_step / 2, theta_step))
    phi = np.deg2rad(np.arange(0, 360 + phi_step / 2, phi_step))
    return theta, phi


def rad_power(accE, accH, point_meta, num_points):
    """Radiated power through the NF2FF surface from one frequency's
    accumulators (6 floats per point: re/im of 3 components)."""
    E = np.asarray(accE, np.float64)[:6 * num_points].reshape(num_points, 3, 2)
    H = np.asarray(accH, np.float64)[:6 * num_points].reshape(num_points, 3, 2)
    m = np.asarray(point_meta, np.float64)[:8 * num_points].reshape(num_points, 8)
    n = m[:, 3].astype(np.int64)
    nP, nPP = (n + 1) % 3, (n + 2) % 3
    k = np.arange(num_points)
    s = (E[k, nP, 0] * H[k, nPP, 0] + E[k, nP, 1] * H[k, nPP, 1]) - \
        (E[k, nPP, 0] * H[k, nP, 0] + E[k, nPP, 1] * H[k, nP, 1])
    return float(np.sum(0.5 * m[:, 5] * s * m[:, 4]))


def directivity(ff, n_theta, n_phi, power, radius=1.0):
    """Dmax (linear and dBi), the |E| pattern and the dBi map of a far-field
    result (5 values per angle, theta-major)."""
    f = np.asarray(ff, np.float64).reshape(n_theta * n_phi, 5)
    U = f[:, 0]
    e_norm = np.sqrt(f[:, 1] ** 2 + f[:, 2] ** 2 + f[:, 3] ** 2 + f[:, 4] ** 2)
    dmax = U.max() * 4 * np.pi * radius * radius / power if power > 0 else 0.0
    dmax_dbi = 10 * np.log10(max(dmax, 1e-15))
    emax = e_norm.max()
    d_dbi = 20 * np.log10(np.maximum(e_norm / emax, 1e-15)) + dmax_dbi if emax > 0 else np.full_like(e_norm, dmax_dbi)
    return {"Dmax": dmax, "DmaxdBi": dmax_dbi, "E_norm": e_norm, "directivity_dBi": d_dbi}


def cuts(e_norm, dmax_dbi, n_theta, n_phi, theta_step=2, phi_step=2):
    """xz and yz cuts (dBi) at theta = -180..178 degrees; negative theta
    reads the opposite half-plane (phi + 180)."""
    e = np.asarray(e_norm, np.float64)
    emax = e.max()
    theta_deg = np.arange(-180, 180, 2)
    phi90 = int(round(90 / phi_step))
    xz, yz = [], []
    for t in theta_deg:
        frac_i = abs(t) / theta_step
        i0 = min(int(np.floor(frac_i)), n_theta - 1)
        i1 = min(i0 + 1, n_theta - 1)
        w = frac_i - i0
        off = int(round(180 / phi_step)) if t < 0 else 0
        px, py = off % n_phi, (phi90 + off) % n_phi
        vx = e[i0 * n_phi + px] * (1 - w) + e[i1 * n_phi + px] * w
        vy = e[i0 * n_phi + py] * (1 - w) + e[i1 * n_phi + py] * w
        xz.append(20 * np.log10(max(vx / emax, 1e-15)) + dmax_dbi)
        yz.append(20 * np.log10(max(vy / emax, 1e-15)) + dmax_dbi)
    return theta_deg, np.array(xz), np.array(yz)


def intensity_power(ff, theta, phi):
    """Radiated power by integrating U over the sphere (trapezoid in theta
    and phi on the sphere_grid angles) -- a check on rad_power: both are the
    same power for a closed surface."""
    U = np.asarray(ff, np.float64).reshape(len(theta), len(phi), 5)[:, :, 0]
    g = U * np.sin(theta)[:, None]
    return float(np.trapezoid(np.trapezoid(g, phi, axis=1), theta))


def incident_spectrum(signal, omega, dT, e0=1.0):
    """|E_inc| at each angular frequency with the NF2FF accumulation's DFT
    (sum over steps of field x exp(-j w t) dT): the plane wave at any point
    is the source signal delayed, which changes only the phase."""
    s = np.asarray(signal, np.float64)
    k = np.arange(len(s))
    om = np.atleast_1d(np.asarray(omega, np.float64))
    return np.abs(np.exp(-1j * np.outer(om, k) * dT) @ s) * dT * abs(e0)


def rcs(ff, e_inc):
    """Bistatic radar cross section (m^2) per angle from a far-field result
    (5 values per angle, U first, radius 1) and the incident |E| at the same
    frequency: sigma = 4 pi U / (|E_inc|^2 / (2 Z0))."""
    U = np.asarray(ff, np.float64).reshape(-1, 5)[:, 0]
    z0 = 376.73031346177066
    return 4 * np.pi * U / (e_inc * e_inc / (2 * z0))


def circular(ff):
    """Circular-polarisation magnitudes per angle from a far-field result
    (5 values per angle), with the reference's labels (src/nf2ff.mjs:
    554-573): E_cprh = |E_theta - j E_phi| / sqrt(2), E_cplh =
    |E_theta + j E_phi| / sqrt(2)."""
    f = np.asarray(ff, np.float64).reshape(-1, 5)
    et = f[:, 1] + 1j * f[:, 2]
    ep = f[:, 3] + 1j * f[:, 4]
    return np.abs(et - 1j * ep) / np.sqrt(2.0), np.abs(et + 1j * ep) / np.sqrt(2.0)
