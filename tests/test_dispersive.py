"""Plane-wave source and dispersive materials of the FP64 builder
(efs_amd/operator.py tfsf_plane_wave, ade_orders), run on the FP64 oracle,
against closed forms:

- TFSF: in an empty UPML box the total-field region carries the plane wave
  (peak |E| dl = e0 dl) and the scattered-field shell stays below 0.3% of it,
  for +x, -y and +z propagation (a wrong sign, index or half-step in the
  point lists leaves a wave of order 1 outside the box);
- ADE (openEMS write-back form): a 10 mm Lorentz (eps_inf 2, delta eps 3,
  6 GHz pole) or Drude (12 GHz plasma frequency) slab under a normally
  incident plane wave, quasi-1-D (periodic in y and z, Mur in x), transmits
  T = (1 - r^2) exp(-j (n - 1) k0 d) / (1 - r^2 exp(-2 j n k0 d)),
  r = (1 - n) / (1 + n), n = sqrt(eps(w)), measured as the ratio of probe
  spectra with and without the slab: within 0.1 dB and 1 degree away from
  the pole;
- the stability limit: ade_orders refuses a dT above Mesh.timestep(wp_max);
- the Mie checker itself (tests/mie.py): optical theorem and the Rayleigh
  limit.
"""
import numpy as np
import pytest

from efs_amd import operator as op

import mie


def _tfsf_run(orc, prop_axis, prop_sign, e_axis):
    n, pml = 40, 6
    mesh = op.uniform_mesh((n, n, n), 1e-3)
    dT = mesh.timestep()
    mat = op.Materials(mesh)
    op.pec_boundary(mat)
    co = op.build_coefficients(mesh, mat, dT)
    regions = op.build_upml(mesh, co, [pml] * 6, dT)
    sig = op.gaussian_pulse(8e9, 6e9, dT)
    lo, hi = [12] * 3, [27] * 3
    cfg = op.tfsf_plane_wave(mesh, co, lo, hi, dT, sig, prop_axis, prop_sign, e_axis)
    o = orc.Oracle([n] * 3, op.to_f32(co), fp64=True)
    o.configurePML(regions)
    o.configureTFSF(cfg)
    tf = np.zeros((n, n, n), bool)
    tf[lo[0] + 1:hi[0], lo[1] + 1:hi[1], lo[2] + 1:hi[2]] = True
    sf = np.zeros((n, n, n), bool)
    sf[pml + 1:n - pml - 1, pml + 1:n - pml - 1, pml + 1:n - pml - 1] = True
    sf[lo[0] - 1:hi[0] + 2, lo[1] - 1:hi[1] + 2, lo[2] - 1:hi[2] + 2] = False
    peak_tf = peak_sf = 0.0
    for _ in range(12):
        o.iterate(20)
        V = o.getFields()["volt"].reshape(3, n, n, n)
        a = np.abs(V).max(axis=0)
        peak_tf = max(peak_tf, a[tf].max())
        peak_sf = max(peak_sf, a[sf].max())
    return peak_tf, peak_sf


@pytest.mark.parametrize("prop_axis,prop_sign,e_axis", [(0, 1, 2), (1, -1, 0), (2, 1, 1)])
def test_tfsf_plane_wave(oracle_mod, prop_axis, prop_sign, e_axis):
    peak_tf, peak_sf = _tfsf_run(oracle_mod, prop_axis, prop_sign, e_axis)
    assert abs(peak_tf - 1e-3) < 2e-5  # e0 = 1 V/m on 1 mm edges
    assert peak_sf < 3e-3 * peak_tf


def _slab_trace(orc, kind, slab):
    nx, nt, d = 400, 6, 0.5e-3
    mesh = op.uniform_mesh((nx, nt, nt), d)
    poles = []
    mat = op.Materials(mesh)
    if slab:
        mask = np.zeros(mat.eps.shape, bool)
        mask[200:220] = True
        if kind == "lorentz":
            mat.eps[mask] = 2.0
            w0 = 2 * np.pi * 6e9
            poles = [{"mask": mask, "wp": w0 * np.sqrt(3.0), "w0": w0, "gamma": 1e9}]
        else:
            poles = [{"mask": mask, "wp": 2 * np.pi * 12e9, "w0": 0.0, "gamma": 1e9}]
    dT = mesh.timestep(wp_max=2 * np.pi * 12e9)
    co = op.build_coefficients(mesh, mat, dT)
    sig = op.gaussian_pulse(7e9, 6e9, dT)
    # total field over all of y and z (periodic), from x = 20 to nx - 21
    cfg = op.tfsf_plane_wave(mesh, co, [20, -1, -1], [nx - 21, nt, nt], dT, sig, 0, 1, 2)
    o = orc.Oracle(mesh.shape, op.to_f32(co), fp64=True)
    o.configureMur(op.mur_config(mesh, [0, 1], dT))
    o.configurePBC({"axes": [{"axis": 1}, {"axis": 2}]})
    o.configureTFSF(cfg)
    if poles:
        o.configureLorentz(op.ade_orders(mesh, co, dT, poles), apply=True)
    N = nx * nt * nt
    steps = 8000
    o.configureProbeGatherBuffered([2 * N + (300 * nt + 2) * nt + 2], steps, 1)
    o.iterate(steps)
    return o.readProbeGatherBuffered()[:, 0], dT, 20 * d


@pytest.mark.parametrize("kind", ["lorentz", "drude"])
def test_ade_slab_transmission(oracle_mod, kind):
    a, dT, L = _slab_trace(oracle_mod, kind, True)
    b, _, _ = _slab_trace(oracle_mod, kind, False)
    f = np.fft.rfftfreq(len(a), dT)
    A, B = np.fft.rfft(a), np.fft.rfft(b)
    for fg in (2, 3, 4, 9, 10, 11, 12):
        i = int(np.argmin(abs(f - fg * 1e9)))
        w = 2 * np.pi * f[i]
        if kind == "lorentz":
            w0 = 2 * np.pi * 6e9
            eps = 2.0 + 3.0 * w0 ** 2 / (w0 ** 2 - w ** 2 + 1j * 1e9 * w)  # exp(+j w t)
        else:
            wp = 2 * np.pi * 12e9
            eps = 1.0 - wp ** 2 / (w ** 2 - 1j * 1e9 * w)
        n = np.sqrt(complex(eps))
        if n.imag > 0:
            n = -n
        k0 = w / op.C0
        r = (1 - n) / (1 + n)
        T = (1 - r * r) * np.exp(-1j * (n - 1) * k0 * L) / (1 - r * r * np.exp(-2j * n * k0 * L))
        m = A[i] / B[i]
        assert abs(20 * np.log10(abs(m) / abs(T))) < 0.1, (fg, m, T)
        assert abs(np.angle(m / T, deg=True)) < 1.0, (fg, m, T)


def test_ade_stability_guard():
    mesh = op.uniform_mesh((8, 8, 8), 1e-3)
    mat = op.Materials(mesh)
    dT = mesh.timestep()
    co = op.build_coefficients(mesh, mat, dT)
    mask = np.zeros(mat.eps.shape, bool)
    mask[2:5, 2:5, 2:5] = True
    pole = [{"mask": mask, "wp": 2 * np.pi * 12e9, "w0": 0.0, "gamma": 1e9}]
    with pytest.raises(ValueError):
        op.ade_orders(mesh, co, dT, pole)
    dT2 = mesh.timestep(wp_max=2 * np.pi * 12e9)
    co2 = op.build_coefficients(mesh, mat, dT2)
    cfg = op.ade_orders(mesh, co2, dT2, pole)
    o = cfg["orders"][0]
    assert not o["hasLorentz"] and len(o["directions"]) == 3
    # edges of the filled 3x3x3 block and its surface: 4x4x3 per direction
    assert o["numCells"] == len(o["directions"][0]["pos_idx"])
    for d in o["directions"]:
        assert np.count_nonzero(d["v_ext_ADE"]) == 4 * 4 * 3


def test_mie_checker():
    # optical theorem: Q_ext = (4 / x^2) Re S(0) = Q_sca for a real index
    for m, x in ((2.0 + 0j, 2.0), (1.5 + 0j, 5.0)):
        n, a, b = mie.coefficients(m, x)
        qe = 2 / x ** 2 * np.sum((2 * n + 1) * np.real(a + b))
        qs = 2 / x ** 2 * np.sum((2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2))
        S1, S2 = mie.amplitudes(m, x, np.array([0.0]))
        assert abs(qe - qs) < 1e-9 * qe
        assert abs(4 / x ** 2 * S1.real[0] - qe) < 1e-9 * qe
        assert abs(S1[0] - S2[0]) < 1e-12 * abs(S1[0])
    # Rayleigh limit: backscatter 4 pi k^4 a^6 |(eps - 1) / (eps + 2)|^2
    c0 = 299792458.0
    k = 0.02
    s = mie.bistatic_rcs(4.0, 1.0, k * c0 / (2 * np.pi), np.array([np.pi]))[0]
    assert abs(s / (4 * np.pi * k ** 4 * (3 / 6) ** 2) - 1) < 1e-3
