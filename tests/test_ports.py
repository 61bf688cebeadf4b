"""Probe -> port -> S11 chain (efs_amd/ports.py).

Pins:
- the reference's own test on openEMS's coax traces (tests/test_api.mjs:
  573-660): |Z| = |U/I| within [-3 %, +6 %] of the analytic 49.94 ohm over
  0.1-0.5 GHz and S11 < -10 dB against 50 ohm;
- dft_time2freq against closed forms: a sampled exponential is a geometric
  series, exp(-a t) -> dt / (1 - exp(-(a + j w) dt)) (pins the -j sign and
  the dt weight), and the first-sample weight rule on a non-uniform grid;
- calc_port against its circuit identities: P_inc - P_ref = P_acc, matched
  (U = Z I) -> no reflection, open -> S11 = 1, short -> S11 = -1, and
  uf_ref/uf_inc equal to compute_s11's (Z - Z0)/(Z + Z0);
- build_gather_indices: a current probe's loop sum equals the sum of the
  discrete curls of the enclosed edges (discrete Stokes, the curl being the
  one the voltage update applies), for all three normal directions, and a
  loop whose legs leave the grid drops exactly those legs.
"""
import json
import os

import numpy as np
import pytest

from efs_amd import ports

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_coax_fixture_impedance_and_s11():
    ref = json.load(open(os.path.join(GOLDEN, "coax_openems.json")))
    ut = (ref["probe_ut1"]["time_s"], ref["probe_ut1"]["voltage"])
    it = (ref["probe_it1"]["time_s"], ref["probe_it1"]["voltage"])
    freqs = np.linspace(0.1e9, 0.5e9, 200)
    ur, ui = ports.dft_time2freq(*ut, freqs)
    ir, ii = ports.dft_time2freq(*it, freqs)
    Z = np.abs((ur + 1j * ui) / (ir + 1j * ii))
    z0 = ref["Z0_analytical_ohm"]
    tol = ref["tolerances"]
    assert np.all(Z <= z0 * (1 + tol["upper_error"])) and np.all(Z >= z0 * (1 - tol["lower_error"])), \
        (Z.min(), Z.max())
    p = ports.calc_port(ur + 1j * ui, ir + 1j * ii, 50.0)
    s11 = np.abs(p["uf_ref"] / p["uf_inc"])
    assert 20 * np.log10(s11.max()) < -10
    # the same through compute_s11's impedance form
    r = ports.s11_at(ut, it, freqs)
    np.testing.assert_allclose(r["s11_dB"], 20 * np.log10(s11), atol=1e-9)


def test_dft_geometric_series():
    dt, a = 1e-11, 3e9
    t = np.arange(20000) * dt
    v = np.exp(-a * t)
    f = np.array([0.0, 0.3e9, 1e9, 4.2e9])
    re, im = ports.dft_time2freq(t, v, f)
    w = 2 * np.pi * f
    want = dt / (1 - np.exp(-(a + 1j * w) * dt))
    np.testing.assert_allclose(re + 1j * im, want, rtol=1e-9)


def test_dft_nonuniform_weights_and_periodic():
    t = np.array([0.0, 1.0, 3.0, 4.5])
    v = np.array([2.0, -1.0, 0.5, 3.0])
    f = np.array([0.0, 0.1])
    re, im = ports.dft_time2freq(t, v, f)
    # dt weights: first sample uses t1 - t0 = 1, then 1, 2, 1.5
    wts = np.array([1.0, 1.0, 2.0, 1.5])
    ph = 2 * np.pi * 0.1 * t
    assert re[0] == pytest.approx(np.sum(v * wts)) and im[0] == 0.0
    assert re[1] == pytest.approx(np.sum(v * wts * np.cos(ph)))
    assert im[1] == pytest.approx(-np.sum(v * wts * np.sin(ph)))
    re, im = ports.dft_time2freq(t, v, f, "periodic")
    assert re[0] == pytest.approx(np.sum(v))


def test_calc_port_identities():
    rng = np.random.default_rng(5)
    U = rng.normal(size=8) + 1j * rng.normal(size=8)
    I = rng.normal(size=8) + 1j * rng.normal(size=8)
    p = ports.calc_port(U, I, 50.0)
    np.testing.assert_allclose(p["P_inc"] - p["P_ref"], p["P_acc"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(p["uf_inc"] + p["uf_ref"], U, rtol=1e-12)
    np.testing.assert_allclose(p["if_inc"] - p["if_ref"], I, rtol=1e-12)
    m = ports.calc_port(50.0 * I, I, 50.0)
    np.testing.assert_allclose(m["uf_ref"], 0, atol=1e-12)
    assert np.all(m["P_ref"] == pytest.approx(0, abs=1e-20))
    o = ports.calc_port(U, np.zeros(8), 50.0)
    np.testing.assert_allclose(o["uf_ref"] / o["uf_inc"], 1.0)
    s = ports.calc_port(np.zeros(8), I, 50.0)
    np.testing.assert_allclose(s["uf_ref"] / s["uf_inc"], -1.0)
    # the port form and the impedance form of S11 agree
    Z = U / I
    np.testing.assert_allclose(p["uf_ref"] / p["uf_inc"], (Z - 50) / (Z + 50), rtol=1e-12)
    with pytest.raises(ValueError, match="Z_ref must be positive"):
        ports.calc_port(U, I, 0.0)


def curl(curr, n, x, y, z):
    """Discrete curl at the E edge (n; x,y,z) as the voltage update forms
    it: H_nPP - H_nPP[-nP] - H_nP + H_nP[-nPP]."""
    nP, nPP = (n + 1) % 3, (n + 2) % 3
    p = [x, y, z]

    def h(c, sh):
        q = list(p)
        if sh is not None:
            q[sh] -= 1
        return curr[c][tuple(q)]
    return h(nPP, None) - h(nPP, nP) - h(nP, None) + h(nP, nPP)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_current_loop_is_discrete_stokes(d):
    dims = (9, 8, 10)
    rng = np.random.default_rng(40 + d)
    curr = rng.normal(size=(3,) + dims)
    a, b = (d + 1) % 3, (d + 2) % 3
    s, t = [0, 0, 0], [0, 0, 0]
    s[a], t[a] = 2, 6
    s[b], t[b] = 1, 5
    s[d] = t[d] = 4
    probe = {"type": 1, "start": s, "stop": t, "normDir": d, "weight": 1.0}
    idx, sl = ports.build_gather_indices([probe], dims)
    N = int(np.prod(dims))
    flat = curr.ravel()
    gathered = np.zeros(2 * len(idx))
    gathered[1::2] = flat[idx.astype(np.int64)]
    got = ports.compute_probe_integrals(gathered, [probe], sl)[0]
    # the loop on dual lines s..t encloses the E_d edges at nodes s+1..t
    want = 0.0
    for i in range(s[a] + 1, t[a] + 1):
        for j in range(s[b] + 1, t[b] + 1):
            q = [0, 0, 0]
            q[a], q[b], q[d] = i, j, s[d]
            want += curl(curr, d, *q)
    assert got == pytest.approx(want, rel=1e-12, abs=1e-12)
    assert sl[0]["count"] == 2 * (t[a] - s[a]) + 2 * (t[b] - s[b])
    assert N == curr[0].size


def test_current_loop_drops_outside_legs():
    dims = (6, 6, 6)
    probe = {"type": 1, "start": [1, 1, 2], "stop": [3, 4, 2], "normDir": 2, "weight": 1.0,
             "startInside": [True, False, True], "stopInside": [True, True, True]}
    idx, sl = ports.build_gather_indices([probe], dims)
    # si[1] false drops the +x leg at the low-y line (2 edges); the rest stay
    assert sl[0]["count"] == 3 + 2 + 3
    assert list(sl[0]["signs"]) == [1.0] * 3 + [-1.0] * 2 + [-1.0] * 3


def test_voltage_probe_and_traces():
    dims = (5, 6, 7)
    N = int(np.prod(dims))
    probe = {"type": 0, "start": [2, 3, 5], "stop": [2, 3, 1], "weight": -1.0}
    cp = {"type": 1, "start": [1, 2, 2], "stop": [2, 3, 2], "normDir": 2, "weight": 1.0}
    idx, sl = ports.build_gather_indices([probe, cp], dims)
    assert list(idx[:4]) == [2 * N + 2 * 42 + 3 * 7 + z for z in (1, 2, 3, 4)]
    assert list(sl[0]["signs"]) == [-1.0] * 4
    rng = np.random.default_rng(3)
    data = rng.normal(size=(3, 2 * len(idx)))
    tr = ports.probe_traces(data.ravel(), 3, 2 * len(idx), [probe, cp], sl, 10, 2e-12)
    np.testing.assert_allclose(tr[0][0], np.array([10, 20, 30]) * 2e-12)
    np.testing.assert_allclose(tr[1][0], np.array([10.5, 20.5, 30.5]) * 2e-12)
    np.testing.assert_allclose(tr[0][1], data[:, 0:8:2].sum(axis=1))
