"""Pins the FP64 coefficient builder (efs_amd/operator.py) and the FP32
oracle stepping it with the physics of a PEC cavity: the reference's cavity
fixture (tests/fixtures/cavity, a native openEMS run; geometry and source
from tests/generate_fixtures.py:63-130) gives the analytic TE/TM resonances,
their tolerances, and openEMS's own probe trace.

The same 50 x 20 x 60 mm cavity is built here (same mesh lines, PEC walls,
soft source on the edges at node (17, 7, 21), probe ut1z = V_z at node
(6, 5, 6)), run for the fixture's duration, and the probe spectrum must
peak at the analytic TM modes within the fixture's tolerances and at the
same frequencies as openEMS's trace.  A wrong capacitance or inductance
moves every peak; a missing field update leaves no peaks at all.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def cavity():
    with open(os.path.join(HERE, "golden", "cavity_openems.json")) as f:
        return json.load(f)


def peak_near(sig, dt, f_guess, rel_window=0.03):
    """Frequency of the spectral peak within +-rel_window of f_guess
    (Hann window, 16x zero padding, parabolic interpolation in log |X|)."""
    x = np.asarray(sig, np.float64)
    x = (x - x.mean()) * np.hanning(len(x))
    nfft = 16 * int(2 ** np.ceil(np.log2(len(x))))
    X = np.abs(np.fft.rfft(x, nfft))
    f = np.fft.rfftfreq(nfft, dt)
    sel = np.nonzero((f > f_guess * (1 - rel_window)) & (f < f_guess * (1 + rel_window)))[0]
    k = sel[np.argmax(X[sel])]
    a, b, c = np.log(X[k - 1:k + 2])
    d = 0.5 * (a - c) / (a - 2 * b + c)
    return f[k] + d * (f[1] - f[0]), X[k] / X.max()


@pytest.fixture(scope="module")
def simulated(cavity, oracle_mod):
    from efs_amd import operator as op
    dims = cavity["dimensions_m"]
    npts = cavity["mesh_points"]
    mesh = op.Mesh(np.linspace(0, dims["a"], npts["x"]), np.linspace(0, dims["b"], npts["y"]),
                   np.linspace(0, dims["d"], npts["z"]))
    dT = mesh.timestep()
    mat = op.Materials(mesh)
    op.pec_boundary(mat)
    co = op.build_coefficients(mesh, mat, dT)
    nx, ny, nz = mesh.shape
    N = nx * ny * nz
    ex = (nx * 2 // 3, ny * 2 // 3, nz * 2 // 3)  # generator: len(mesh)*2//3 per axis
    lin = ex[0] * ny * nz + ex[1] * nz + ex[2]
    sig = op.gaussian_pulse(5.5e9, 4.5e9, dT)
    o = oracle_mod.Oracle(mesh.shape, co)
    o.configureExcitation({"signal": sig, "delay": np.zeros(3, np.uint32), "amp": np.ones(3, np.float32),
                           "dir": np.array([0, 1, 2], np.uint32), "pos": np.full(3, lin, np.uint32), "period": 0})
    pz = cavity["probe"]["from_node"]
    probe = 2 * N + pz[0] * ny * nz + pz[1] * nz + pz[2]
    duration = cavity["probe"]["sample_dt"] * (len(cavity["probe"]["voltage"]) - 1)
    steps = int(duration / dT)
    o.configureProbeGatherBuffered(np.array([probe], np.uint32), steps, 1)
    o.iterate(steps)
    trace = o.readProbeGatherBuffered()[:, 0]
    return trace, dT


@pytest.mark.parametrize("mode", ["TM110", "TM111"])
def test_cavity_tm_resonances(cavity, simulated, mode):
    """The probe is V_z, so the TM modes (E_z != 0) show, as in the
    reference's own check of this fixture (tests/test_wasm.mjs:140-240):
    peak within [f (1 - tm_freq_lower_rel), f (1 + tm_freq_upper_rel)] of the
    analytic frequency with normalised amplitude >= tm_min_amplitude."""
    trace, dT = simulated
    assert np.all(np.isfinite(trace)) and np.max(np.abs(trace)) > 0
    tol = cavity["tolerances"]
    f_exact = cavity["tm_modes"][mode]
    f_sim, amp = peak_near(trace, dT, f_exact)
    assert f_exact * (1 - tol["tm_freq_lower_rel"]) <= f_sim <= f_exact * (1 + tol["tm_freq_upper_rel"]), \
        (mode, f_sim, f_exact)
    assert amp >= tol["tm_min_amplitude"], (mode, amp)
    # the same numerical dispersion as openEMS (same timestep, same vacuum
    # coefficients): the peak sits where openEMS's trace peaks
    f_native, _ = peak_near(cavity["probe"]["voltage"], cavity["probe"]["sample_dt"], f_exact)
    assert abs(f_sim - f_native) / f_native < 2e-4, (mode, f_sim, f_native)


def test_cavity_strong_peaks_are_modes(cavity, simulated):
    """Every spectral peak above 20 % of the largest lies within 1 % of an
    analytic cavity mode (tests/test_wasm.mjs:205-240)."""
    trace, dT = simulated
    a, b, d = (cavity["dimensions_m"][k] for k in ("a", "b", "d"))
    c0 = 299792458.0
    modes = [c0 / 2 * np.sqrt((m / a) ** 2 + (n / b) ** 2 + (q / d) ** 2)
             for m in range(5) for n in range(5) for q in range(5) if m or n]
    x = np.asarray(trace, np.float64)
    x = (x - x.mean()) * np.hanning(len(x))
    nfft = 16 * int(2 ** np.ceil(np.log2(len(x))))
    X = np.abs(np.fft.rfft(x, nfft))
    f = np.fft.rfftfreq(nfft, dT)
    sel = np.nonzero((f > 1e9) & (f < 10e9))[0]
    Xs = X[sel] / X[sel].max()
    strong = [sel[i] for i in range(1, len(sel) - 1) if Xs[i] > 0.2 and Xs[i] >= Xs[i - 1] and Xs[i] >= Xs[i + 1]]
    assert strong
    for k in strong:
        assert min(abs(f[k] - fm) / fm for fm in modes) < 0.01, f[k]
