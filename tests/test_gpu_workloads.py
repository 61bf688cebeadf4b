"""The workloads at their full lengths and sizes on the HIP path.

- The reference's patch antenna (86 x 87 x 71 graded mesh, Mur on six
  faces, 50-ohm lumped port, NF2FF at 2.4 GHz) over all 30,000 steps
  against the f32 oracle (the reference CPU engine's arithmetic): port
  traces within 1e-5 relative L2 (the FP32 engine tolerance), S11 within
  0.05 dB over 1-3 GHz and Dmax within 0.1 dBi.
- The large mixed-precision workload's HIP path (FP16 storage, 16-bit
  coefficient indices, 4 x 4 lumped-port array, 32-frequency NF2FF with
  K = 64) on a 96 x 96 x 72 sample of the same builder against the FP64
  oracle: active S11 of all 16 ports within 0.1 dB, Dmax within 0.1 dBi.
- The same workload at its full size (800 x 800 x 600, 384 M cells),
  1,500 steps in FP16 storage and in FP32: finite fields and NF2FF
  accumulators, and FP16 active S11 of all 16 ports within 0.1 dB of the
  FP32 run (the FP64 oracle cannot run 384 M cells in a test's time).
"""
import os

import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else n


def traces(eng, w, slices):
    from efs_amd import ports
    r = eng.readProbeGatherBuffered()
    if isinstance(r, dict):
        data, ns, stride = r["data"], r["numSamples"], r["stride"]
    else:
        data, ns, stride = r.ravel(), r.shape[0], r.shape[1]
    return ports.probe_traces(data, ns, stride, w["probes"], slices, 1, w["dT"])


def s11_db(tr, nports, freqs):
    from efs_amd import ports
    return np.stack([20 * np.log10(np.abs(ports.port_s11(tr[2 * p], tr[2 * p + 1], freqs))) for p in range(nports)])


def dmax(accE, accH, w, ff, theta, phi):
    from efs_amd import farfield
    p = farfield.rad_power(accE, accH, w["pointMeta"], w["nf2ff"]["numPoints"])
    return float(farfield.directivity(ff, len(theta), len(phi), p)["DmaxdBi"])


@pytest.mark.timeout(900)
def test_patch_full_run_matches_oracle(oracle_mod):
    from efs_amd import engine, farfield, workloads
    engine.load_library()
    w = workloads.patch_antenna()
    steps = w["nTS"]
    say("patch full run vs f32 oracle", w["gridSize"], steps)
    g = engine.HIPEngine()
    assert g.initGPU()
    sg = workloads.configure_patch(g, w, nf2ff=True, max_samples=steps)
    g.iterate(steps)
    tg = traces(g, w, sg)
    ga = g.readNF2FFAccumulation()
    theta, phi = farfield.sphere_grid()
    cfg = {"pointMeta": w["pointMeta"], "theta": theta, "phi": phi, "center": [0, 0, 0],
           "frequency": float(w["freqs"][0]), "radius": 1.0}
    ffg = g.computeNF2FFfarField(cfg)
    g.destroy()
    oracle_mod.Oracle.set_threads(host_threads())
    o = oracle_mod.Oracle(w["gridSize"], w["coefficients"])
    so = workloads.configure_patch(o, w, nf2ff=True, max_samples=steps)
    o.iterate(steps)
    say("oracle done")
    to = traces(o, w, so)
    oe, oh = o.readNF2FF()
    ffo = oracle_mod.farfield(oe[0], oh[0], w["pointMeta"], theta, phi, np.zeros(3), float(w["freqs"][0]), 1.0)
    for k in range(2):
        assert len(tg[k][1]) == steps
        np.testing.assert_array_equal(tg[k][0], to[k][0])
        err = rel_l2(tg[k][1], to[k][1])
        say("port trace", k, err)
        assert err <= 1e-5
    for got, ref, nm in ((ga["accumE"][0], oe[0], "E"), (ga["accumH"][0], oh[0], "H")):
        err = rel_l2(np.asarray(got, np.float64), ref)
        say("nf2ff", nm, err)
        assert err <= 2e-5
    f = np.linspace(1e9, 3e9, 201)
    ds = np.max(np.abs(s11_db(tg, 1, f) - s11_db(to, 1, f)))
    dg, do = dmax(ga["accumE"][0], ga["accumH"][0], w, ffg, theta, phi), dmax(oe[0], oh[0], w, ffo, theta, phi)
    say("S11 max |err| dB", ds, "Dmax", dg, do)
    assert ds <= 0.05
    assert abs(dg - do) <= 0.1


@pytest.fixture(scope="module")
def large_sample():
    from efs_amd import workloads
    return workloads.large_array(dims=(96, 96, 72), port_pitch=8, n_boxes=20, n_ts=1500, nfreq=32, k_block=64,
                                 threads=host_threads())


@pytest.mark.timeout(600)
def test_large_sample_fp16_matches_fp64_oracle(oracle_mod, large_sample, monkeypatch):
    from efs_amd import engine, farfield, workloads
    engine.load_library()
    w = large_sample
    steps = w["nTS"]
    nports = len(w["probes"]) // 2
    assert nports == 16 and len(w["freqs"]) == 32 and w["nf2ff"]["kBlock"] == 64
    fi = int(np.argmin(np.abs(w["freqs"] - w["f0"])))
    theta, phi = farfield.sphere_grid(5, 5)
    cfg = {"pointMeta": w["pointMeta"], "theta": theta, "phi": phi, "center": [0.024, 0.024, 0.018],
           "frequency": float(w["freqs"][fi]), "radius": 1.0, "freqIndex": fi}
    # the 16-bit coefficient indices of the full-size grid, here too
    monkeypatch.setenv("EFS_COMPRESS", "16")
    say("large sample fp16", w["gridSize"], steps)
    g = engine.HIPEngine()
    assert g.initGPU()
    sg = workloads.configure_large(g, w, storage="fp16", max_samples=steps)
    flags = g.pathFlags()
    assert {"f16", "coef_idx16", "vec4"} <= flags and "coef_idx" not in flags
    g.iterate(steps)
    tg = traces(g, w, sg)
    ga = g.readNF2FFAccumulation()
    ffg = g.computeNF2FFfarField(cfg)
    g.destroy()
    oracle_mod.Oracle.set_threads(host_threads())
    o = oracle_mod.Oracle(w["gridSize"], w["coefficients"], fp64=True)
    so = workloads.configure_large(o, w, max_samples=steps)
    o.iterate(steps)
    to = traces(o, w, so)
    oe, oh = o.readNF2FF()
    ffo = oracle_mod.farfield(oe[fi], oh[fi], w["pointMeta"], theta, phi, np.array(cfg["center"]),
                              cfg["frequency"], 1.0)
    f = np.linspace(2e9, 18e9, 81)
    err = np.abs(s11_db(tg, nports, f) - s11_db(to, nports, f))
    say("active S11 max |err| dB per port", np.round(err.max(axis=1), 4))
    assert np.all(np.isfinite(err)) and err.max() <= 0.1
    dg = dmax(ga["accumE"][fi], ga["accumH"][fi], w, ffg, theta, phi)
    do = dmax(oe[fi], oh[fi], w, ffo, theta, phi)
    say("Dmax fp16", dg, "fp64", do)
    assert abs(dg - do) <= 0.1


@pytest.mark.timeout(900)
def test_large_full_size_fp16_against_fp32():
    from efs_amd import engine, workloads
    engine.load_library()
    say("building the 800 x 800 x 600 workload")
    w = workloads.large_array(threads=host_threads(), n_ts=1500)
    steps = w["nTS"]
    nports = len(w["probes"]) // 2
    f = np.linspace(2e9, 18e9, 81)
    out = {}
    for storage in ("fp16", "fp32"):
        say("large full size", storage, steps)
        g = engine.HIPEngine()
        assert g.initGPU()
        sl = workloads.configure_large(g, w, storage=storage, max_samples=steps)
        flags = g.pathFlags()
        assert "coef_idx16" in flags and ("f16" in flags) == (storage == "fp16")
        done = 0
        while done < steps:
            k = min(500, steps - done)
            g.iterate(k)
            g.sync()
            done += k
            say(storage, "steps", done)
        tr = traces(g, w, sl)
        acc = g.readNF2FFAccumulation()
        finite = bool(np.all(np.isfinite(acc["accumE"])) and np.all(np.isfinite(acc["accumH"])))
        e = g.energy()
        g.destroy()
        assert finite and np.isfinite(e) and e > 0
        out[storage] = s11_db(tr, nports, f)
    err = np.abs(out["fp16"] - out["fp32"])
    say("FP16 vs FP32 active S11 max |err| dB", float(err.max()))
    assert np.all(np.isfinite(err)) and err.max() <= 0.1


@pytest.mark.timeout(300)
def test_nf2ff_compensated_accumulation(oracle_mod):
    """Static fields (coefficients 1 and 0 keep every value), 30,000 steps:
    each accumulator is the sample times a geometric series, which the FP64
    oracle sums in double.  Near DC every fold adds a block of the same sign
    to a growing total, the case where plain FP32 += drifts with the fold
    count; the compensated form (TwoSum per fold, errors kept) stays at the
    rounding of one block.  At 0.37 and 2.4 GHz the rounding inside each
    block (FP32 twiddles and matrix-core sums) dominates, and the two forms
    are close; the compensated one must not be worse."""
    from efs_amd import engine
    engine.load_library()
    n = 16
    N3 = 3 * n ** 3
    co = {"vv": np.ones(N3, np.float32), "vi": np.zeros(N3, np.float32), "ii": np.ones(N3, np.float32),
          "iv": np.zeros(N3, np.float32)}
    # small integer fields and power-of-two edge lengths: the surface
    # samples (means and interpolations of neighbours over edge lengths) are
    # exact in FP32, so only the accumulation differs from the oracle's
    rng = np.random.default_rng(5)
    v0 = rng.integers(-8, 9, N3).astype(np.float32)
    c0 = rng.integers(-8, 9, N3).astype(np.float32)
    pts = np.array([(x, y, z) for x in range(3, 13) for y in range(3, 13) for z in (3, 12)], np.uint32).ravel()
    el = [np.full(n, 2.0 ** -10, np.float32)] * 3
    dT, steps = 1.9e-12, 30000
    omegas = 2 * np.pi * np.array([1e6, 0.37e9, 2.4e9])
    o = oracle_mod.Oracle([n] * 3, co, fp64=True)
    o.uploadFields(v0, c0)
    o.configureNF2FF(pts, omegas, dT, el, el)
    o.iterate(steps)
    oe, oh = o.readNF2FF()
    errs = {}
    for comp in (True, False):
        g = engine.HIPEngine()
        assert g.initGPU()
        g.init([n] * 3, co)
        g.uploadFields(v0, c0)
        g.configureNF2FFAccumulation({"surfaceIndices": pts, "omega": list(omegas), "dT": dT, "primalEdgeLens": el,
                                      "dualEdgeLens": el, "compensated": comp})
        g.iterate(steps)
        a = g.readNF2FFAccumulation()
        g.destroy()
        errs[comp] = [rel_l2(np.concatenate([np.asarray(a["accumE"][k], np.float64),
                                             np.asarray(a["accumH"][k], np.float64)]),
                             np.concatenate([oe[k], oh[k]])) for k in range(len(omegas))]
    say("relative L2 vs FP64, compensated", errs[True], "plain", errs[False])
    assert errs[True][0] <= 0.25 * errs[False][0] and errs[True][0] <= 1e-7
    for k in (1, 2):
        assert errs[True][k] <= 1.2 * errs[False][k] + 1e-9 and errs[True][k] <= 1e-5
