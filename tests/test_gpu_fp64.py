"""The FP64 storage form (EFS_STORAGE_F64): the precision reference run of
the same engine -- fields, UPML flux, Mur state and NF2FF buffer, fold and
accumulators in double on the 1-wide sweeps -- against the FP64 oracle (the
C restatement in double).  Both evaluate the same formulas in double from
the same FP32 coefficients; they differ by operation order and fused
multiply-adds only, so the tolerance is 1e-12 relative L2 (measured values
are printed), six orders below the FP32 engine's 1e-5 bar.  ADE and TFSF
run in this form too; the engine refuses the extensions it does not run
(RLC, periodic boundaries, steady-state detection)."""
import numpy as np
import pytest

from cases import configure, random_case, rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def say(*a):
    print(*a, flush=True)


def nf2ff_cfg(dims, inset=2):
    pts = []
    lo, hi = [inset] * 3, [d - inset - 1 for d in dims]
    for x in range(lo[0], hi[0] + 1):
        for y in range(lo[1], hi[1] + 1):
            for z in range(lo[2], hi[2] + 1):
                if x in (lo[0], hi[0]) or y in (lo[1], hi[1]) or z in (lo[2], hi[2]):
                    pts.append((x, y, z))
    rng = np.random.default_rng(7)
    primal = [rng.uniform(0.8e-3, 1.2e-3, d).astype(np.float32) for d in dims]
    dual = [rng.uniform(0.8e-3, 1.2e-3, d).astype(np.float32) for d in dims]
    return np.array(pts, np.uint32).ravel(), primal, dual


CASES = [
    ("upml_exc", 61, (20, 18, 22), dict(pml_width=3)),
    ("mur_exc_odd", 62, (17, 19, 13), dict(mur=True)),
    ("upml_mur", 63, (24, 20, 16), dict(pml_width=4, mur=True)),
]


@pytest.mark.parametrize("name,seed,dims,kw", CASES, ids=[c[0] for c in CASES])
def test_fp64_engine_matches_fp64_oracle(oracle_mod, name, seed, dims, kw):
    from efs_amd import engine
    engine.load_library()
    case = random_case(seed, dims, **kw)
    pts, primal, dual = nf2ff_cfg(dims)
    dT, steps = 1.9e-12, 90
    omegas = 2 * np.pi * np.array([1.5e9, 6e9, 21e9])
    say("case fp64", name, dims)
    g = engine.HIPEngine()
    assert g.initGPU()
    g.init(case["dims"], case["coeffs"], storage="fp64")
    configure(g, case)
    assert "f64" in g.pathFlags() and "vec4" not in g.pathFlags()
    g.configureNF2FFAccumulation({"surfaceIndices": pts, "omega": list(omegas), "dT": dT, "primalEdgeLens": primal,
                                  "dualEdgeLens": dual, "kBlock": 16})
    idx = np.array([5, 77, 3 * int(np.prod(dims)) - 9], np.uint32)
    g.configureProbeGatherBuffered(idx, steps, 1)
    o = oracle_mod.Oracle(case["dims"], case["coeffs"], fp64=True)
    configure(o, case)
    o.configureNF2FF(pts, omegas, dT, primal, dual)
    o.configureProbeGatherBuffered(idx, steps, 1)
    g.iterate(steps)
    o.iterate(steps)
    gf, of = g.getFields(fp64=True), o.getFields()
    for k in ("volt", "curr"):
        err = rel_l2(gf[k], of[k])
        say(k, err)
        assert np.all(np.isfinite(gf[k])) and err <= TOL64
    ge, oe = g.energy(), o.energy()
    say("energy", ge, oe)
    assert abs(ge - oe) <= TOL64 * abs(oe)
    acc = g.readNF2FFAccumulation(fp64=True)
    re, rh = o.readNF2FF()
    for f in range(len(omegas)):
        err = rel_l2(np.concatenate([acc["accumE"][f], acc["accumH"][f]]), np.concatenate([re[f], rh[f]]))
        say("nf2ff", f, err)
        assert err <= TOL64
    rg = g.readProbeGatherBuffered()
    ro = o.readProbeGatherBuffered()
    assert rg["numSamples"] == ro.shape[0]
    # the ring holds FP32 copies of the FP64 values
    np.testing.assert_allclose(np.asarray(rg["data"], np.float64).reshape(ro.shape), ro, rtol=1e-7, atol=1e-30)
    # the far field from the FP64 accumulators (FP32 per-point terms, FP64 sums)
    theta, phi = np.linspace(0.1, 3.0, 7), np.linspace(0, 6.0, 9)
    P = len(pts) // 3
    meta = np.zeros((P, 8), np.float32)
    rng = np.random.default_rng(3)
    meta[:, :3] = rng.uniform(-0.01, 0.01, (P, 3))
    meta[:, 3] = rng.integers(0, 3, P)
    meta[:, 4] = rng.choice([-1.0, 1.0], P)
    meta[:, 5] = 1e-6
    ff = g.computeNF2FFfarField({"pointMeta": meta.ravel(), "theta": theta, "phi": phi, "center": [0, 0, 0],
                                 "frequency": omegas[1] / (2 * np.pi), "radius": 1.0, "freqIndex": 1})
    ref = oracle_mod.farfield(re[1], rh[1], meta.ravel(), theta, phi, np.zeros(3), omegas[1] / (2 * np.pi), 1.0)
    assert rel_l2(ff, ref) <= 1e-5
    g.destroy()


@pytest.mark.parametrize("apply", [False, True], ids=["ade", "ade_write_back"])
def test_fp64_ade_tfsf_match_fp64_oracle(oracle_mod, apply):
    """ADE (Drude and Lorentz orders; with and without the openEMS
    write-back) and TFSF on UPML + Mur in FP64 storage, 30 steps: at this
    length the write-back changes the fields by ~3 %, so the 1e-12 bar
    checks the ADE arithmetic (later the random coefficients grow the
    fields past 1e18)."""
    from efs_amd import engine
    engine.load_library()
    case = random_case(65, (18, 16, 20), pml_width=3, ade=True, tfsf=True, mur=True)
    say("case fp64 ade/tfsf, write-back", apply)
    g = engine.HIPEngine()
    assert g.initGPU()
    g.init(case["dims"], case["coeffs"], storage="fp64")
    configure(g, case, ade_apply=apply)
    o = oracle_mod.Oracle(case["dims"], case["coeffs"], fp64=True)
    configure(o, case, ade_apply=apply)
    g.iterate(30)
    o.iterate(30)
    gf, of = g.getFields(fp64=True), o.getFields()
    for k in ("volt", "curr"):
        err = rel_l2(gf[k], of[k])
        say(k, err)
        assert np.all(np.isfinite(gf[k])) and err <= TOL64
    g.destroy()


def test_fp64_refuses_unsupported_extensions():
    from efs_amd import engine
    engine.load_library()
    case = random_case(64, (12, 12, 12), rlc=True, pbc=[(0, 0.0)])
    g = engine.HIPEngine()
    assert g.initGPU()
    g.init(case["dims"], case["coeffs"], storage="fp64")
    for fn, arg in ((g.configureRLC, case["rlc"]), (g.configurePBC, case["pbc"])):
        with pytest.raises(Exception, match="FP64"):
            fn(arg)
    g.destroy()
