"""Point work between the sweeps in one launch per phase against the
separate kernels (EFS_FUSE_HOOKS=0):
  k_post_e  Mur post + apply and the excitation,
  k_post_h  NF2FF sampling, the probe gather and the next step's Mur pre.
The bar is the same bits (fields, buffered gather, NF2FF accumulators):
the fused launches keep the reference's phase order for every cell that
two operations share.  The Mur box case puts excitation entries on Mur
normal cells and on a cell that is both normal and shifted cell, which
exercises the deferred lists (phase B, or the first workgroups of k_post_e
in the group form, EFS_POST_GROUP) and, on the default path for grids this
small, the end of the E-sweep blocks (EFS_SWEEP_EPILOGUE); iterate calls of 1, 7, 64 and 128
steps check that the Mur pre pass runs exactly once per step across call
boundaries.  The fused path is also held to the FP64 oracle (1e-5)."""
import numpy as np
import pytest

from cases import rel_l2
from efs_amd import operator as op

pytestmark = pytest.mark.gpu

CHUNKS = (1, 7, 64, 128)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def HIPEngine():
    from efs_amd import engine
    engine.load_library()
    if engine.device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    return engine.HIPEngine


def mur_box_case(dims=(28, 24, 20)):
    """Vacuum box, Mur on 6 faces (corner entries alias), a dielectric
    block, excitation on two Mur normal cells, on a corner cell and inside,
    buffered probe gather, NF2FF box."""
    mesh = op.uniform_mesh(dims, 1e-3)
    mat = op.Materials(mesh)
    mat.box((8, 6, 5), (14, 12, 9), eps=3.0)
    dT = mesh.timestep(0.95)
    co = op.to_f32(op.build_coefficients(mesh, mat, dT))
    nx, ny, nz = dims
    N = nx * ny * nz

    def lin(x, y, z):
        return x * ny * nz + y * nz + z

    # (component, cell): Ey on the x = 0 face (a normal cell), Ez on the
    # y = Ny-1 face, Ez at (1, 0, 5) (normal cell of the y = 0 face and
    # shifted cell of the x = 0 face's entry at (0, 0, 5)), two inside
    cells = [(1, lin(0, 7, 9)), (2, lin(11, ny - 1, 4)), (2, lin(1, 0, 5)), (2, lin(14, 12, 10)),
             (0, lin(5, 5, 5))]
    sig = op.gaussian_pulse(20e9, 15e9, dT)
    exc = {"signal": sig, "delay": np.array([0, 3, 1, 0, 2], np.uint32),
           "amp": np.array([1.0, -0.5, 0.8, 1.2, 0.3], np.float32),
           "dir": np.array([c for c, _ in cells], np.uint32), "pos": np.array([p for _, p in cells], np.uint32),
           "period": 0}
    mur = op.mur_config(mesh, range(6), dT)
    probe = np.array([c * N + p for c, p in cells] + [lin(20, 12, 10), N + lin(3, 3, 3)], np.uint32)
    pts, _, primal, dual = op.nf2ff_box(mesh, [3] * 3, [nx - 4, ny - 4, nz - 4])
    nf2ff = {"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": [2 * np.pi * 20e9, 2 * np.pi * 30e9],
             "dT": dT, "primalEdgeLens": primal, "dualEdgeLens": dual, "kBlock": 16}
    return {"dims": list(dims), "coeffs": co, "excitation": exc, "mur": mur, "probe": probe, "nf2ff": nf2ff}


def run_case(eng, case, storage="fp32", chunks=CHUNKS):
    eng.init(case["dims"], case["coeffs"], storage=storage)
    eng.configureExcitation(case["excitation"])
    eng.configureMur(case["mur"])
    eng.configureProbeGatherBuffered(case["probe"], 512, 1)
    eng.configureNF2FFAccumulation(case["nf2ff"])
    snaps = []
    for n in chunks:
        eng.iterate(n)
        f = eng.getFields()
        snaps.append((f["volt"].copy(), f["curr"].copy()))
    g = eng.readProbeGatherBuffered()
    acc = eng.readNF2FFAccumulation()
    return snaps, g["data"].copy(), np.concatenate([acc["accumE"].ravel(), acc["accumH"].ravel()])


@pytest.mark.parametrize("storage,group,epi", [("fp32", "1", "1"), ("fp16", "1", "1"), ("fp64", "1", "1"),
                                               ("fp32", "1", "0"), ("fp16", "1", "0"), ("fp32", "0", "0"),
                                               ("fp16", "0", "0")])
def test_fused_hooks_match_separate_kernels(HIPEngine, monkeypatch, storage, group, epi):
    case = mur_box_case()
    say("case mur box", storage, group, epi)
    g = HIPEngine()
    assert g.initGPU()
    monkeypatch.setenv("EFS_FUSE_HOOKS", "0")
    ref = run_case(g, case, storage)
    monkeypatch.setenv("EFS_FUSE_HOOKS", "1")
    monkeypatch.setenv("EFS_POST_GROUP", group)
    monkeypatch.setenv("EFS_SWEEP_EPILOGUE", epi)
    out = run_case(g, case, storage)
    # the Mur pairs and excitation cells of the box lie inside single sweep
    # blocks, so the E-sweep epilogue takes them (FP32 / FP16 storage)
    assert ("sweep_epilogue" in g.pathFlags()) == (epi == "1" and storage != "fp64")
    for k, ((v0, c0), (v1, c1)) in enumerate(zip(ref[0], out[0])):
        assert np.array_equal(v0, v1) and np.array_equal(c0, c1), f"fields differ after chunk {k}"
    assert np.array_equal(ref[1], out[1]), "buffered gather differs"
    assert np.array_equal(ref[2], out[2]), "NF2FF accumulators differ"
    assert np.abs(ref[0][-1][0]).max() > 0


def test_fused_hooks_against_fp64_oracle(HIPEngine, oracle_mod):
    """The fused FP32 path on the Mur box case against the FP64 oracle:
    fields and the probe trace within 1e-5 (relative L2) after 100 steps.
    Later the excitation leaves a static E field while H decays to 1e-3 of
    its peak, and H's relative error is then FP32 rounding over a vanishing
    norm (the FP32 oracle itself is 1.6e-4 from FP64 at step 200)."""
    case = mur_box_case()
    o = oracle_mod.Oracle(case["dims"], case["coeffs"], fp64=True)
    o.configureExcitation(case["excitation"])
    o.configureMur(case["mur"])
    g = HIPEngine()
    assert g.initGPU()
    g.init(case["dims"], case["coeffs"])
    g.configureExcitation(case["excitation"])
    g.configureMur(case["mur"])
    g.configureProbeGatherBuffered(case["probe"], 512, 1)
    g.configureNF2FFAccumulation(case["nf2ff"])
    trace = []
    for _ in range(100):
        o.iterate(1)
        r = o.getFields()
        trace.append(np.stack([r["volt"][case["probe"]], r["curr"][case["probe"]]], 1).ravel())
    g.iterate(100)
    f = g.getFields()
    ev, ec = rel_l2(f["volt"], r["volt"]), rel_l2(f["curr"], r["curr"])
    gt = g.readProbeGatherBuffered()["data"]
    et = rel_l2(gt, np.concatenate(trace))
    say(f"fused vs fp64 oracle: volt {ev:.2e} curr {ec:.2e} probes {et:.2e}")
    assert ev < 1e-5 and ec < 1e-5 and et < 1e-5


def test_patch_fused_hooks_match_separate_kernels(HIPEngine, monkeypatch):
    """The reference's patch example (Mur on 6 faces, lumped port, buffered
    port gather, NF2FF): 400 steps in calls of 1, 7, 64, 128 and 200, the
    same bits with and without the fused launches."""
    from efs_amd import workloads
    w = workloads.patch_antenna()
    say("case patch antenna", w["gridSize"])
    res = {}
    for fuse in ("0", "1"):
        monkeypatch.setenv("EFS_FUSE_HOOKS", fuse)
        g = HIPEngine()
        assert g.initGPU()
        workloads.configure_patch(g, w, nf2ff=True, max_samples=400)
        for n in (1, 7, 64, 128, 200):
            g.iterate(n)
        if fuse == "1":  # Mur and the port's excitation at the end of the E-sweep blocks
            assert {"side_work", "sweep_epilogue"} <= g.pathFlags()
        f = g.getFields()
        acc = g.readNF2FFAccumulation()
        res[fuse] = (f["volt"], f["curr"], g.readProbeGatherBuffered()["data"],
                     np.concatenate([acc["accumE"].ravel(), acc["accumH"].ravel()]))
        g.destroy()
    for a, b, what in zip(res["0"], res["1"], ("volt", "curr", "gather", "nf2ff")):
        assert np.array_equal(a, b), what
