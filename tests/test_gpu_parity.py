"""GPU parity of the HIP engine (through the C ABI) against the CPU oracle
and the reference's own step vectors.

Tolerance (stated by the task for FP32 engines): relative L2 of every field
array <= 1e-5 after all steps; per-element |gpu - ref| <= 1e-5 * max|ref|.
The GPU computes in FP32 with fused multiply-adds while the oracle follows
the reference's JavaScript arithmetic (double evaluation, float storage), so
bit equality is not expected; integer work (sample counts, step counters)
is compared exactly.
"""
import numpy as np
import pytest

from cases import configure, golden_cases, random_case, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def HIPEngine():
    from efs_amd import engine
    engine.load_library()
    if engine.device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    return engine.HIPEngine


def make_gpu(HIPEngine, case, ade_apply=False):
    g = HIPEngine()
    assert g.initGPU()
    g.init(case["dims"], case["coeffs"])
    configure(g, case, ade_apply)
    return g


def make_oracle(orc, case, ade_apply=False, fp64=False):
    o = orc.Oracle(case["dims"], case["coeffs"], fp64=fp64)
    configure(o, case, ade_apply)
    return o


def assert_fields_close(gf, of, tol=TOL, what=""):
    for key in ("volt", "curr"):
        g = np.asarray(gf[key], np.float64)
        o = np.asarray(of[key], np.float64)
        assert np.all(np.isfinite(g)), f"{what} {key}: non-finite values"
        err = rel_l2(g, o)
        scale = max(np.max(np.abs(o)), 1e-30)
        mx = np.max(np.abs(g - o)) / scale
        assert err <= tol and mx <= tol, f"{what} {key}: relL2={err:.3e} maxrel={mx:.3e}"


@pytest.mark.parametrize("name", ["core", "all", "all_periodic"])
def test_reference_vectors(HIPEngine, name):
    """GPU vs outputs written by the reference's CPUFDTDEngine itself."""
    case = golden_cases()[name]
    say("case", name, case["dims"], case["steps"])
    g = make_gpu(HIPEngine, case)
    g.iterate(case["steps"])
    assert g.numTS == case["out"]["numTS"]
    assert_fields_close(g.getFields(), case["out"], what=name)
    g.destroy()


CASES = [
    ("core_odd", 1, (37, 29, 41), dict(pml_width=0)),
    ("core_wide_cols", 2, (13, 3, 300), dict(pml_width=0)),
    ("pml_shell", 3, (20, 18, 22), dict(pml_width=3)),
    ("pml_exc_mur_alias", 4, (16, 12, 14), dict(pml_width=2, mur=True)),
    ("everything", 5, (18, 16, 20), dict(pml_width=3, mur=True, ade=True, tfsf=True, rlc=True)),
    ("thin_x", 6, (2, 40, 33), dict(pml_width=1)),
    ("line_8x1x1", 7, (8, 1, 1), dict(pml_width=0)),
    ("single_cell", 8, (1, 1, 1), dict(pml_width=0)),
    ("flat_y1", 9, (11, 1, 9), dict(pml_width=0)),
    ("v4_pml_aligned", 31, (16, 12, 16), dict(pml_width=4, mur=True)),
    ("v4_pml_split", 32, (20, 18, 24), dict(pml_width=3, ade=True, tfsf=True)),
    ("v4_ny1", 34, (9, 1, 12), dict(pml_width=0)),
    ("v4_nz4", 35, (6, 5, 4), dict(pml_width=1, mur=True)),
    ("v4_rlc_tfsf", 36, (12, 10, 8), dict(pml_width=2, tfsf=True, rlc=True, ade=True)),
    ("split_pml8_all", 37, (24, 20, 32), dict(pml_width=8, mur=True, ade=True, tfsf=True, rlc=True)),
    ("split_pml4_tall", 38, (40, 9, 12), dict(pml_width=4, mur=True)),
]


@pytest.mark.parametrize("name,seed,dims,kw", CASES, ids=[c[0] for c in CASES])
def test_random_cases(HIPEngine, oracle_mod, name, seed, dims, kw):
    case = random_case(seed, dims, **kw)
    steps = 40
    say("case", name, seed, dims, steps)
    o = make_oracle(oracle_mod, case)
    g = make_gpu(HIPEngine, case)
    o.iterate(steps)
    g.iterate(steps)
    assert g.numTS == o.numTS == steps
    assert_fields_close(g.getFields(), o.getFields(), what=name)
    g.destroy()


@pytest.mark.parametrize("name,seed,dims,kw", [c for c in CASES if c[2][2] % 4 == 0],
                         ids=[c[0] for c in CASES if c[2][2] % 4 == 0])
def test_scalar_lane_path(HIPEngine, oracle_mod, monkeypatch, name, seed, dims, kw):
    """The same cases with the one-cell-per-lane sweeps forced (EFS_VEC=1)
    and no layout padding (EFS_PAD=0)."""
    monkeypatch.setenv("EFS_VEC", "1")
    monkeypatch.setenv("EFS_PAD", "0")
    monkeypatch.setenv("EFS_STAGGER", "0")
    case = random_case(seed, dims, **kw)
    say("case scalar-lane", name, seed, dims)
    o = make_oracle(oracle_mod, case)
    g = make_gpu(HIPEngine, case)
    o.iterate(30)
    g.iterate(30)
    assert_fields_close(g.getFields(), o.getFields(), what=name + "/vec1")
    g.destroy()


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_chunk_lengths_same_bits(HIPEngine, monkeypatch, storage):
    """The x chunk a workgroup marches (1 plane each, the default on small
    grids, or long chunks from a small block target) changes only which x
    neighbours come from registers: fields stay bit-identical."""
    case = random_case(61, (26, 20, 32), pml_width=4, mur=True, ade=True, tfsf=True)
    out = []
    for blocks, planes in (("2048", "1"), ("16", "1"), ("2048", "4")):
        monkeypatch.setenv("EFS_SWEEP_BLOCKS", blocks)
        monkeypatch.setenv("EFS_SWEEP_MIN_PLANES", planes)
        say("case chunks", storage, blocks, planes)
        g = HIPEngine()
        assert g.initGPU()
        g.init(case["dims"], case["coeffs"], storage=storage)
        configure(g, case)
        g.iterate(40)
        out.append(g.getFields())
        g.destroy()
    for f in out[1:]:
        np.testing.assert_array_equal(f["volt"], out[0]["volt"])
        np.testing.assert_array_equal(f["curr"], out[0]["curr"])


QCASES = [
    ("q_core", 41, (16, 14, 32), dict(quantize=23)),
    ("q_pml_split", 42, (20, 18, 24), dict(pml_width=3, quantize=50, mur=True, ade=True, tfsf=True, rlc=True)),
    ("q_pml_aligned", 43, (16, 16, 16), dict(pml_width=4, quantize=200, mur=True)),
    ("q_split_pml8", 44, (26, 22, 36), dict(pml_width=8, quantize=120, mur=True, ade=True, tfsf=True, rlc=True)),
]


@pytest.mark.parametrize("name,seed,dims,kw", QCASES, ids=[c[0] for c in QCASES])
def test_compressed_coefficients(HIPEngine, oracle_mod, monkeypatch, name, seed, dims, kw):
    """Few distinct coefficients: the table-compressed sweeps must give the
    same bits as the FP32-array sweeps (same values, same arithmetic) and
    match the oracle."""
    case = random_case(seed, dims, **kw)
    say("case compressed", name, seed, dims)
    g = make_gpu(HIPEngine, case)
    monkeypatch.setenv("EFS_COMPRESS", "0")
    u = make_gpu(HIPEngine, case)
    o = make_oracle(oracle_mod, case)
    for e in (g, u, o):
        e.iterate(30)
    fg, fu = g.getFields(), u.getFields()
    np.testing.assert_array_equal(fg["volt"], fu["volt"])
    np.testing.assert_array_equal(fg["curr"], fu["curr"])
    assert_fields_close(fg, o.getFields(), what=name)


Q16CASES = [
    ("q16_core", 51, (20, 18, 32), dict(quantize=300)),
    ("q16_pml_fused", 52, (24, 20, 32), dict(pml_width=4, quantize=400, mur=True, ade=True, tfsf=True)),
    ("q16_odd_nz", 53, (18, 16, 27), dict(quantize=250)),
]


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("name,seed,dims,kw", Q16CASES, ids=[c[0] for c in Q16CASES])
def test_coefficient_idx16(HIPEngine, oracle_mod, monkeypatch, name, seed, dims, kw, storage):
    """More than 256 distinct (a, b) pairs per component but at most 65,536:
    the 16-bit index form, with the same bits as the FP32 arrays (same
    values, same arithmetic), and (FP32) the oracle's fields."""
    case = random_case(seed, dims, **kw)
    say("case idx16", name, seed, dims, storage)
    g = HIPEngine()
    assert g.initGPU()
    g.init(case["dims"], case["coeffs"], storage=storage)
    configure(g, case)
    assert "coef_idx16" in g.pathFlags() and "coef_idx" not in g.pathFlags()
    monkeypatch.setenv("EFS_COMPRESS", "0")
    u = HIPEngine()
    assert u.initGPU()
    u.init(case["dims"], case["coeffs"], storage=storage)
    configure(u, case)
    assert "coef_idx16" not in u.pathFlags()
    for e in (g, u):
        e.iterate(30)
    fg, fu = g.getFields(), u.getFields()
    np.testing.assert_array_equal(fg["volt"], fu["volt"])
    np.testing.assert_array_equal(fg["curr"], fu["curr"])
    if storage == "fp32":
        o = make_oracle(oracle_mod, case)
        o.iterate(30)
        assert_fields_close(fg, o.getFields(), what=name)
    g.destroy()
    u.destroy()


@pytest.mark.parametrize("name,seed,dims,kw", [
    ("odd_nz_extensions", 47, (23, 19, 29), dict(pml_width=0, mur=True, tfsf=True, ade=True, rlc=True)),
    ("odd_nz_upml", 48, (21, 18, 30), dict(pml_width=3, mur=True)),
], ids=["odd_nz_extensions", "odd_nz_upml"])
def test_padded_rows_match_unpadded(HIPEngine, monkeypatch, name, seed, dims, kw):
    """Nz % 4 != 0: device rows padded to a multiple of 4 (4-wide sweeps, or
    1-wide on the padded rows when the UPML is not aligned) give the bits
    of the unpadded 1-wide sweeps (EFS_VEC=1): same per-cell arithmetic."""
    case = random_case(seed, dims, **kw)
    say("case padded", name, dims)
    a = make_gpu(HIPEngine, case)
    monkeypatch.setenv("EFS_VEC", "1")
    b = make_gpu(HIPEngine, case)
    monkeypatch.delenv("EFS_VEC")
    if kw["pml_width"] == 0:
        assert "vec4" in a.pathFlags()
    assert "vec4" not in b.pathFlags()
    a.iterate(30)
    b.iterate(30)
    fa, fb = a.getFields(), b.getFields()
    np.testing.assert_array_equal(fa["volt"], fb["volt"])
    np.testing.assert_array_equal(fa["curr"], fb["curr"])
    assert a.energy() == b.energy()


def test_split_matches_unsplit(HIPEngine, monkeypatch):
    """Interior + UPML launches vs the one-cell-per-lane fused sweeps on the
    same aligned case: same arithmetic per cell, so identical bits."""
    case = random_case(45, (24, 20, 32), pml_width=8, mur=True, tfsf=True)
    say("case split vs vec1")
    a = make_gpu(HIPEngine, case)
    monkeypatch.setenv("EFS_VEC", "1")
    b = make_gpu(HIPEngine, case)
    assert "upml4" in a.pathFlags() and "vec4" not in b.pathFlags()
    a.iterate(25)
    b.iterate(25)
    fa, fb = a.getFields(), b.getFields()
    np.testing.assert_array_equal(fa["volt"], fb["volt"])
    np.testing.assert_array_equal(fa["curr"], fb["curr"])


@pytest.mark.parametrize("name,seed,dims,kw", [
    ("fused_pml8_all", 46, (24, 20, 32), dict(pml_width=8, mur=True, tfsf=True, ade=True, rlc=True)),
    ("fused_pml4_q", 47, (20, 16, 24), dict(pml_width=4, mur=True, quantize=5)),
    ("fused_tall", 48, (40, 9, 12), dict(pml_width=4)),
])
def test_fused_matches_two_launch(HIPEngine, oracle_mod, monkeypatch, name, seed, dims, kw):
    """One grid launch doing core and UPML cells vs the interior launch plus
    the region launch: same arithmetic per cell, so identical bits; both
    against the oracle too."""
    case = random_case(seed, dims, **kw)
    say(f"case {name} fused vs two launches")
    a = make_gpu(HIPEngine, case)
    monkeypatch.setenv("EFS_PML_FUSED", "0")
    b = make_gpu(HIPEngine, case)
    assert "fused" in a.pathFlags() and {"upml4"} <= b.pathFlags() and "fused" not in b.pathFlags()
    o = make_oracle(oracle_mod, case)
    a.iterate(30)
    b.iterate(30)
    o.iterate(30)
    fa, fb = a.getFields(), b.getFields()
    np.testing.assert_array_equal(fa["volt"], fb["volt"])
    np.testing.assert_array_equal(fa["curr"], fb["curr"])
    assert_fields_close(fa, o.getFields(), what=name)


PQCASES = [
    # per region arrays of 3 levels each -> at most 27 triples per region and component
    ("pq_pml8_all", 51, (24, 20, 32), dict(pml_width=8, pml_quantize=3, mur=True, tfsf=True, ade=True, rlc=True)),
    ("pq_both_tables", 52, (20, 16, 24), dict(pml_width=4, pml_quantize=3, quantize=6, mur=True)),
    ("pq_tall", 53, (40, 9, 12), dict(pml_width=4, pml_quantize=2)),
    # 6 regions x 6^3 triples > 256 per component: tables do not fit, FP32 UPML arrays are used
    ("pq_overflow", 54, (20, 16, 24), dict(pml_width=4, pml_quantize=6)),
]


@pytest.mark.parametrize("name,seed,dims,kw", PQCASES, ids=[c[0] for c in PQCASES])
def test_compressed_upml_coefficients(HIPEngine, oracle_mod, monkeypatch, name, seed, dims, kw):
    """UPML coefficients as byte indices into LDS tables vs the FP32 region
    arrays: same values and arithmetic, so identical bits; and the oracle."""
    case = random_case(seed, dims, **kw)
    say(f"case {name} compressed UPML")
    a = make_gpu(HIPEngine, case)
    monkeypatch.setenv("EFS_PML_COMPRESS", "0")
    b = make_gpu(HIPEngine, case)
    assert ("upml_idx" in a.pathFlags()) == (name != "pq_overflow"), a.pathFlags()
    assert "upml_idx" not in b.pathFlags() and "fused" in b.pathFlags()
    o = make_oracle(oracle_mod, case)
    for e in (a, b, o):
        e.iterate(30)
    fa, fb = a.getFields(), b.getFields()
    np.testing.assert_array_equal(fa["volt"], fb["volt"])
    np.testing.assert_array_equal(fa["curr"], fb["curr"])
    assert_fields_close(fa, o.getFields(), what=name)


PBC_CASES = [
    ("pbc_x", 61, (12, 10, 16), dict(pbc=[(0, 0.0)])),
    ("pbc_bloch_yz", 62, (10, 12, 16), dict(pbc=[(1, 0.6), (2, -1.1)], tfsf=True)),
    ("pbc_x_pml_mur", 63, (14, 10, 12), dict(pbc=[(0, 0.3)], pml_width=2, mur=True, ade=True)),
    ("pbc_xyz_vec1", 64, (9, 8, 7), dict(pbc=[(2, 0.0), (0, 2.5), (1, 0.0)], rlc=True)),
]


@pytest.mark.parametrize("name,seed,dims,kw", PBC_CASES, ids=[c[0] for c in PBC_CASES])
def test_periodic_boundaries(HIPEngine, oracle_mod, name, seed, dims, kw):
    """configurePBC (plain and Bloch, several axes) against the oracle."""
    case = random_case(seed, dims, **kw)
    say("case", name)
    o = make_oracle(oracle_mod, case)
    g = make_gpu(HIPEngine, case)
    o.iterate(40)
    g.iterate(40)
    assert_fields_close(g.getFields(), o.getFields(), what=name)


def test_steady_state(HIPEngine, oracle_mod):
    """configureSteadyState: recording window, period split, counter and
    convergence rule against the oracle."""
    case = random_case(65, (12, 11, 16), pml_width=2)
    N = 12 * 11 * 16
    probes = np.array([5, N + 700, 2 * N + 1333, 900], np.uint32)
    say("case steady state")
    o = make_oracle(oracle_mod, case)
    g = make_gpu(HIPEngine, case)
    cfg = {"probe_idx": probes, "periodSamples": 7, "threshold": 0.5}
    o.configureSteadyState(cfg)
    g.configureSteadyState(cfg)
    for e in (o, g):
        e.iterate(5)            # not recording: nothing accumulates
    assert g.readSteadyState()["currentSample"] == 0
    o.ssSet(True, 0)
    g.ssRecording = True
    g.ssCurrentSample = 0
    for e in (o, g):
        e.iterate(10)
    mid_o, mid_g = o.ssRead(), g.readSteadyState()
    assert mid_g["currentSample"] == mid_o["currentSample"] == 10
    assert not g.checkConvergence() and not o.checkConvergence()  # period 2 not complete
    for e in (o, g):
        e.iterate(6)
    so, sg = o.ssRead(), g.readSteadyState()
    assert sg["currentSample"] == so["currentSample"] == 16
    for k in ("energy_period1", "energy_period2"):
        np.testing.assert_allclose(sg[k], so[k], rtol=1e-5, atol=1e-30)
    assert g.checkConvergence() == o.checkConvergence()
    g.ssRecording = False
    g.iterate(3)
    assert g.readSteadyState()["currentSample"] == 16


def test_ade_write_back(HIPEngine, oracle_mod):
    """Optional openEMS ADE write-back, same switch on both sides."""
    case = random_case(11, (14, 12, 10), pml_width=2, ade=True)
    say("case ade_apply")
    o = make_oracle(oracle_mod, case, ade_apply=True)
    g = make_gpu(HIPEngine, case, ade_apply=True)
    o.iterate(30)
    g.iterate(30)
    assert_fields_close(g.getFields(), o.getFields(), what="ade_apply")


def test_iterate_split_equals_single_call(HIPEngine):
    case = random_case(12, (15, 14, 13), pml_width=2, mur=True, tfsf=True)
    say("case split iterate")
    a = make_gpu(HIPEngine, case)
    b = make_gpu(HIPEngine, case)
    a.iterate(25)
    for n in (1, 7, 17):
        b.iterate(n)
    fa, fb = a.getFields(), b.getFields()
    np.testing.assert_array_equal(fa["volt"], fb["volt"])
    np.testing.assert_array_equal(fa["curr"], fb["curr"])


def nf2ff_setup(dims, inset=2, shuffle=False):
    nx, ny, nz = dims
    pts = []
    lo = [inset] * 3
    hi = [nx - 1 - inset, ny - 1 - inset, nz - 1 - inset]
    for ax in range(3):
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        for side in (lo[ax], hi[ax]):
            for i in range(lo[a1], hi[a1] + 1):
                for j in range(lo[a2], hi[a2] + 1):
                    p = [0, 0, 0]
                    p[ax], p[a1], p[a2] = side, i, j
                    pts.append(p)
    rng = np.random.default_rng(99)
    if shuffle:  # any point order: the engine sorts internally and returns caller order
        pts = [pts[i] for i in rng.permutation(len(pts))]
    primal = [rng.uniform(0.8e-3, 1.2e-3, d).astype(np.float32) for d in dims]
    dual = [rng.uniform(0.8e-3, 1.2e-3, d).astype(np.float32) for d in dims]
    return np.array(pts, np.uint32).ravel(), primal, dual


@pytest.mark.parametrize("kblock,steps,window,shuffle", [(64, 100, 0, False), (7, 45, 1, False), (16, 16, 0, False),
                                                         (16, 40, 0, True)])
def test_nf2ff_accumulation(HIPEngine, oracle_mod, kblock, steps, window, shuffle):
    dims = (14, 13, 15)
    case = random_case(21, dims, pml_width=2)
    pts, primal, dual = nf2ff_setup(dims, shuffle=shuffle)
    dT = 1.9e-12
    omegas = 2 * np.pi * np.array([1.0e9, 2.5e9, 4.0e9, 6.3e9, 9.0e9, 11e9, 12e9, 20e9, 31e9])
    say("case nf2ff", kblock, steps, window)
    o = make_oracle(oracle_mod, case, fp64=False)
    o.configureNF2FF(pts, omegas, dT, primal, dual, maxTS=steps, windowType=window)
    g = make_gpu(HIPEngine, case)
    g.configureNF2FFAccumulation({"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": list(omegas),
                                  "dT": dT, "primalEdgeLens": primal, "dualEdgeLens": dual, "maxTS": steps,
                                  "windowType": window, "kBlock": kblock})
    o.iterate(steps)
    g.iterate(steps)
    oe, oh = o.readNF2FF()
    acc = g.readNF2FFAccumulation()
    for F in range(len(omegas)):
        for got, ref, nm in ((acc["accumE"][F], oe[F], "E"), (acc["accumH"][F], oh[F], "H")):
            err = rel_l2(got, ref)
            assert err < 2e-5, (F, nm, err)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_nf2ff_zface_capture(HIPEngine, oracle_mod, monkeypatch, storage):
    """Surface points on z planes read slabs the H sweep wrote (the planes
    have z and z+1 in one 4-cell lane, the rows no UPML): the accumulators
    equal those of the plain gather bit for bit, and the oracle's within the
    FP32 tolerance."""
    dims = (30, 28, 40)
    case = random_case(31, dims, pml_width=4)  # UPML z-bounds multiples of 4 (FP16)
    pts, primal, dual = nf2ff_setup(dims, inset=5)
    dT = 1.9e-12
    omegas = 2 * np.pi * np.array([1.0e9, 3.3e9, 7.0e9])
    steps = 70
    cfg = {"surfaceIndices": pts, "omega": list(omegas), "dT": dT, "primalEdgeLens": primal,
           "dualEdgeLens": dual, "kBlock": 16}
    out = {}
    for cap in ("1", "0"):
        monkeypatch.setenv("EFS_NF2FF_CAPTURE", cap)
        say("case nf2ff capture", cap, storage)
        g = HIPEngine()
        assert g.initGPU()
        g.init(case["dims"], case["coeffs"], storage=storage)
        configure(g, case)
        g.configureNF2FFAccumulation(cfg)
        assert ("nf2ff_capture" in g.pathFlags()) == (cap == "1")
        g.iterate(steps)
        out[cap] = g.readNF2FFAccumulation()
        g.destroy()
    for k in ("accumE", "accumH"):
        np.testing.assert_array_equal(np.asarray(out["1"][k]), np.asarray(out["0"][k]))
    if storage == "fp32":
        o = make_oracle(oracle_mod, case, fp64=False)
        o.configureNF2FF(pts, omegas, dT, primal, dual)
        o.iterate(steps)
        refE, refH = o.readNF2FF()
        for f in range(len(omegas)):
            got = np.concatenate([np.asarray(out["1"]["accumE"][f]), np.asarray(out["1"]["accumH"][f])])
            want = np.concatenate([refE[f], refH[f]])
            assert rel_l2(got, want) <= 2e-5


@pytest.mark.parametrize("storage,pml", [("fp32", 4), ("fp16", 4), ("fp32", 0)])
def test_side_work_same_bits(HIPEngine, monkeypatch, storage, pml):
    """NF2FF sampling, the buffered probe gather and the Mur pre pass run by
    extra workgroups of the sweep kernels (the E half of the sample beside
    the H sweep, the H half beside the next E sweep or in the flush at the
    end of an iterate call) give the fields, accumulators and gather samples
    of the separate point launches bit for bit: z-face capture, Mur, FP16
    range control every 5 steps, gathers every step and every 3rd step,
    iterate calls of 1 and of several steps."""
    dims = (30, 28, 40)
    case = random_case(41, dims, pml_width=pml, mur=True)
    pts, primal, dual = nf2ff_setup(dims, inset=5)
    dT = 1.9e-12
    omegas = 2 * np.pi * np.array([1.0e9, 3.3e9, 7.0e9])
    cfg = {"surfaceIndices": pts, "omega": list(omegas), "dT": dT, "primalEdgeLens": primal,
           "dualEdgeLens": dual, "kBlock": 8}
    N = int(np.prod(dims))
    probes = np.array([5, 7 * 40 + 3, N + 1234, 2 * N + 999, 17 * 28 * 40 + 11 * 40 + 20], np.uint32)
    monkeypatch.setenv("EFS_RANGE_INTERVAL", "5")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("EFS_SIDE_WORK", mode)
        say("case side work", mode, storage, pml)
        res = []
        for interval in (1, 3):
            g = HIPEngine()
            assert g.initGPU()
            g.init(case["dims"], case["coeffs"], storage=storage)
            configure(g, case)
            g.configureNF2FFAccumulation(cfg)
            g.configureProbeGatherBuffered(probes, 64, interval)
            assert ("side_work" in g.pathFlags()) == (mode == "1")
            for n in (13, 1, 1, 30, 2):
                g.iterate(n)
            f = g.getFields()
            acc = g.readNF2FFAccumulation()
            gb = g.readProbeGatherBuffered()
            res += [np.asarray(f["volt"]), np.asarray(f["curr"]), np.asarray(acc["accumE"]),
                    np.asarray(acc["accumH"]), np.asarray(gb["data"]), np.array([gb["numSamples"]])]
            g.destroy()
        out[mode] = res
    for a, b in zip(out["1"], out["0"]):
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a, b)


def test_farfield_matches_fp64_oracle(HIPEngine, oracle_mod):
    dims = (14, 13, 15)
    case = random_case(22, dims, pml_width=2)
    pts, primal, dual = nf2ff_setup(dims, shuffle=True)
    dT = 1.9e-12
    freq = 3.1e9
    g = make_gpu(HIPEngine, case)
    g.configureNF2FFAccumulation({"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": [2 * np.pi * freq],
                                  "dT": dT, "primalEdgeLens": primal, "dualEdgeLens": dual})
    say("case farfield")
    g.iterate(60)
    acc = g.readNF2FFAccumulation()
    P = len(pts) // 3
    rng = np.random.default_rng(5)
    meta = np.zeros((P, 8), np.float32)
    meta[:, :3] = pts.reshape(P, 3) * 1e-3
    meta[:, 3] = rng.integers(0, 3, P)
    meta[:, 4] = rng.choice([-1.0, 1.0], P)
    meta[:, 5] = rng.uniform(0.5e-6, 1.5e-6, P)
    theta = np.linspace(0, np.pi, 19, dtype=np.float32)
    phi = np.linspace(0, 2 * np.pi, 24, endpoint=False, dtype=np.float32)
    center = [7e-3, 6.5e-3, 7.5e-3]
    out = g.computeNF2FFfarField({"pointMeta": meta.ravel(), "theta": theta, "phi": phi, "center": center,
                                  "frequency": freq, "radius": 1.0})
    ref = oracle_mod.farfield(acc["accumE"][0].astype(np.float64), acc["accumH"][0].astype(np.float64), meta.ravel(),
                              theta, phi, center, freq, 1.0)
    out = out.reshape(-1, 5)
    ref = ref.reshape(-1, 5)
    assert rel_l2(out[:, 0], ref[:, 0]) < 1e-5
    assert rel_l2(out[:, 1:], ref[:, 1:]) < 1e-5
    # the mirror option (image in a PEC / PMC plane; the reference's rule by
    # default, image-theory normals on request) against the FP64 oracle's
    # restatement
    for mirror in ({"type": "PEC", "direction": 2, "position": 1e-3},
                   {"type": "PMC", "direction": 0, "position": -2e-3, "imageNormals": True},
                   {"type": "PEC", "direction": 1, "position": 0.0, "imageNormals": True}):
        out = g.computeNF2FFfarField({"pointMeta": meta.ravel(), "theta": theta, "phi": phi, "center": center,
                                      "frequency": freq, "radius": 1.0, "mirror": mirror}).reshape(-1, 5)
        ref = oracle_mod.farfield(acc["accumE"][0].astype(np.float64), acc["accumH"][0].astype(np.float64),
                                  meta.ravel(), theta, phi, center, freq, 1.0, mirror=mirror).reshape(-1, 5)
        assert rel_l2(out[:, 0], ref[:, 0]) < 1e-5, mirror
        assert rel_l2(out[:, 1:], ref[:, 1:]) < 1e-5, mirror


def test_energy_and_lag(HIPEngine, oracle_mod):
    case = random_case(23, (30, 27, 33), pml_width=3)
    o = make_oracle(oracle_mod, case)
    g = make_gpu(HIPEngine, case)
    say("case energy")
    e_initial = o.energy()
    assert g.computeEnergy() == 0.0  # first call returns 0, as the reference's lagged readback
    for _ in range(3):
        o.iterate(10)
        g.iterate(10)
        e_ref = o.energy()
        e_gpu = g.energy()
        assert abs(e_gpu - e_ref) <= 1e-5 * abs(e_ref), (e_gpu, e_ref)
    lag1 = g.computeEnergy()  # value measured by the first call (step 0)
    lag2 = g.computeEnergy()  # value measured by the previous call (now)
    now = g.energy()
    assert abs(lag1 - e_initial) <= 1e-6 * e_initial
    assert abs(lag2 - now) <= 1e-12 * now
    assert abs(g.flushEnergy() - now) <= 1e-12 * now


def test_probe_gather_buffered(HIPEngine, oracle_mod):
    case = random_case(24, (12, 11, 10), pml_width=2, mur=True)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, 3 * 12 * 11 * 10, 9).astype(np.uint32)
    o = make_oracle(oracle_mod, case)
    g = make_gpu(HIPEngine, case)
    o.configureProbeGatherBuffered(idx, 6, 4)
    g.configureProbeGatherBuffered(idx, 6, 4)
    g.configureProbeGather(idx)
    say("case gather")
    o.iterate(30)
    g.iterate(30)
    ref = o.readProbeGatherBuffered()
    got = g.readProbeGatherBuffered()
    assert got["numSamples"] == ref.shape[0] == 6  # steps 4,8,..,24 -> 6 samples, capped at 6
    assert got["stride"] == 18
    assert rel_l2(got["data"], ref.ravel()) <= TOL
    now = g.readProbeGather()
    f = g.getFields()
    np.testing.assert_array_equal(now[0::2], f["volt"][idx])
    np.testing.assert_array_equal(now[1::2], f["curr"][idx])


def test_invalid_configuration_is_rejected(HIPEngine):
    """Out-of-range indices never reach the GPU: the engine raises."""
    case = random_case(25, (6, 5, 4))
    g = make_gpu(HIPEngine, case)
    N = 6 * 5 * 4
    with pytest.raises(RuntimeError):
        g.configureExcitation({"signal": [1.0], "delay": [0], "amp": [1.0], "dir": [3], "pos": [0]})
    with pytest.raises(RuntimeError):
        g.configureExcitation({"signal": [1.0], "delay": [0], "amp": [1.0], "dir": [0], "pos": [N]})
    with pytest.raises(RuntimeError):
        g.configureMur({"coeff": 0.1, "normal_idx": [3 * N], "shifted_idx": [0]})
    with pytest.raises(RuntimeError):
        g.configurePML([{"startPos": [0, 0, 0], "numLines": [7, 1, 1], "vv": np.ones(21), "vvfo": np.ones(21),
                         "vvfn": np.ones(21), "ii": np.ones(21), "iifo": np.ones(21), "iifn": np.ones(21)}])
    ones = np.ones(3 * 4)
    with pytest.raises(RuntimeError):  # overlapping regions
        g.configurePML([{"startPos": [0, 0, 0], "numLines": [2, 2, 1], "vv": ones, "vvfo": ones, "vvfn": ones,
                         "ii": ones, "iifo": ones, "iifn": ones}] * 2)
    fresh = HIPEngine()
    assert fresh.initGPU()
    with pytest.raises(RuntimeError):
        fresh.iterate(1)  # before init
    g.iterate(2)  # engine still usable after rejected calls
    assert g.numTS == 2


@pytest.mark.parametrize("upload_after", [False, True])
def test_widened_upml_regions(HIPEngine, oracle_mod, monkeypatch, upload_after):
    """openEMS's PML_n slabs are n + 1 lines deep (efs_amd.operator.
    pml_regions_for_shell), so their z-bounds miss the 4-cell lanes; the
    engine widens them to lane boundaries with identity padding cells whose
    flux follows the field.  The 4-wide fused sweeps then give the bits of
    the 1-wide sweeps on the regions as given (EFS_VEC=1), whether the
    fields were uploaded before the UPML was configured or after, and the
    oracle's fields within the FP32 tolerance; FP16 storage takes them."""
    from efs_amd import operator as op
    from cases import random_pml
    dims = (20, 18, 24)
    regions = op.pml_regions_for_shell(op.uniform_mesh(dims, 1e-3), [4] * 6)
    assert any(s[2] % 4 or n[2] % 4 for s, n in regions)
    case = random_case(71, dims, mur=False)
    case["pml"] = random_pml(np.random.default_rng(72), regions)
    steps = 40

    def run(storage="fp32"):
        g = HIPEngine()
        assert g.initGPU()
        g.init(case["dims"], case["coeffs"], storage=storage)
        if not upload_after:
            g.uploadFields(case["volt0"], case["curr0"])
        g.configurePML(case["pml"])
        if upload_after:
            g.uploadFields(case["volt0"], case["curr0"])
        g.configureExcitation(case["excitation"])
        flags = g.pathFlags()
        g.iterate(steps)
        f = g.getFields()
        g.destroy()
        return flags, f
    say("case widened upml, upload after", upload_after)
    flags4, f4 = run()
    assert {"vec4", "upml4", "fused"} <= flags4
    monkeypatch.setenv("EFS_VEC", "1")
    flags1, f1 = run()
    assert "vec4" not in flags1
    monkeypatch.delenv("EFS_VEC")
    for k in ("volt", "curr"):
        np.testing.assert_array_equal(f4[k], f1[k])
    o = oracle_mod.Oracle(case["dims"], case["coeffs"])
    o.uploadFields(case["volt0"], case["curr0"])
    o.configurePML(case["pml"])
    o.configureExcitation(case["excitation"])
    o.iterate(steps)
    assert_fields_close(f4, o.getFields(), what="widened upml")
    # FP16 storage takes the widened regions on the fused path (its accuracy
    # on them is the stress-grid test of tests/test_gpu_fp16.py, whose UPML
    # slabs are widened too; this case's random coefficients grow without
    # bound, which FP16 cannot hold)
    flags16, _ = run("fp16")
    assert "f16" in flags16 and "fused" in flags16
