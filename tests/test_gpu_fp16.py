"""FP16 field storage (per-component power-of-two scales, FP32 updates):
error against the FP64 oracle, which is the precision form's reference.

Tolerances are stated per case.  FP16 keeps an 11-bit significand: each
stored value carries a relative rounding of 2^-12, and in a well-resolved
wave the per-step increment vi*curl is small against the stored value, so
the rounding of the increments accumulates -- about 1e-3 of the field level
for E and 1e-2 for H after 150 steps on the stress grid.  Quantities
integrated over time are much closer: the NF2FF spectra and the energy stay
within a few 1e-4.  The FP32 engine on the same case is within 1e-5.
"""
import numpy as np
import pytest

from cases import configure, random_case, rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def HIPEngine():
    from efs_amd import engine
    engine.load_library()
    if engine.device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    return engine.HIPEngine


@pytest.fixture(scope="module")
def small_stress():
    from efs_amd import workloads
    return workloads.stress_grid(48, pml=8, nfreq=3, n_boxes=6, n_plates=2)


def stress_engine(HIPEngine, w, storage):
    g = HIPEngine()
    assert g.initGPU()
    g.init(w["gridSize"], w["coefficients"], storage=storage)
    g.configurePML(w["pmlRegions"])
    g.configureExcitation(w["excitation"])
    g.configureNF2FFAccumulation(w["nf2ff"])
    return g


def stress_oracle(oracle_mod, w):
    o = oracle_mod.Oracle(w["gridSize"], w["coefficients"], fp64=True)
    o.configurePML(w["pmlRegions"])
    o.configureExcitation(w["excitation"])
    nf = w["nf2ff"]
    o.configureNF2FF(nf["surfaceIndices"], nf["omega"], nf["dT"], nf["primalEdgeLens"], nf["dualEdgeLens"])
    return o


def test_fp16_stress_grid_against_fp64(HIPEngine, oracle_mod, small_stress):
    """48^3 stress grid (UPML-8, dielectric boxes, PEC plates, lumped port,
    NF2FF at 3 frequencies), 150 steps."""
    w = small_stress
    say("case fp16 stress 48^3")
    o = stress_oracle(oracle_mod, w)
    o.iterate(150)
    ref = o.getFields()
    e_ref = o.energy()
    accE, accH = o.readNF2FF()
    g16 = stress_engine(HIPEngine, w, "fp16")
    assert {"f16", "fused", "upml_idx"} <= g16.pathFlags()
    g16.iterate(150)
    f = g16.getFields()
    err_v, err_c = rel_l2(f["volt"], ref["volt"]), rel_l2(f["curr"], ref["curr"])
    err_e = abs(g16.energy() - e_ref) / e_ref
    acc = g16.readNF2FFAccumulation()
    err_nf = max(rel_l2(np.concatenate([acc["accumE"][k], acc["accumH"][k]]),
                        np.concatenate([accE[k], accH[k]])) for k in range(3))
    say(f"fp16 vs fp64: volt {err_v:.2e} curr {err_c:.2e} energy {err_e:.2e} nf2ff {err_nf:.2e}")
    assert err_v < 3e-3 and err_c < 2e-2 and err_e < 1e-3 and err_nf < 2e-3
    # the FP32 engine on the same case, for scale
    g32 = stress_engine(HIPEngine, w, "fp32")
    g32.iterate(150)
    f32 = g32.getFields()
    assert rel_l2(f32["volt"], ref["volt"]) < 1e-5 and rel_l2(f32["curr"], ref["curr"]) < 1e-5


@pytest.mark.parametrize("name,seed,dims,kw", [
    ("f16_all", 71, (20, 16, 24), dict(pml_width=4, mur=True, ade=True, tfsf=True, rlc=True, pbc=[(1, 0.4)])),
    ("f16_pq", 72, (24, 20, 32), dict(pml_width=8, pml_quantize=3, quantize=6, mur=True)),
    ("f16_core", 73, (16, 12, 20), dict()),
])
def test_fp16_all_extensions(HIPEngine, oracle_mod, monkeypatch, name, seed, dims, kw):
    """Every extension on FP16 storage.  The random coefficients let the
    fields grow by orders of magnitude within a few steps, so the range
    check runs every step here (EFS_RANGE_INTERVAL=1); 30 steps."""
    monkeypatch.setenv("EFS_RANGE_INTERVAL", "1")
    case = random_case(seed, dims, **kw)
    say("case", name)
    o = oracle_mod.Oracle(case["dims"], case["coeffs"], fp64=True)
    configure(o, case)
    g = HIPEngine()
    assert g.initGPU()
    g.init(case["dims"], case["coeffs"], storage="fp16")
    configure(g, case)
    o.iterate(30)
    g.iterate(30)
    f, r = g.getFields(), o.getFields()
    ev, ec = rel_l2(f["volt"], r["volt"]), rel_l2(f["curr"], r["curr"])
    say(f"{name}: volt {ev:.2e} curr {ec:.2e}")
    assert np.isfinite(f["volt"]).all() and ev < 5e-3 and ec < 5e-3


def test_fp16_scales_follow_the_fields(HIPEngine):
    """Uploaded fields of very different magnitudes per component: each gets
    its own power-of-two scale and the values come back to FP16 precision."""
    dims = (8, 8, 12)
    N = 8 * 8 * 12
    rng = np.random.default_rng(4)
    v = rng.normal(size=3 * N).astype(np.float32)
    c = rng.normal(size=3 * N).astype(np.float32)
    v[:N] *= 1e-6
    v[N:2 * N] *= 1e3
    c[2 * N:] *= 1e-9
    co = {"vv": np.ones(3 * N, np.float32), "vi": np.zeros(3 * N, np.float32),
          "ii": np.ones(3 * N, np.float32), "iv": np.zeros(3 * N, np.float32)}
    g = HIPEngine()
    assert g.initGPU()
    g.init(dims, co, storage="fp16")
    g.uploadFields(v, c)
    sc = g.fieldScales()
    assert sc[0] < 1e-8 and sc[1] > 0.1 and sc[5] < 1e-11
    assert np.all(sc == 2.0 ** np.round(np.log2(sc)))  # powers of two
    g.iterate(5)  # static fields (vi = iv = 0): values unchanged
    f = g.getFields()
    for a, b in ((f["volt"], v), (f["curr"], c)):
        for n in range(3):
            seg = slice(n * N, (n + 1) * N)
            assert np.max(np.abs(a[seg] - b[seg])) <= 2.0 ** -11 * np.max(np.abs(b[seg]))


def test_fp16_needs_aligned_upml(HIPEngine, monkeypatch):
    """Any Nz works (device rows are padded to a multiple of 4); UPML
    regions whose z-bounds miss the 4-cell lanes are widened to them (the
    fused sweeps run), and are refused only with the widening switched off
    (EFS_PML_WIDEN=0); the 1-wide sweeps (forced) have no FP16 form."""
    g = HIPEngine()
    assert g.initGPU()
    N = 8 * 8 * 10
    co = {k: np.ones(3 * N, np.float32) for k in ("vv", "vi", "ii", "iv")}
    g.init((8, 8, 10), co, storage="fp16")
    assert "f16" in g.pathFlags() and "vec4" in g.pathFlags()
    monkeypatch.setenv("EFS_VEC", "1")
    with pytest.raises(RuntimeError, match="4-wide"):
        g.init((8, 8, 10), co, storage="fp16")
    monkeypatch.delenv("EFS_VEC")
    case = random_case(5, (12, 12, 16), pml_width=3)
    g.init(case["dims"], case["coeffs"], storage="fp16")
    g.configurePML(case["pml"])
    assert {"f16", "fused"} <= g.pathFlags()
    monkeypatch.setenv("EFS_PML_WIDEN", "0")
    g.init(case["dims"], case["coeffs"], storage="fp16")
    with pytest.raises(RuntimeError, match="4-cell boundaries"):
        g.configurePML(case["pml"])
