"""Dispersive scatterers under a TFSF plane wave on the HIP engine.

- Parity: a 48^3 version of the dispersive workload (Drude sphere,
  Lorentz slab, ADE with write-back, TFSF, UPML, NF2FF at 4 frequencies)
  against the FP32 oracle after 400 steps: fields relative L2 <= 1e-5,
  NF2FF accumulators <= 1e-4 (FP32 phase arithmetic over the steps).
- Physics: the bistatic RCS of a lone sphere (xy plane, 16 angles) from the
  engine's NF2FF far field against the Mie series of the same material
  (tests/mie.py).  Staircased spheres of 14-16 cells radius: within 1 dB
  (Lorentz sphere, eps about 6-7 at 2.5 and 3.5 GHz) and 1.5 dB (Drude
  sphere at 3 GHz, eps = -15) wherever the Mie RCS is within 15 dB of its
  maximum.  Frequencies near the spheres' plasmon resonances (eps near -2)
  are left out: there the staircase error is of order the resonance shift.
"""
import numpy as np
import pytest

from cases import rel_l2

import mie

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


def run_dispersive_chunks(HIPEngine, w, chunks, fused):
    from efs_amd import workloads
    g = HIPEngine()
    assert g.initGPU()
    workloads.configure_dispersive(g, w)
    for n in chunks:
        g.iterate(n)
    flags = g.pathFlags()
    assert ("fuse_tfsf_ade" in flags) == fused and ("fuse_tfsf_hooks" in flags) == fused, flags
    assert "side_work" not in flags
    f = g.getFields()
    acc = g.readNF2FFAccumulation()
    g.destroy()
    return f, acc


def test_fused_tfsf_ade_launches_are_bit_identical(HIPEngine, oracle_mod, monkeypatch):
    """TFSF-E with the ADE write-back in one launch, TFSF-H with the NF2FF
    sampling and the next step's ADE update in another (the path of grids
    without side work, forced here on 40^3), against the separate kernels
    (EFS_FUSE_POINTS=0): the same bits, over iterate calls of 1, 7, 64
    and 128 steps (the ADE update that runs ahead must not be repeated or
    lost at a call boundary); and within 1e-5 of the FP32 oracle."""
    from efs_amd import workloads
    w = workloads.dispersive_scatter(40, pml=6)
    chunks = (1, 7, 64, 128)
    monkeypatch.setenv("EFS_SIDE_WORK", "0")
    monkeypatch.setenv("EFS_FUSE_POINTS", "1")
    fa, aa = run_dispersive_chunks(HIPEngine, w, chunks, True)
    monkeypatch.setenv("EFS_FUSE_POINTS", "0")
    fb, ab = run_dispersive_chunks(HIPEngine, w, chunks, False)
    for k in ("volt", "curr"):
        np.testing.assert_array_equal(fa[k], fb[k])
    for k in ("accumE", "accumH"):
        np.testing.assert_array_equal(np.asarray(aa[k]), np.asarray(ab[k]))
    o = oracle_mod.Oracle(w["gridSize"], w["coefficients"], fp64=False)
    workloads.configure_dispersive(o, w)
    o.iterate(sum(chunks))
    ref = o.getFields()
    assert np.abs(ref["volt"]).max() > 1e-5
    for k in ("volt", "curr"):
        e = rel_l2(fa[k], ref[k])
        say(f"fused dispersive 40^3 vs FP32 oracle: {k} {e:.2e}")
        assert e <= 1e-5, k


def test_dispersive_matches_oracle(HIPEngine, oracle_mod):
    from efs_amd import workloads
    w = workloads.dispersive_scatter(48, pml=6)
    steps = 400
    say("dispersive 48^3", w["info"], "tfsf points", len(w["tfsf"]["voltagePoints"]["field_idx"]),
        len(w["tfsf"]["currentPoints"]["field_idx"]))
    g = HIPEngine()
    assert g.initGPU()
    workloads.configure_dispersive(g, w)
    g.iterate(steps)
    f = g.getFields()
    acc = g.readNF2FFAccumulation()
    g.destroy()
    o = oracle_mod.Oracle(w["gridSize"], w["coefficients"], fp64=False)
    workloads.configure_dispersive(o, w)
    o.iterate(steps)
    ref = o.getFields()
    aE, aH = o.readNF2FF()
    assert np.abs(ref["volt"]).max() > 1e-5
    for k in ("volt", "curr"):
        assert rel_l2(f[k], ref[k]) <= 1e-5, k
    assert rel_l2(acc["accumE"], aE) <= 1e-4
    assert rel_l2(acc["accumH"], aH) <= 1e-4


@pytest.mark.parametrize("kind,n,radius,freqs,tol", [
    ("lorentz", 96, 14.0, (2.5e9, 3.5e9), 1.0),
    ("drude", 112, 16.0, (3e9,), 1.5),
])
def test_sphere_rcs_matches_mie(HIPEngine, kind, n, radius, freqs, tol):
    from efs_amd import farfield, workloads
    w = workloads.dispersive_scatter(n, sphere=kind, slab=False, radius=radius, rcs_freqs=freqs, n_angles=16)
    sph = w["info"]["sphere"]
    say("mie", kind, n, sph)
    g = HIPEngine()
    assert g.initGPU()
    workloads.configure_dispersive(g, w)
    g.iterate(5000)
    g.sync()
    einc = farfield.incident_spectrum(w["signal"], 2 * np.pi * w["freqs"], w["dT"])
    centre = [c * 1e-3 for c in sph["centre"]]
    for k, fr in enumerate(w["freqs"]):
        ff = g.computeNF2FFfarField({"pointMeta": w["pointMeta"], "theta": w["theta"], "phi": w["phi"],
                                     "center": centre, "frequency": fr, "radius": 1.0, "freqIndex": k})
        got = 10 * np.log10(farfield.rcs(ff, einc[k]))
        ref = 10 * np.log10(mie.bistatic_rcs(mie.material_eps(fr, sph["material"]), radius * 1e-3, fr, w["phi"]))
        say(kind, fr, np.round(got, 2), np.round(ref, 2))
        sel = ref >= ref.max() - 15
        assert np.all(np.abs(got - ref)[sel] <= tol), (fr, got, ref)
    g.destroy()


def test_capture_with_tfsf(HIPEngine, monkeypatch):
    """TFSF current points off the NF2FF z planes leave the z-face capture
    on, with the bits of the plain gather; a z face of the NF2FF box inside
    the total-field box's z range (its y faces carry TFSF current points in
    the captured rows) turns it off."""
    from efs_amd import operator as op
    from efs_amd import workloads
    w = workloads.dispersive_scatter(48, rcs_freqs=(3e9, 7e9))  # UPML-8: split UPML, 4-wide sweeps
    out = {}
    for cap in ("1", "0"):
        monkeypatch.setenv("EFS_NF2FF_CAPTURE", cap)
        say("dispersive capture", cap)
        g = HIPEngine()
        assert g.initGPU()
        workloads.configure_dispersive(g, w)
        g.iterate(150)
        assert ("nf2ff_capture" in g.pathFlags()) == (cap == "1")
        out[cap] = g.readNF2FFAccumulation()
        g.destroy()
    for k in ("accumE", "accumH"):
        np.testing.assert_array_equal(np.asarray(out["1"][k]), np.asarray(out["0"][k]))
    monkeypatch.setenv("EFS_NF2FF_CAPTURE", "1")
    # total-field box 16..31; the captured z face is z = 28
    pts, _, primal, dual = op.nf2ff_box(w["mesh"], [15] * 3, [32, 32, 28])
    nf = dict(w["nf2ff"], surfaceIndices=pts, numPoints=len(pts) // 3, primalEdgeLens=primal, dualEdgeLens=dual)
    g = HIPEngine()
    assert g.initGPU()
    workloads.configure_dispersive(g, dict(w, nf2ff=nf))
    g.iterate(2)
    assert "nf2ff_capture" not in g.pathFlags()
    g.destroy()
