"""Sweep-shape tuner (efs_tune_sweeps): it changes only how the sweeps are
cut into workgroups, so an engine tuned after some steps must leave its
fields exactly as they were and then step bit for bit like an untuned one.

The tuner runs on grids without side work (larger than a few million
cells); EFS_SIDE_WORK=0 gives a 64^3 grid that path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def HIPEngine():
    from efs_amd import engine
    engine.load_library()
    if engine.device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    return engine.HIPEngine


@pytest.fixture(scope="module")
def stress64():
    from efs_amd import workloads
    return workloads.stress_grid(64, pml=8, nfreq=3, n_boxes=8, n_plates=3)


def engine_for(HIPEngine, w, storage):
    from efs_amd import workloads
    g = HIPEngine()
    assert g.initGPU()
    workloads.configure_engine(g, w, storage=storage)
    return g


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_tuned_engine_steps_like_untuned(HIPEngine, stress64, storage, monkeypatch):
    monkeypatch.setenv("EFS_SIDE_WORK", "0")
    a = engine_for(HIPEngine, stress64, storage)
    b = engine_for(HIPEngine, stress64, storage)
    try:
        assert "side_work" not in a.pathFlags()
        a.iterate(40)
        b.iterate(40)
        before = a.getFields()
        scale = a.tuneSweeps(3)
        print(f"{storage}: chosen factor {scale}", flush=True)
        assert scale in (0.5, 1.0, 2.0, 3.0, 4.0)
        after = a.getFields()
        for k in ("volt", "curr"):
            np.testing.assert_array_equal(after[k], before[k])
        a.iterate(60)
        b.iterate(60)
        fa, fb = a.getFields(), b.getFields()
        for k in ("volt", "curr"):
            assert np.isfinite(fa[k]).all()
            np.testing.assert_array_equal(fa[k], fb[k])
        ea, eb = a.readNF2FFAccumulation(), b.readNF2FFAccumulation()
        np.testing.assert_array_equal(np.asarray(ea["accumE"]), np.asarray(eb["accumE"]))
    finally:
        a.destroy()
        b.destroy()


def test_tuner_keeps_small_grids(HIPEngine):
    """A grid with side work keeps the planner's shape (factor 1)."""
    from efs_amd import workloads
    w = workloads.stress_grid(32, pml=4, nfreq=2, n_boxes=3, n_plates=1)
    g = engine_for(HIPEngine, w, "fp32")
    try:
        assert g.tuneSweeps(3) == 1.0
    finally:
        g.destroy()
