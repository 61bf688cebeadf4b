"""The batched form (include/efs_batch.h) on the GPU.

- Each instance of a batch gives the same bits as the single-grid engine
  run alone on the same arrays (fields and probe traces): both run the same
  per-cell fused multiply-adds, the batch only changes the launch shape and
  pads Nz to a multiple of 4 (instances here have Nz % 4 in {0, 2, 3}).
- Graph replays (blocks of 32 steps) give the same bits as direct launches.
- Mur on the z faces applied inside the E sweep gives the same bits as the
  Mur point kernels (FP32 and FP16).
- The autotuner leaves a reset batch whose results do not depend on the
  chosen block table.
- Every launch shape gives the same bits: the persistent kernel (grid
  barriers between the phases, both barrier forms) and the launch form at
  each column tile width, in FP32 and FP16 (the FP16 range control runs
  inside the persistent launch), over iterate calls that start and end off
  the 32- and 64-step boundaries.
- FP16 storage against the FP64 oracle per instance: relative L2 of the
  port traces and fields within the stated tolerances.
- Misuse fails with a message and launches nothing.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 150


def say(*a):
    print(*a, flush=True)


def small_patch(W):
    from efs_amd import workloads
    return workloads.patch_antenna(patch_w=W, patch_l=1.25 * W, sub_w=1.875 * W, sub_l=1.875 * W,
                                   feed_x=-0.1875 * W, sim_box=(3 * W + 20, 3 * W + 16, 2 * W + 10))


@pytest.fixture(scope="module")
def patches():
    return [small_patch(W) for W in (14.0, 17.0, 20.0, 23.0)]


def make_batch(ws, storage="fp32", steps=STEPS):
    from efs_amd import batch, workloads
    b = batch.BatchEngine(storage=storage)
    sl = [workloads.add_patch(b, w)[1] for w in ws]
    b.finalize(1, steps)
    return b, sl


def traces(r, w, slices):
    from efs_amd import ports
    return ports.probe_traces(r["data"], r["numSamples"], r["stride"], w["probes"], slices, 1, w["dT"])


def test_batch_matches_single_engine(patches):
    from efs_amd import engine, workloads
    engine.load_library()
    say("instances", [w["gridSize"] for w in patches], "steps", STEPS)
    b, sl = make_batch(patches)
    info = b.info()
    say("info", info)
    assert info["graph_steps"] == 32 and info["instances"] == len(patches)
    b.iterate(STEPS)
    assert b.numTS == STEPS
    for i, w in enumerate(patches):
        g = engine.HIPEngine()
        assert g.initGPU()
        workloads.configure_patch(g, w, nf2ff=False, max_samples=STEPS)
        g.iterate(STEPS)
        fs, fb = g.getFields(), b.getFields(i)
        for k in ("volt", "curr"):
            np.testing.assert_array_equal(fb[k], fs[k], err_msg=f"instance {i} {k}")
        rs, rb = g.readProbeGatherBuffered(), b.readProbeGatherBuffered(i)
        assert rb["numSamples"] == rs["numSamples"] == STEPS
        np.testing.assert_array_equal(rb["data"], rs["data"])
        g.destroy()
    b.destroy()


def test_graph_replay_matches_direct_launches(patches):
    ws = patches[:2]
    a, _ = make_batch(ws, steps=100)
    d, _ = make_batch(ws, steps=100)
    d.profile(True)  # profiling launches every step directly
    a.iterate(100)
    d.iterate(100)
    p = d.profileRead()
    assert p["e_n"] == 100 and p["h_n"] == 100
    for i in range(len(ws)):
        fa, fd = a.getFields(i), d.getFields(i)
        for k in ("volt", "curr"):
            np.testing.assert_array_equal(fa[k], fd[k])
        np.testing.assert_array_equal(a.readProbeGatherBuffered(i)["data"], d.readProbeGatherBuffered(i)["data"])
    a.destroy()
    d.destroy()


def test_tune_then_same_results(patches):
    ws = patches[1:3]
    a, _ = make_batch(ws, steps=64)
    t, _ = make_batch(ws, steps=64)
    shape = t.tune(steps=4)
    say("tuned", shape, t.info())
    assert shape["mode"] in ("launch", "persistent") and shape["blocks"] > 0
    assert shape["waves"] in ((4, 5) if shape["mode"] == "launch" else (3,))
    assert t.numTS == 0
    z = t.getFields(0)
    assert not np.any(z["volt"]) and not np.any(z["curr"])
    a.iterate(64)
    t.iterate(64)
    for i in range(len(ws)):
        fa, ft = a.getFields(i), t.getFields(i)
        for k in ("volt", "curr"):
            np.testing.assert_array_equal(fa[k], ft[k])
    a.destroy()
    t.destroy()


SHAPES = [("launch", 128, 4, 4, 1024), ("launch", 512, 4, 5, 4096), ("persistent", 256, 4, 3, 768),
          ("persistent", 256, 4, 3, 6000), ("persistent-flat", 256, 4, 3, 2048)]
SHAPES_Z8 = [("launch", 256, 8, 3, 2048), ("launch", 512, 8, 3, 1024)]


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_launch_shapes_same_bits(patches, storage, monkeypatch):
    """The reference shape (launch form, 256-lane tiles of 4-cell lanes, 4
    waves, 2048 blocks) against other tile widths, 8-cell lanes (FP16) and
    the persistent kernel, over
    iterate calls of 37 and 70 steps (off the graph and persistent-launch
    boundaries; FP16 crosses three range checks)."""
    ws = patches[:3]
    calls = (37, 70)
    ref, _ = make_batch(ws, storage=storage, steps=sum(calls))
    say("reference shape", ref.shape())
    for n in calls:
        ref.iterate(n)
    ref.sync()
    want = [ref.getFields(i) for i in range(len(ws))]
    want_g = [ref.readProbeGatherBuffered(i)["data"] for i in range(len(ws))]
    ref.destroy()
    for mode, tpb, zc, waves, blocks in SHAPES + SHAPES_Z8:
        if mode == "persistent-flat":
            monkeypatch.setenv("EFS_BATCH_BARRIER", "flat")
        else:
            monkeypatch.delenv("EFS_BATCH_BARRIER", raising=False)
        b, _ = make_batch(ws, storage=storage, steps=sum(calls))
        b.setShape(mode.split("-")[0], tpb, zc, waves, blocks)
        say("shape", mode, b.shape())
        for n in calls:
            b.iterate(n)
        b.sync()
        assert b.numTS == sum(calls)
        for i in range(len(ws)):
            got = b.getFields(i)
            for k in ("volt", "curr"):
                np.testing.assert_array_equal(got[k], want[i][k], err_msg=f"{mode} {tpb} instance {i} {k}")
            np.testing.assert_array_equal(b.readProbeGatherBuffered(i)["data"], want_g[i])
        b.destroy()


def test_shape_rejects_invalid(patches, monkeypatch):
    b, _ = make_batch(patches[:1], steps=4)
    for args in (("launch", 384, 4, 4, 100), ("launch", 256, 4, 3, 100), ("persistent", 512, 4, 3, 100),
                 ("persistent", 256, 4, 4, 100), ("launch", 256, 4, 4, 0), ("launch", 256, 8, 4, 100),
                 ("launch", 256, 6, 4, 100), ("launch", 128, 8, 3, 100)):
        with pytest.raises(RuntimeError):
            b.setShape(*args)
    b.destroy()
    # rows padded to 4 cells: no 8-cell lanes
    monkeypatch.setenv("EFS_BATCH_ZPAD", "4")
    b, _ = make_batch(patches[:1], steps=4)
    with pytest.raises(RuntimeError):
        b.setShape("launch", 256, 8, 3, 100)
    b.setShape("launch", 256, 4, 4, 100)
    b.destroy()


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_zface_mur_in_sweep_same_bits(patches, storage, monkeypatch):
    """The z-face Mur entries the E sweep applies itself give the same bits
    as the same entries run through the Mur point kernels
    (EFS_BATCH_ZMUR=0), in FP32 and FP16 storage."""
    ws = patches[:3]
    say("zmur", storage, [w["gridSize"] for w in ws])
    a, _ = make_batch(ws, storage=storage, steps=120)
    monkeypatch.setenv("EFS_BATCH_ZMUR", "0")
    d, _ = make_batch(ws, storage=storage, steps=120)
    sa, sd = a.murSplit(), d.murSplit()
    say("mur split", sa, sd)
    assert sa["in_sweep"] > 0 and sd["in_sweep"] == 0
    assert sa["listed"] + sa["in_sweep"] == sd["listed"]
    a.iterate(120)
    d.iterate(120)
    for i in range(len(ws)):
        fa, fd = a.getFields(i), d.getFields(i)
        for k in ("volt", "curr"):
            np.testing.assert_array_equal(fa[k], fd[k], err_msg=f"instance {i} {k}")
        np.testing.assert_array_equal(a.readProbeGatherBuffered(i)["data"], d.readProbeGatherBuffered(i)["data"])
    a.destroy()
    d.destroy()


def six_face_mur(dims, order, coef):
    """Mur lists on the six faces of a grid in the given face order: each face
    (axis a, side) lists its two tangential components over the whole face,
    shifted one cell inward, with coefficient coef[(a, side)]."""
    nx, ny, nz = dims
    N = nx * ny * nz
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    nidx, sidx, co = [], [], []
    for a, side in order:
        c = [x, y, z]
        on = (c[a] == (0 if side == 0 else dims[a] - 1)).ravel()
        sh = [v.copy() for v in c]
        sh[a] = sh[a] + (1 if side == 0 else -1)
        lin_n = (c[0] * ny * nz + c[1] * nz + c[2]).ravel()[on]
        lin_s = (sh[0] * ny * nz + sh[1] * nz + sh[2]).ravel()[on]
        for n in range(3):
            if n == a:
                continue
            nidx.append(n * N + lin_n)
            sidx.append(n * N + lin_s)
            co.append(np.full(lin_n.size, coef[(a, side)], np.float32))
    return {"normal_idx": np.concatenate(nidx).astype(np.uint32), "shifted_idx": np.concatenate(sidx).astype(np.uint32),
            "coeff": np.concatenate(co)}


def synthetic_instances():
    """Instances whose z-face Mur entries take every branch of the split:
    both faces in the sweep (Nz = 14), the Nz-1 face left in the lists
    ((Nz-1) % 4 == 0, Nz = 13), a face with an excitation cell on it and a
    face with one odd coefficient (both stay listed), Nz = 5 (nothing in the
    sweep) and the z faces listed first (the x and y faces then win the
    shared edge cells)."""
    import cases
    faces = [(a, s) for a in range(3) for s in (0, 1)]
    out = []
    for k, (dims, order) in enumerate([((10, 9, 14), faces), ((9, 8, 13), faces), ((8, 10, 11), faces),
                                       ((7, 7, 5), faces), ((6, 12, 10), faces[::-1])]):
        rng = np.random.default_rng(700 + k)
        N = int(np.prod(dims))
        coef = {f: float(np.float32(rng.uniform(0.2, 0.5))) for f in faces}
        mur = six_face_mur(dims, order, coef)
        pos = rng.integers(0, N, 3).astype(np.uint32)
        exc = {"signal": cases.f32(rng, 17, -1, 1), "delay": np.zeros(3, np.uint32), "amp": cases.f32(rng, 3, 0.5, 1),
               "dir": rng.integers(0, 3, 3).astype(np.uint32), "pos": pos, "period": 0}
        if k == 2:
            nx, ny, nz = dims
            # an Ex excitation cell on the z = 0 face, and one odd coefficient on
            # the z = Nz-1 face's Ey entries
            exc["pos"] = np.append(pos, np.uint32(cases.lin(dims, 3, 4, 0)))
            exc["dir"] = np.append(exc["dir"], np.uint32(0))
            exc["delay"] = np.append(exc["delay"], np.uint32(0))
            exc["amp"] = np.append(exc["amp"], np.float32(0.7))
            odd = np.flatnonzero(mur["normal_idx"] == N + cases.lin(dims, 2, 5, nz - 1))
            mur["coeff"][odd] = np.float32(0.123)
        out.append({"dims": dims, "coeffs": cases.random_coeffs(rng, N), "mur": mur, "exc": exc,
                    "probes": rng.integers(0, 3 * N, 5).astype(np.uint32)})
    return out


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_zface_mur_split_branches(storage, monkeypatch):
    """Synthetic instances through every branch of the z-face split: the
    batch with the split gives the bits of the batch without it, and in FP32
    the bits of the single-grid engine."""
    from efs_amd import batch, engine
    engine.load_library()
    insts = synthetic_instances()
    steps = 80

    def make():
        b = batch.BatchEngine(storage=storage)
        for c in insts:
            i = b.add(c["dims"], c["coeffs"])
            b.configureMur(i, c["mur"])
            b.configureExcitation(i, c["exc"])
            b.configureProbeGather(i, c["probes"])
        b.finalize(1, steps)
        return b
    say("split branches", storage, [c["dims"] for c in insts])
    a = make()
    monkeypatch.setenv("EFS_BATCH_ZMUR", "0")
    d = make()
    monkeypatch.delenv("EFS_BATCH_ZMUR")
    sa, sd = a.murSplit(), d.murSplit()
    say("mur split", sa, sd)
    assert sa["in_sweep"] > 0 and sd["in_sweep"] == 0 and sa["listed"] + sa["in_sweep"] == sd["listed"]
    a.iterate(steps)
    d.iterate(steps)
    for i, c in enumerate(insts):
        fa, fd = a.getFields(i), d.getFields(i)
        for k in ("volt", "curr"):
            assert np.all(np.isfinite(fa[k]))
            np.testing.assert_array_equal(fa[k], fd[k], err_msg=f"instance {i} {k}")
        ga = a.readProbeGatherBuffered(i)["data"]
        np.testing.assert_array_equal(ga, d.readProbeGatherBuffered(i)["data"])
        if storage == "fp32":
            g = engine.HIPEngine()
            assert g.initGPU()
            g.init(c["dims"], c["coeffs"])
            g.configureMur(c["mur"])
            g.configureExcitation(c["exc"])
            g.configureProbeGatherBuffered(c["probes"], steps, 1)
            g.iterate(steps)
            fs = g.getFields()
            for k in ("volt", "curr"):
                np.testing.assert_array_equal(fa[k], fs[k], err_msg=f"instance {i} {k} vs single engine")
            np.testing.assert_array_equal(ga, g.readProbeGatherBuffered()["data"])
            g.destroy()
    a.destroy()
    d.destroy()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_batch_fp16_against_fp64(patches, oracle_mod):
    """FP16 storage (FP32 updates) per instance against the FP64 oracle:
    port traces within 5e-3 relative L2, fields within 1e-2 (volt) and
    3e-2 (curr) after 400 steps."""
    from efs_amd import workloads
    steps = 400
    b, sl = make_batch(patches, storage="fp16", steps=steps)
    b.iterate(steps)
    for i, w in enumerate(patches):
        o = oracle_mod.Oracle(w["gridSize"], w["coefficients"], fp64=True)
        so = workloads.configure_patch(o, w, nf2ff=False, max_samples=steps)
        o.iterate(steps)
        r = o.readProbeGatherBuffered()
        to = traces({"data": r.ravel(), "numSamples": r.shape[0], "stride": r.shape[1]}, w, so)
        tb = traces(b.readProbeGatherBuffered(i), w, sl[i])
        fo, fb = o.getFields(), b.getFields(i)
        ev, ec = rel_l2(fb["volt"], fo["volt"]), rel_l2(fb["curr"], fo["curr"])
        eu, ei = rel_l2(tb[0][1], to[0][1]), rel_l2(tb[1][1], to[1][1])
        say("instance", i, w["gridSize"], "volt", ev, "curr", ec, "ut", eu, "it", ei)
        assert np.all(np.isfinite(fb["volt"])) and np.all(np.isfinite(fb["curr"]))
        assert eu <= 5e-3 and ei <= 5e-3
        assert ev <= 1e-2 and ec <= 3e-2
    b.destroy()


def test_batch_misuse(patches):
    from efs_amd import batch
    w = patches[0]
    b = batch.BatchEngine()
    with pytest.raises(RuntimeError, match="not finalized"):
        b.iterate(1)
    i = b.add(w["gridSize"], w["coefficients"])
    with pytest.raises(RuntimeError, match="before adding"):
        b._check(b.lib.efs_batch_set_storage(b.h, 1))
    with pytest.raises(RuntimeError, match="out of range"):
        b.configureMur(i, {"coeff": [0.5], "normal_idx": [10 ** 9], "shifted_idx": [0]})
    with pytest.raises(RuntimeError, match="does not exist"):
        b.configureProbeGather(7, [0])
    b.finalize(1, 4)
    with pytest.raises(RuntimeError, match="already finalized"):
        b.add(w["gridSize"], w["coefficients"])
    b.iterate(3)
    assert b.numTS == 3
    b.destroy()
