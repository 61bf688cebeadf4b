"""Graded mesh lines (efs_amd/meshing.py) and the patch-antenna workload.

Pins:
- the reference's patch example (app/examples.mjs:17-103) meshes to
  86 x 87 x 71 lines (README.md:133, STATUS.md:132), and its NF2FF box
  (3 lines inside a Mur boundary, app/ems-api.mjs:908-975) holds 33,890
  surface points -- both reproduced only when the smoother places every
  line as the reference's does;
- properties every smoothed set must have: the fixed lines survive, no
  spacing exceeds max_res, neighbouring spacings grow by at most the ratio
  away from the fixed lines, and a symmetric input gives a symmetric mesh;
- the reference's quick checks of a smoother (tests/test_api.mjs:1059-1083:
  [0, 1, 5, 6] at 2 gains lines with no gap above 2, [0, 1, 2, 3] stays,
  a single line stays).
"""
import numpy as np
import pytest

from efs_amd import meshing, workloads


def test_patch_mesh_matches_reference_size():
    w = workloads.patch_antenna()
    assert w["gridSize"] == [86, 87, 71]
    assert w["nf2ff"]["numPoints"] == 33890


@pytest.mark.parametrize("lines,res,ratio", [
    ([-100, 100, -16, 16, -30, 30], 2.5, 1.4),
    ([-50, 100, 0, 0.381, 0.762, 1.143, 1.524], 2.5, 1.4),
    ([0, 3.3, 3.5, 40, 41], 1.7, 1.5),
    ([0, 1, 5, 6], 2.0, 1.5),
])
def test_smooth_properties(lines, res, ratio):
    out = np.array(meshing.smooth_mesh_lines(lines, res, ratio))
    assert np.all(np.diff(out) > 0)
    for v in lines:
        assert np.min(np.abs(out - v)) < 1e-9
    d = np.diff(out)
    assert d.max() <= res * (1 + 1e-9)
    # growth: a spacing larger than its neighbour by more than the ratio only
    # where the neighbour was pinned by two fixed lines
    fixed = np.array(sorted(set(lines)))
    for k in range(len(d) - 1):
        big, small = max(d[k], d[k + 1]), min(d[k], d[k + 1])
        if big > small * ratio * (1 + 1e-6):
            j = k if d[k] < d[k + 1] else k + 1
            ends = out[j], out[j + 1]
            assert all(np.min(np.abs(fixed - e)) < 1e-9 for e in ends), (k, d[k], d[k + 1])


def test_symmetric_input_gives_symmetric_mesh():
    for lines in ([-100, 100, -16, 16, -30, 30], [-10, -1, 0, 1, 10]):
        out = np.array(meshing.smooth_mesh_lines(lines, 2.5, 1.4))
        c = 0.5 * (out[0] + out[-1])
        np.testing.assert_allclose(out - c, -(out[::-1] - c), atol=1e-9)


def test_reference_quick_checks():
    s = meshing.smooth_mesh_lines([0, 1, 5, 6], 2)
    assert len(s) > 4 and max(np.diff(s)) <= 2.0 + 1e-10
    assert len(meshing.smooth_mesh_lines([0, 1, 2, 3], 2)) == 4
    assert len(meshing.smooth_mesh_lines([5], 1)) == 1


def test_unique_and_symmetry_helpers():
    assert meshing.unique_lines([3, 1, 2, 2, 1 + 1e-12]) == [1 + 1e-12, 2, 3]
    assert meshing.check_symmetry([-2, -1, 1, 2]) == 2
    assert meshing.check_symmetry([-2, 0, 2]) == 1
    assert meshing.check_symmetry([-2, 0, 3]) == 0
    assert meshing.check_symmetry([0, 1]) == 0


def test_patch_port_and_boundaries():
    w = workloads.patch_antenna()
    nx, ny, nz = w["gridSize"]
    ut, it = w["probes"]
    # 4 substrate edges in series along z at the feed
    assert ut["stop"][2] - ut["start"][2] == 4 and len(w["excitation"]["pos"]) == 4
    assert np.all(w["excitation"]["amp"] < 0)
    assert it["normDir"] == 2 and it["stop"][0] - it["start"][0] == 1
    # Mur on six faces: two tangential components per face point
    assert len(w["mur"]["normal_idx_nyP"]) == 2 * (ny * nz + nx * nz + nx * ny)
    c = w["coefficients"]
    N = nx * ny * nz
    vi = c["vi"].reshape(3, nx, ny, nz)
    # PEC patch and ground: tangential E zero on their planes
    assert vi[0, 40, 43, 27] == 0.0 and vi[0, 40, 43, 23] == 0.0
    assert np.all(np.isfinite(c["vi"])) and c["vi"].size == 3 * N
