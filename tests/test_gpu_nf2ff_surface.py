"""computeNF2FF over surface fields on the GPU (efs_amd.nf2ff ->
efs_farfield_surface) against the FP64 oracle (oracle/nf2ff_surface.py,
pinned in tests/test_nf2ff_surface.py): Cartesian boxes and cylindrical
meshes, with and without PEC/PMC mirrors.  Bar: 1e-5 relative L2 on
E_theta and E_phi (the north-star's nf2ff.mjs bar), Dmax within 1e-5."""
import numpy as np
import pytest

from test_nf2ff_surface import ANG, box_faces, cyl_ring, cylinder_faces, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["box", "cylinder", "ring"])
@pytest.mark.parametrize("mirror", [None, {"type": "PEC", "direction": 2, "position": -0.02},
                                    {"type": "PMC", "direction": 0, "position": 0.03}])
def test_surface_farfield_matches_oracle(which, mirror):
    from efs_amd import nf2ff
    from efs_amd.engine import HIPEngine
    from oracle import nf2ff_surface as orc
    rng = np.random.default_rng({"box": 31, "cylinder": 32, "ring": 33}[which])
    faces = {"box": lambda: box_faces(rng), "cylinder": lambda: cylinder_faces(rng), "ring": lambda: [cyl_ring()]}[which]()
    theta = np.linspace(0.0, np.pi, 37)
    phi = np.linspace(0.0, 2 * np.pi, 25)
    center, freq, radius = (0.001, -0.002, 0.003), 3.1e9, 1.0
    opts = {"mirror": mirror} if mirror else {}
    print("case", which, mirror, flush=True)
    g = HIPEngine()
    assert g.initGPU()
    got = nf2ff.compute_nf2ff({"faces": faces}, freq, theta, phi, center, radius, opts, engine=g)
    g.destroy()
    want = orc.compute_nf2ff({"faces": faces}, freq, theta, phi, center, radius, opts)
    for a, b in (("E_theta_re", "E_theta_im"), ("E_phi_re", "E_phi_im")):
        e = rel(got[a] + 1j * got[b], want[a] + 1j * want[b])
        print(a[:7], f"{e:.2e}", flush=True)
        assert e < 1e-5
    assert abs(got["Prad"] - want["Prad"]) <= 1e-12 * abs(want["Prad"])
    if want["Dmax"] > 0:
        assert abs(got["Dmax"] - want["Dmax"]) <= 1e-5 * want["Dmax"]
    assert rel(got["P_rad"], want["P_rad"]) < 2e-5


def test_surface_farfield_without_init_and_bad_input():
    from efs_amd.engine import HIPEngine
    g = HIPEngine()
    assert g.initGPU()
    meta = np.zeros(8, np.float32)
    meta[3] = 5  # not an axis
    with pytest.raises(RuntimeError):
        g.farfieldSurface(meta, np.zeros(6), np.zeros(6), [0.1], [0.2], [0, 0, 0], 1e9)
    meta[3] = 2
    meta[4] = 1
    meta[5] = 1e-6
    ah = np.zeros(6)
    ah[0] = -1.0  # z x (-x) = -y current element
    out = g.farfieldSurface(meta, np.zeros(6), ah, [np.pi / 2], [0.0], [0, 0, 0], 1e9)
    assert out.shape == (5,) and out[0] > 0
    g.destroy()


def test_js_compute_nf2ff_matches_python(tmp_path):
    """efs_amd/js/nf2ff.mjs (the JS drop-in for src/nf2ff.mjs's
    computeNF2FF) against the Python path on the same cylindrical surface
    with a PMC mirror: the same points reach the same kernel."""
    import json
    import os
    import shutil
    import subprocess
    from efs_amd import nf2ff
    node = shutil.which("node")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if node is None or not os.path.exists(os.path.join(root, "efs_amd", "js", "efs_engine.node")):
        pytest.skip("node or the addon is missing")
    rng = np.random.default_rng(41)
    faces = cylinder_faces(rng)
    theta, phi = np.linspace(0.0, np.pi, 19), np.linspace(0.0, 2 * np.pi, 13)
    opts = {"mirror": {"type": "PMC", "direction": 0, "position": 0.03}}
    case = {"faces": [{"E": [a.tolist() for a in f["E"]], "H": [a.tolist() for a in f["H"]],
                       "mesh": {k: np.asarray(v).tolist() for k, v in f["mesh"].items()}, "normal": f["normal"],
                       "meshType": f.get("meshType", 0)} for f in faces],
            "freq": 2.2e9, "theta": theta.tolist(), "phi": phi.tolist(), "center": [0.0, 0.001, 0.0],
            "radius": 1.0, "opts": opts}
    src, dst = tmp_path / "surf.json", tmp_path / "out.json"
    src.write_text(json.dumps(case))
    print("node computeNF2FF", flush=True)
    res = subprocess.run([node, os.path.join(root, "tests", "js", "run_nf2ff_surface.mjs"), str(src), str(dst)],
                         capture_output=True, text=True, timeout=120, cwd=root)
    assert res.returncode == 0, res.stderr
    got = json.loads(dst.read_text())
    want = nf2ff.compute_nf2ff({"faces": faces}, 2.2e9, theta, phi, (0.0, 0.001, 0.0), 1.0, opts)
    for k in ("E_theta_re", "E_theta_im", "E_phi_re", "E_phi_im", "P_rad"):
        assert rel(np.asarray(got[k]), want[k]) < 1e-6, k
    assert abs(got["Prad"] - want["Prad"]) <= 1e-12 * abs(want["Prad"])
    assert abs(got["Dmax"] - want["Dmax"]) <= 1e-6 * want["Dmax"]
