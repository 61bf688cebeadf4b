"""computeNF2FF over surface fields (src/nf2ff.mjs:102-436): the oracle's
restatement (oracle/nf2ff_surface.py) and the host side of the GPU path
(efs_amd/nf2ff.py: faces -> Cartesian surface points).

The oracle is pinned by
- the reference's own far-field tests (tests/test_api.mjs): the
  infinitesimal dipole (1697-1780: Dmax from the angular integral within
  10 % of 1.5, E_theta null at theta = 90 deg, phi = 0), the cylindrical
  mesh (1909-1990: differs from the same numbers read as Cartesian), the
  PEC and PMC mirrors (2000-2097: each changes E_theta, and they differ);
- the textbook radiation integrals (Balanis 12-10/12-12) evaluated with
  explicit Cartesian normals, positions and fields for every face
  orientation of a cylindrical mesh and of a Cartesian box, with and
  without a mirror: this fixes the rotation of the currents, the alpha each
  point uses, the rho factors of the areas, the transposed index of y
  faces, the mirror signs and the power.

The host point construction of the GPU path is checked on the CPU: its
points, summed with the same textbook integrals, give the oracle's result.
"""
import numpy as np
import pytest

from efs_amd import nf2ff as host
from oracle import nf2ff_surface as orc

C0 = orc.C0
ETA = orc.Z0


def face(mesh, normal, rng=None, mesh_type=None, fill=None):
    n = len(mesh["x"]) * len(mesh["y"]) * len(mesh["z"])
    f = {"E": [np.zeros(2 * n) for _ in range(3)], "H": [np.zeros(2 * n) for _ in range(3)],
         "mesh": {k: np.asarray(v, float) for k, v in mesh.items()}, "normal": list(normal)}
    if rng is not None:
        for c in range(3):
            f["E"][c] = rng.normal(size=2 * n)
            f["H"][c] = rng.normal(size=2 * n) / ETA
    if fill:
        fill(f)
    if mesh_type is not None:
        f["meshType"] = mesh_type
    return f


def textbook(points, theta, phi, center, freq, radius):
    """E_theta, E_phi per angle from points (pos[3] Cartesian, n[3]
    Cartesian unit normal times sign, area, E[3], H[3] complex)."""
    k = 2 * np.pi * freq / C0
    et, ep = [], []
    for th in theta:
        for ph in phi:
            rhat = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
            that = np.array([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
            phat = np.array([-np.sin(ph), np.cos(ph), 0.0])
            Nv = np.zeros(3, complex)
            Lv = np.zeros(3, complex)
            for pos, n, area, E, H in points:
                w = area * np.exp(1j * k * rhat @ (pos - center))
                Nv += np.cross(n, H) * w
                Lv += -np.cross(n, E) * w
            pre = 1j * k * np.exp(-1j * k * radius) / (4 * np.pi * radius)
            et.append(-pre * (Lv @ phat + ETA * (Nv @ that)))
            ep.append(pre * (Lv @ that - ETA * (Nv @ phat)))
    return np.array(et), np.array(ep)


def physical_points(faces, mirror=None, default_mesh_type=0):
    """The surface as physical points, written from the geometry: a
    cylindrical patch's area is the product of its physical edge lengths
    (rho * d_alpha along alpha), its normal and fields are rotated by its
    alpha into Cartesian."""
    out = []
    passes = [(np.ones(3), np.ones(3), -1, 0.0)]
    if mirror:
        n = mirror["direction"]
        ef, hf = np.ones(3), np.ones(3)
        t = [(n + 1) % 3, (n + 2) % 3]
        if mirror["type"] == "PEC":
            hf[n] = -1
            ef[t] = -1
        else:
            ef[n] = -1
            hf[t] = -1
        passes.append((ef, hf, n, mirror["position"]))
    for ef, hf, mdir, mpos in passes:
        for f in faces:
            cyl = f.get("meshType", default_mesh_type) == 1
            m = [f["mesh"][a] for a in ("x", "y", "z")]
            ny = [i for i in range(3) if f["normal"][i]][0]
            ax = [(ny + 1) % 3, (ny + 2) % 3]
            L = [orc.edge_lengths(m[a]) for a in range(3)]
            for i0 in range(len(m[0])):
                for i1 in range(len(m[1])):
                    for i2 in range(len(m[2])):
                        ii = [i0, i1, i2]
                        idx = (i0 * len(m[1]) + i1) * len(m[2]) + i2  # (x, y, z) order of the dump
                        lens = [L[a][ii[a]] for a in ax]
                        if cyl:
                            rho = m[0][i0]
                            lens = [lens[j] * (rho if ax[j] == 1 else 1.0) for j in range(2)]
                        area = lens[0] * lens[1]
                        E = np.array([f["E"][c][2 * idx] + 1j * f["E"][c][2 * idx + 1] for c in range(3)]) * ef
                        H = np.array([f["H"][c][2 * idx] + 1j * f["H"][c][2 * idx + 1] for c in range(3)]) * hf
                        n = np.zeros(3)
                        n[ny] = f["normal"][ny]
                        pos = np.array([m[0][i0], m[1][i1], m[2][i2]])
                        if cyl:
                            a = m[1][i1]
                            R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
                            E, H, n = R @ E, R @ H, R @ n
                            pos = np.array([pos[0] * np.cos(a), pos[0] * np.sin(a), pos[2]])
                        if mdir >= 0:
                            pos[mdir] = 2 * mpos - pos[mdir]
                        out.append((pos, n, area, E, H))
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


ANG = (np.linspace(0.05, 3.1, 7), np.linspace(0.0, 6.2, 6))


# ---- the reference's own tests --------------------------------------------

def test_reference_infinitesimal_dipole():
    """tests/test_api.mjs:1697-1780."""
    f = face({"x": [-0.001, 0, 0.001], "y": [-0.001, 0, 0.001], "z": [0.0]}, [0, 0, 1],
             fill=lambda f: f["H"][1].__setitem__(slice(0, None, 2), -1.0))
    nT, nP = 91, 73
    theta = np.arange(nT) * np.pi / (nT - 1)
    phi = np.arange(nP) * 2 * np.pi / (nP - 1)
    r = orc.compute_nf2ff({"faces": [f]}, 1e9, theta, phi)
    P = r["P_rad"]
    assert P.max() > 0
    assert P[(nT // 2) * nP] / P[0] < 0.05
    prad = (P.reshape(nT, nP) * np.sin(theta)[:, None]).sum() * (np.pi / (nT - 1)) * (2 * np.pi / (nP - 1))
    assert abs(4 * np.pi * P.max() / prad - 1.5) / 1.5 < 0.10


def cyl_ring():
    """tests/test_api.mjs:1909-1944: a rho face at 10 mm, 8 alphas, 3 z
    lines, H_z = 1."""
    alpha = np.arange(8) * 2 * np.pi / 8
    return face({"x": [0.01], "y": alpha, "z": [-0.001, 0, 0.001]}, [1, 0, 0], mesh_type=1,
                fill=lambda f: f["H"][2].__setitem__(slice(0, None, 2), 1.0))


def test_reference_cylindrical_differs_from_cartesian():
    """tests/test_api.mjs:1909-1990."""
    f = cyl_ring()
    theta, phi = np.array([np.pi / 4]), np.array([0, np.pi / 2])
    rc = orc.compute_nf2ff({"faces": [f]}, 1e9, theta, phi)
    g = dict(f)
    g.pop("meshType")
    rk = orc.compute_nf2ff({"faces": [g]}, 1e9, theta, phi)
    assert rc["P_rad"].max() > 0 and rk["P_rad"].max() > 0
    d = abs(rc["E_theta_re"][0] - rk["E_theta_re"][0]) + abs(rc["E_phi_re"][0] - rk["E_phi_re"][0])
    m = max(abs(rc["E_theta_re"][0]), abs(rk["E_theta_re"][0]), abs(rc["E_phi_re"][0]), abs(rk["E_phi_re"][0]))
    assert d / m > 0.01


def test_reference_mirrors_change_the_field():
    """tests/test_api.mjs:2000-2097."""
    f = face({"x": [-0.001, 0, 0.001], "y": [-0.001, 0, 0.001], "z": [0.01]}, [0, 0, 1],
             fill=lambda f: f["H"][1].__setitem__(slice(0, None, 2), -1.0))
    theta, phi = np.array([np.pi / 4]), np.array([0.0])
    et = {}
    for t in (None, "PEC", "PMC"):
        opts = {"mirror": {"type": t, "direction": 2, "position": 0}} if t else {}
        r = orc.compute_nf2ff({"faces": [f]}, 1e9, theta, phi, (0, 0, 0), 1, opts)
        et[t] = np.hypot(r["E_theta_re"][0], r["E_theta_im"][0])
    assert et[None] > 0 and et["PEC"] > 0 and et["PMC"] > 0
    assert abs(et["PEC"] - et[None]) > 1e-20 and abs(et["PMC"] - et[None]) > 1e-20
    assert abs(et["PMC"] - et["PEC"]) > 1e-20


# ---- the oracle against the textbook integrals ------------------------------

def cylinder_faces(rng):
    """A closed-ish cylindrical surface: an outer rho face, two alpha faces
    (a 200-degree sector) and two z faces, random fields."""
    rho = [0.004, 0.007, 0.01]
    alpha = np.deg2rad([10, 60, 110, 160, 210])
    z = [-0.003, 0.0, 0.002, 0.004]
    return [face({"x": [0.01], "y": alpha, "z": z}, [1, 0, 0], rng, mesh_type=1),
            face({"x": rho, "y": [alpha[0]], "z": z}, [0, -1, 0], rng, mesh_type=1),
            face({"x": rho, "y": [alpha[-1]], "z": z}, [0, 1, 0], rng, mesh_type=1),
            face({"x": rho, "y": alpha, "z": [z[0]]}, [0, 0, -1], rng, mesh_type=1),
            face({"x": rho, "y": alpha, "z": [z[-1]]}, [0, 0, 1], rng, mesh_type=1)]


def box_faces(rng):
    xs, ys, zs = [-0.01, -0.002, 0.005, 0.011], [-0.008, 0.0, 0.009], [-0.006, 0.001, 0.004, 0.01, 0.012]
    out = []
    for ny, lines in enumerate((xs, ys, zs)):
        for side, v in ((-1, lines[0]), (1, lines[-1])):
            mesh = {"x": xs, "y": ys, "z": zs}
            mesh["xyz"[ny]] = [v]
            nrm = [0, 0, 0]
            nrm[ny] = side
            out.append(face(mesh, nrm, rng))
    return out


@pytest.mark.parametrize("which", ["box", "cylinder"])
@pytest.mark.parametrize("mirror", [None, {"type": "PEC", "direction": 2, "position": -0.02},
                                    {"type": "PMC", "direction": 0, "position": 0.03}])
def test_oracle_matches_textbook(which, mirror):
    rng = np.random.default_rng(11 if which == "box" else 12)
    faces = box_faces(rng) if which == "box" else cylinder_faces(rng)
    theta, phi = ANG
    center, freq, radius = np.array([0.001, -0.002, 0.003]), 3.1e9, 2.0
    got = orc.compute_nf2ff({"faces": faces}, freq, theta, phi, center, radius,
                            {"mirror": mirror} if mirror else {})
    et, ep = textbook(physical_points(faces, mirror), theta, phi, center, freq, radius)
    assert rel(got["E_theta_re"] + 1j * got["E_theta_im"], et) < 1e-12
    assert rel(got["E_phi_re"] + 1j * got["E_phi_im"], ep) < 1e-12
    # power: the Poynting flux of the unreflected surface, from the physical points
    pw = sum(0.5 * area * np.real(np.cross(E, np.conj(H)) @ n) for pos, n, area, E, H in physical_points(faces))
    assert abs(got["Prad"] - pw) <= 1e-12 * abs(pw)


def test_oracle_cylindrical_center():
    """meshType 1 in the options: the phase center is given as (rho, alpha, z)."""
    rng = np.random.default_rng(5)
    faces = cylinder_faces(rng)
    theta, phi = ANG
    c_cyl = (0.003, 0.7, -0.001)
    c_xyz = np.array([0.003 * np.cos(0.7), 0.003 * np.sin(0.7), -0.001])
    a = orc.compute_nf2ff({"faces": faces}, 2e9, theta, phi, c_cyl, 1.0, {"meshType": 1})
    et, ep = textbook(physical_points(faces, None, 1), theta, phi, c_xyz, 2e9, 1.0)
    assert rel(a["E_theta_re"] + 1j * a["E_theta_im"], et) < 1e-12


# ---- the host side of the GPU path -------------------------------------------

def sum_points(meta, accE, accH, theta, phi, center, freq, radius):
    pts = []
    for p in range(len(meta)):
        n = np.zeros(3)
        n[int(meta[p, 3])] = meta[p, 4]
        E = accE[p, 0::2] + 1j * accE[p, 1::2]
        H = accH[p, 0::2] + 1j * accH[p, 1::2]
        pts.append((meta[p, :3].astype(float), n, float(meta[p, 5]), E, H))
    return textbook(pts, theta, phi, np.asarray(center, float), freq, radius)


@pytest.mark.parametrize("which", ["box", "cylinder"])
@pytest.mark.parametrize("mirror", [None, {"type": "PEC", "direction": 1, "position": 0.02}])
def test_host_points_give_the_oracle(which, mirror):
    rng = np.random.default_rng(21)
    faces = box_faces(rng) if which == "box" else cylinder_faces(rng)
    theta, phi = ANG
    center, freq, radius = [0.002, 0.001, -0.001], 2.7e9, 1.5
    meta, accE, accH, power = host.surface_points({"faces": faces}, 0, mirror)
    assert meta.shape[1] == 8 and accE.shape == (len(meta), 6) and accH.shape == accE.shape
    assert set(np.unique(meta[:, 3])) <= {0.0, 1.0, 2.0} and np.all(np.abs(meta[:, 4]) == 1)
    want = orc.compute_nf2ff({"faces": faces}, freq, theta, phi, center, radius,
                             {"mirror": mirror} if mirror else {})
    assert abs(power - want["Prad"]) <= 1e-12 * abs(want["Prad"])
    # positions and areas go to the GPU as FP32, fields as FP64
    et, ep = sum_points(meta, accE, accH, theta, phi, center, freq, radius)
    assert rel(et, want["E_theta_re"] + 1j * want["E_theta_im"]) < 1e-5
    assert rel(ep, want["E_phi_re"] + 1j * want["E_phi_im"]) < 1e-5


def test_host_rejects_nothing_but_empty_normals():
    f = face({"x": [0.0, 0.01], "y": [0.0], "z": [0.0, 0.01]}, [0, 0, 0])
    meta, accE, accH, power = host.surface_points({"faces": [f]})
    assert len(meta) == 0 and power == 0.0


def test_textbook_check_catches_a_planted_mistake():
    """The oracle with the sign of its cylindrical rotation flipped must
    fail the textbook comparison (the check has teeth)."""
    import os
    import types
    path = os.path.join(os.path.dirname(orc.__file__), "nf2ff_surface.py")
    src = open(path).read()
    good = "js = [js[0] * ca - js[1] * sa"
    assert good in src
    planted = types.ModuleType("nf2ff_surface_planted")
    exec(compile(src.replace(good, "js = [js[0] * ca + js[1] * sa"), path, "exec"), planted.__dict__)
    rng = np.random.default_rng(12)
    faces = cylinder_faces(rng)
    theta, phi = ANG
    center, freq, radius = np.array([0.001, -0.002, 0.003]), 3.1e9, 2.0
    got = planted.compute_nf2ff({"faces": faces}, freq, theta, phi, center, radius)
    et, _ = textbook(physical_points(faces), theta, phi, center, freq, radius)
    assert rel(got["E_theta_re"] + 1j * got["E_theta_im"], et) > 1e-3


def test_js_points_match_python(tmp_path):
    """The JS drop-in (efs_amd/js/nf2ff.mjs) builds the same surface points
    as the Python path (CPU only: no far field is run)."""
    import json
    import os
    import shutil
    import subprocess
    node = shutil.which("node")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if node is None or not os.path.exists(os.path.join(root, "efs_amd", "js", "efs_engine.node")):
        pytest.skip("node or the addon is missing")
    rng = np.random.default_rng(51)
    faces = cylinder_faces(rng) + box_faces(rng)
    mirror = {"type": "PEC", "direction": 1, "position": -0.01}
    case = {"faces": [{"E": [a.tolist() for a in f["E"]], "H": [a.tolist() for a in f["H"]],
                       "mesh": {k: np.asarray(v).tolist() for k, v in f["mesh"].items()}, "normal": f["normal"],
                       "meshType": f.get("meshType", 0)} for f in faces],
            "opts": {"mirror": mirror}}
    src, dst = tmp_path / "surf.json", tmp_path / "pts.json"
    src.write_text(json.dumps(case))
    res = subprocess.run([node, os.path.join(root, "tests", "js", "run_nf2ff_surface.mjs"), str(src), str(dst),
                          "points"], capture_output=True, text=True, timeout=120, cwd=root)
    assert res.returncode == 0, res.stderr
    got = json.loads(dst.read_text())
    meta, accE, accH, power = host.surface_points({"faces": faces}, 0, mirror)
    np.testing.assert_array_equal(np.asarray(got["meta"], np.float32).reshape(-1, 8), meta)
    np.testing.assert_allclose(np.asarray(got["accE"]).reshape(-1, 6), accE, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(np.asarray(got["accH"]).reshape(-1, 6), accH, rtol=1e-13, atol=1e-300)
    assert abs(got["power"] - power) <= 1e-13 * abs(power)
