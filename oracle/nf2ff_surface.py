"""FP64 restatement of the reference's CPU far field over surface fields,
computeNF2FF (src/nf2ff.mjs:102-436), for the tests.

TEST INFRASTRUCTURE ONLY: the product path (efs_amd.nf2ff, which runs the
radiation integrals on the GPU through efs_farfield_surface) never imports
this module; the tests use it as the checker.

Follows the reference step for step, per face and per surface point:
  - edge lengths by the midpoint rule, 0 for a single line
    (computeEdgeLengths, src/nf2ff.mjs:462-475);
  - cylindrical meshes (rho, alpha, z): alpha edges of a rho face scaled by
    rho, areas of a z face by the point's rho (src/nf2ff.mjs:165-195);
  - the transposed point index of y faces (src/nf2ff.mjs:184-187);
  - mirror signs applied to the components as stored (src/nf2ff.mjs:
    200-206, 381-397);
  - Js = n x H, Ms = -n x E in the face's frame, then for cylindrical
    meshes rotated by the point's alpha to Cartesian (src/nf2ff.mjs:208-
    262); positions converted to Cartesian and reflected for the mirror
    pass (src/nf2ff.mjs:264-281);
  - the Poynting power of the main pass only (src/nf2ff.mjs:283-297);
  - the radiation integrals and E_theta, E_phi, P_rad, Dmax (src/nf2ff.mjs:
    299-436), vectorised over the angles here.

Pinned by the reference's own far-field tests (tests/test_api.mjs:1697-
1780, 1909-2097) and, independently, by the textbook radiation integrals
with explicit Cartesian normals (tests/test_nf2ff_surface.py).
"""
import numpy as np

C0 = 299792458.0
Z0 = 376.73031346177066


def edge_lengths(lines):
    """computeEdgeLengths (src/nf2ff.mjs:462-475)."""
    x = np.asarray(lines, np.float64)
    n = len(x)
    out = np.zeros(n)
    if n == 1:
        return out
    out[0] = 0.5 * abs(x[1] - x[0])
    out[1:n - 1] = 0.5 * np.abs(x[2:] - x[:-2])
    out[n - 1] = 0.5 * abs(x[n - 1] - x[n - 2])
    return out


def _accumulate(faces, center_xyz, e_sign, h_sign, mirror_dir, mirror_pos, default_mesh_type, k, theta, phi, acc):
    """accumulateFaces (src/nf2ff.mjs:137-352); returns the radiated power
    of this pass (mirror passes contribute none)."""
    th = np.asarray(theta, np.float64)[:, None]
    ph = np.asarray(phi, np.float64)[None, :]
    sinT, cosT, sinP, cosP = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    rad_power = 0.0
    for face in faces:
        E, H, mesh, normal = face["E"], face["H"], face["mesh"], face["normal"]
        mesh_type = face.get("meshType", default_mesh_type)
        ny = next((n for n in range(3) if normal[n] != 0), -1)
        if ny < 0:
            continue
        nP, nPP = (ny + 1) % 3, (ny + 2) % 3
        m = [np.asarray(mesh[a], np.float64) for a in ("x", "y", "z")]
        numP, numPP = len(m[nP]), len(m[nPP])
        norm_sign = float(normal[ny])
        eP, ePP = edge_lengths(m[nP]), edge_lengths(m[nPP])
        if mesh_type == 1 and ny == 0:
            eP = eP * m[0][0]
        for iP in range(numP):
            for iPP in range(numPP):
                idx = iPP * numP + iP if ny == 1 else iP * numPP + iPP
                area = eP[iP] * ePP[iPP]
                if mesh_type == 1 and ny == 2:
                    area *= m[0][iP]
                e = [complex(E[c][2 * idx], E[c][2 * idx + 1]) * e_sign[c] for c in range(3)]
                h = [complex(H[c][2 * idx], H[c][2 * idx + 1]) * h_sign[c] for c in range(3)]
                nd = [0.0, 0.0, 0.0]
                nd[ny] = norm_sign
                js = [nd[1] * h[2] - nd[2] * h[1], nd[2] * h[0] - nd[0] * h[2], nd[0] * h[1] - nd[1] * h[0]]
                ms = [nd[2] * e[1] - nd[1] * e[2], nd[0] * e[2] - nd[2] * e[0], nd[1] * e[0] - nd[0] * e[1]]
                if mesh_type == 1:
                    pos_alpha = 0 if ny == 1 else (iP if nP == 1 else iPP)
                    a = m[1][pos_alpha]
                    ca, sa = np.cos(a), np.sin(a)
                    js = [js[0] * ca - js[1] * sa, js[0] * sa + js[1] * ca, js[2]]
                    ms = [ms[0] * ca - ms[1] * sa, ms[0] * sa + ms[1] * ca, ms[2]]
                pos = [0.0, 0.0, 0.0]
                pos[ny] = m[ny][0]
                pos[nP] = m[nP][iP]
                pos[nPP] = m[nPP][iPP]
                if mesh_type == 1:
                    pos = [pos[0] * np.cos(pos[1]), pos[0] * np.sin(pos[1]), pos[2]]
                if mirror_dir >= 0:
                    pos[mirror_dir] = 2 * mirror_pos - pos[mirror_dir]
                if mirror_dir < 0:
                    s = (e[nP].real * h[nPP].real + e[nP].imag * h[nPP].imag) - \
                        (e[nPP].real * h[nP].real + e[nPP].imag * h[nP].imag)
                    rad_power += 0.5 * area * s * norm_sign
                r_cos_psi = (pos[0] - center_xyz[0]) * cosP * sinT + (pos[1] - center_xyz[1]) * sinT * sinP + \
                    (pos[2] - center_xyz[2]) * cosT
                ae = area * np.exp(1j * k * r_cos_psi)
                jt = js[0] * cosT * cosP + js[1] * cosT * sinP - js[2] * sinT
                jp = js[1] * cosP - js[0] * sinP
                mt = ms[0] * cosT * cosP + ms[1] * cosT * sinP - ms[2] * sinT
                mp = ms[1] * cosP - ms[0] * sinP
                acc["Nt"] += ae * jt
                acc["Np"] += ae * jp
                acc["Lt"] += ae * mt
                acc["Lp"] += ae * mp
    return rad_power


def compute_nf2ff(surface_data, freq, theta, phi, center=(0.0, 0.0, 0.0), radius=1.0, opts=None):
    """computeNF2FF (src/nf2ff.mjs:102-436): returns E_theta_re/_im,
    E_phi_re/_im and P_rad per angle (theta-major), Prad and Dmax."""
    opts = opts or {}
    nT, nPh = len(theta), len(phi)
    k = 2 * np.pi * freq / C0
    acc = {key: np.zeros((nT, nPh), np.complex128) for key in ("Nt", "Np", "Lt", "Lp")}
    default_mesh_type = opts.get("meshType", 0)
    if default_mesh_type == 1:
        center_xyz = [center[0] * np.cos(center[1]), center[0] * np.sin(center[1]), center[2]]
    else:
        center_xyz = list(center)
    faces = surface_data["faces"]
    rad_power = _accumulate(faces, center_xyz, [1, 1, 1], [1, 1, 1], -1, 0.0, default_mesh_type, k, theta, phi, acc)
    mir = opts.get("mirror")
    if mir:
        n = mir["direction"]
        a1, a2 = (n + 1) % 3, (n + 2) % 3
        ef, hf = [1, 1, 1], [1, 1, 1]
        if mir["type"] == "PEC":
            hf[n] = -1
            ef[a1] = ef[a2] = -1
        elif mir["type"] == "PMC":
            ef[n] = -1
            hf[a1] = hf[a2] = -1
        _accumulate(faces, center_xyz, ef, hf, n, mir["position"], default_mesh_type, k, theta, phi, acc)
    fac = 1j * (k / (4 * np.pi * radius)) * np.exp(-1j * k * radius)
    et = -fac * (acc["Lp"] + Z0 * acc["Nt"])
    ep = fac * (acc["Lt"] - Z0 * acc["Np"])
    p_rad = (np.abs(et) ** 2 + np.abs(ep) ** 2) / (2 * Z0)
    p_max = float(p_rad.max()) if p_rad.size else 0.0
    dmax = p_max * 4 * np.pi * radius * radius / rad_power if rad_power > 0 else 0.0
    return {"E_theta_re": et.real.ravel(), "E_theta_im": et.imag.ravel(), "E_phi_re": ep.real.ravel(),
            "E_phi_im": ep.imag.ravel(), "Prad": rad_power, "Dmax": dmax, "P_rad": p_rad.ravel()}
