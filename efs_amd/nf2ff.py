"""computeNF2FF over surface fields, on the GPU.

The reference's CPU far field (src/nf2ff.mjs:102-436) takes the complex E
and H on the faces of an NF2FF box -- Cartesian faces, or faces of an
openEMS cylindrical mesh (rho, alpha, z) -- and returns E_theta, E_phi, the
radiation intensity per angle, the radiated power and Dmax.  Here the host
turns the faces into surface points and the radiation integrals run on the
GPU (efs_farfield_surface: FP32 phase and currents, FP64 sums).

Points.  Each face point gets the reference's position, area (midpoint
rule; cylindrical: alpha edges times rho) and fields.  A cylindrical point
carries its fields in (rho, alpha, z) components and the normal of its face
in that frame; the reference forms Js = n x H and Ms = -n x E there and
rotates them by the point's alpha (src/nf2ff.mjs:208-262).  A rotation
about z commutes with the cross product, so that equals n' x H' with the
rotated normal n' and fields H'.  The integrals are linear in area * n, so
a point whose rotated normal is n' = (nx, ny, 0) is given to the GPU as the
Cartesian points (axis 0, sign nx, area |nx|) and (axis 1, sign ny, area
|ny|) with the rotated fields.

Mirror (src/nf2ff.mjs:377-398): the reference adds every point again at the
reflected position with PEC/PMC sign factors applied to the components as
stored (for a cylindrical face: to rho, alpha, z) and the face normals kept.
The reflected points are built the same way and added to the point list;
the radiated power is that of the unreflected surface only.
"""
import numpy as np

from . import engine as _eng

C0 = 299792458.0


def _edge_lengths(lines):
    """Midpoint-rule edge lengths, 0 for a single line (computeEdgeLengths,
    src/nf2ff.mjs:462-475)."""
    x = np.asarray(lines, np.float64)
    n = len(x)
    out = np.zeros(n)
    if n > 1:
        out[0] = 0.5 * abs(x[1] - x[0])
        out[1:n - 1] = 0.5 * np.abs(x[2:] - x[:-2])
        out[n - 1] = 0.5 * abs(x[n - 1] - x[n - 2])
    return out


def _face_points(face, default_mesh_type, e_sign, h_sign, mirror_dir, mirror_pos):
    """One face as surface points: (pos[N,3], normal axis, normal sign,
    area, E[N,3] complex, H[N,3] complex), Cartesian, plus the reference's
    per-point Poynting term (components as stored) for the power."""
    E, H, mesh, normal = face["E"], face["H"], face["mesh"], face["normal"]
    mesh_type = face.get("meshType", default_mesh_type)
    ny = next((n for n in range(3) if normal[n] != 0), -1)
    if ny < 0:
        return None
    nP, nPP = (ny + 1) % 3, (ny + 2) % 3
    m = [np.asarray(mesh[a], np.float64) for a in ("x", "y", "z")]
    numP, numPP = len(m[nP]), len(m[nPP])
    sign = float(normal[ny])
    eP, ePP = _edge_lengths(m[nP]), _edge_lengths(m[nPP])
    cyl = mesh_type == 1
    if cyl and ny == 0:
        eP = eP * m[0][0]
    iP, iPP = np.meshgrid(np.arange(numP), np.arange(numPP), indexing="ij")
    iP, iPP = iP.ravel(), iPP.ravel()
    idx = iPP * numP + iP if ny == 1 else iP * numPP + iPP
    area = eP[iP] * ePP[iPP]
    if cyl and ny == 2:
        area = area * m[0][iP]
    Ec = np.stack([(np.asarray(E[c], np.float64)[2 * idx] + 1j * np.asarray(E[c], np.float64)[2 * idx + 1]) * e_sign[c]
                   for c in range(3)], axis=1)
    Hc = np.stack([(np.asarray(H[c], np.float64)[2 * idx] + 1j * np.asarray(H[c], np.float64)[2 * idx + 1]) * h_sign[c]
                   for c in range(3)], axis=1)
    poy = (Ec[:, nP].real * Hc[:, nPP].real + Ec[:, nP].imag * Hc[:, nPP].imag) - \
        (Ec[:, nPP].real * Hc[:, nP].real + Ec[:, nPP].imag * Hc[:, nP].imag)
    power = float(np.sum(0.5 * area * poy * sign))
    pos = np.zeros((len(idx), 3))
    pos[:, ny] = m[ny][0]
    pos[:, nP] = m[nP][iP]
    pos[:, nPP] = m[nPP][iPP]
    nvec = np.zeros((len(idx), 3))
    nvec[:, ny] = sign
    if cyl:
        a = pos[:, 1].copy()  # the point's alpha (src/nf2ff.mjs:222-233 picks the same line)
        ca, sa = np.cos(a), np.sin(a)

        def rot(v):
            return np.stack([v[:, 0] * ca - v[:, 1] * sa, v[:, 0] * sa + v[:, 1] * ca, v[:, 2]], axis=1)

        Ec, Hc, nvec = rot(Ec), rot(Hc), rot(nvec)
        pos = np.stack([pos[:, 0] * ca, pos[:, 0] * sa, pos[:, 2]], axis=1)
    if mirror_dir >= 0:
        pos[:, mirror_dir] = 2 * mirror_pos - pos[:, mirror_dir]
    return pos, nvec, area, Ec, Hc, power


def surface_points(surface_data, default_mesh_type=0, mirror=None):
    """The GPU point list of computeNF2FF's surface: point_meta (8 floats
    per point), accE / accH (6 doubles per point) and the radiated power of
    the unreflected surface."""
    passes = [([1, 1, 1], [1, 1, 1], -1, 0.0)]
    if mirror:
        n = int(mirror["direction"])
        a1, a2 = (n + 1) % 3, (n + 2) % 3
        ef, hf = [1, 1, 1], [1, 1, 1]
        t = str(mirror["type"]).upper()
        if t == "PEC":
            hf[n] = -1
            ef[a1] = ef[a2] = -1
        elif t == "PMC":
            ef[n] = -1
            hf[a1] = hf[a2] = -1
        passes.append((ef, hf, n, float(mirror["position"])))
    metas, accE, accH = [], [], []
    power = 0.0
    for k, (ef, hf, mdir, mpos) in enumerate(passes):
        for face in surface_data["faces"]:
            r = _face_points(face, default_mesh_type, ef, hf, mdir, mpos)
            if r is None:
                continue
            pos, nvec, area, Ec, Hc, pw = r
            if k == 0:
                power += pw
            # a rotated normal splits into its Cartesian components
            for ax in range(3):
                sel = nvec[:, ax] != 0.0
                if not np.any(sel):
                    continue
                cnt = int(np.count_nonzero(sel))
                meta = np.zeros((cnt, 8))
                meta[:, 0:3] = pos[sel]
                meta[:, 3] = ax
                meta[:, 4] = np.sign(nvec[sel, ax])
                meta[:, 5] = area[sel] * np.abs(nvec[sel, ax])
                metas.append(meta)
                accE.append(np.stack([Ec[sel].real, Ec[sel].imag], axis=2).reshape(cnt, 6))
                accH.append(np.stack([Hc[sel].real, Hc[sel].imag], axis=2).reshape(cnt, 6))
    if not metas:
        return np.zeros((0, 8), np.float32), np.zeros((0, 6)), np.zeros((0, 6)), power
    return (np.concatenate(metas).astype(np.float32), np.concatenate(accE), np.concatenate(accH), power)


def compute_nf2ff(surface_data, freq, theta, phi, center=(0.0, 0.0, 0.0), radius=1.0, opts=None, engine=None):
    """computeNF2FF(surfaceData, freq, theta, phi, center, radius, opts)
    (src/nf2ff.mjs:102-436) with the radiation integrals on the GPU.
    opts: meshType (0 Cartesian, 1 cylindrical: faces' default, and center
    given as (rho, alpha, z)), mirror {type: 'PEC'|'PMC', direction,
    position}.  engine: a HIPEngine whose device runs it (a fresh context
    when omitted).  Returns the reference's fields: E_theta_re/_im,
    E_phi_re/_im and P_rad per angle (theta-major), Prad, Dmax."""
    opts = opts or {}
    mesh_type = int(opts.get("meshType", 0))
    if mesh_type == 1:
        cxyz = [center[0] * np.cos(center[1]), center[0] * np.sin(center[1]), center[2]]
    else:
        cxyz = [float(c) for c in center]
    meta, accE, accH, power = surface_points(surface_data, mesh_type, opts.get("mirror"))
    th = np.asarray(theta, np.float64)
    ph = np.asarray(phi, np.float64)
    na = len(th) * len(ph)
    own = engine is None
    if own:
        engine = _eng.HIPEngine()
        if not engine.initGPU():
            raise RuntimeError("computeNF2FF: no HIP device")
    try:
        if len(meta):
            out = engine.farfieldSurface(meta, accE.ravel(), accH.ravel(), th, ph, cxyz, float(freq), float(radius))
        else:
            out = np.zeros(5 * na)
    finally:
        if own:
            engine.destroy()
    f = out.reshape(na, 5)
    p_rad = f[:, 0].copy()
    p_max = float(p_rad.max()) if na else 0.0
    dmax = p_max * 4 * np.pi * radius * radius / power if power > 0 else 0.0
    return {"E_theta_re": f[:, 1].copy(), "E_theta_im": f[:, 2].copy(), "E_phi_re": f[:, 3].copy(),
            "E_phi_im": f[:, 4].copy(), "Prad": power, "Dmax": dmax, "P_rad": p_rad}


computeNF2FF = compute_nf2ff
