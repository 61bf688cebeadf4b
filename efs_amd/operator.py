"""FP64 operator builder: EC-FDTD coefficients, UPML regions, Mur lists,
excitation signals and NF2FF surfaces on a rectilinear mesh.

This plays the role of the reference's upstream boundary -- the openEMS
operator whose arrays the embind getters hand to the engine
(getVV/VI/II/IV, getPMLRegion, getMurRegion, getExcitation*, getSimDT;
src/embind_api.cpp:230-681, 905-908).  openEMS itself is not available, so
the formulas below are written from the equivalent-circuit FDTD model and
Gedney's UPML, in the array layouts those getters produce:

  E edge (n; i,j,k): C = eps0*<eps_r>*A/dl, G = <kappa>*A/dl with A the
      dual face and <.> the area-weighted mean of the (up to 4) touching
      cells; vv = (1-g)/(1+g), vi = (dT/C)/(1+g), g = dT*G/(2C).
  H edge (n; dual, length dl' crossing the primal face A'): the flux
      through A' is continuous along the dual edge, so the edge inductance
      is L = mu0*A' / sum_i(w_i/mu_r,i) over the (up to 2) cells the edge
      crosses (w_i its length in cell i); ii = 1, iv = dT/L.
  PEC edges and edges leaving the grid: vv = vi = 0.
  UPML (component n, nP=(n+1)%3, nPP=(n+2)%3, s_m = sigma_m*dT/(2*eps0)):
      flux   F' = vv_c F + vi_c curl,  vv_c = (1-s_nP)/(1+s_nP), vi_c = vi/(1+s_nP)
      field  V' = vv V - vvfo F + vvfn F'
             vv = (1-s_nPP)/(1+s_nPP), vvfo = (1-s_n)/(1+s_nPP), vvfn = (1+s_n)/(1+s_nPP)
      (same for H with the matched magnetic loss), graded geometrically,
      sigma(d) = -ln(R0) ln(g) / (2 dl Z0 (g^(W/dl) - 1)) * g^(d/dl),
      R0 = 1e-6, g = 2.5.
All arithmetic is FP64; arrays are returned as float32 (the engine format).
"""
import numpy as np

C0 = 299792458.0
MUE0 = 1.2566370614359173e-06
EPS0 = 8.854187817620389e-12
Z0 = 376.73031346177066


class Mesh:
    """Rectilinear mesh: three arrays of line coordinates in metres."""

    def __init__(self, x, y, z):
        self.lines = [np.asarray(a, dtype=np.float64) for a in (x, y, z)]
        self.n = [len(a) for a in self.lines]
        self.primal = []  # edge length from node i to i+1 (0 for the last node)
        self.half_lo = []
        self.half_hi = []
        for a in self.lines:
            d = np.zeros(len(a))
            d[:-1] = np.diff(a)
            self.primal.append(d)
            lo = np.zeros(len(a))
            lo[1:] = d[:-1] / 2
            self.half_lo.append(lo)
            self.half_hi.append(d / 2)
        self.dual = [lo + hi for lo, hi in zip(self.half_lo, self.half_hi)]

    @property
    def shape(self):
        return tuple(self.n)

    def timestep(self, factor=1.0):
        """CFL limit of the smallest cell, times ``factor``."""
        m = [np.min(d[:-1]) if len(d) > 1 else np.inf for d in self.primal]
        s = sum(1.0 / v ** 2 for v in m if np.isfinite(v))
        return factor / (C0 * np.sqrt(s))


def uniform_mesh(n, delta):
    return Mesh(*(np.arange(k) * delta for k in n))


def _bcast(v, axis):
    shape = [1, 1, 1]
    shape[axis] = -1
    return v.reshape(shape)


class Materials:
    """Per-cell eps_r, kappa (S/m), mu_r on the (nx-1)(ny-1)(nz-1) cells plus
    a PEC mask on E edges."""

    def __init__(self, mesh):
        nc = [max(k - 1, 1) for k in mesh.n]
        self.mesh = mesh
        self.eps = np.ones(nc)
        self.kappa = np.zeros(nc)
        self.mu = np.ones(nc)
        self.pec = np.zeros((3,) + mesh.shape, dtype=bool)
        self.port_g = np.zeros((3,) + mesh.shape)  # lumped conductance folded into edges

    def box(self, lo, hi, eps=None, kappa=None):
        """Dielectric box over cell index ranges [lo, hi)."""
        s = tuple(slice(a, b) for a, b in zip(lo, hi))
        if eps is not None:
            self.eps[s] = eps
        if kappa is not None:
            self.kappa[s] = kappa

    def pec_box(self, lo, hi):
        """PEC over node index ranges [lo, hi] (inclusive): every E edge whose
        two end nodes lie in the closed box, so zero-thickness plates work."""
        for n in range(3):
            s = [slice(lo[m], hi[m] + 1) for m in range(3)]
            s[n] = slice(lo[n], hi[n])  # edges from node lo..hi-1 along n
            if s[n].stop > s[n].start:
                self.pec[(n,) + tuple(s)] = True

    def lumped_resistor(self, n, start, length, R):
        """Resistor of `length` edges in series along axis n from node `start`."""
        x, y, z = start
        for k in range(length):
            p = [x, y, z]
            p[n] += k
            self.port_g[n, p[0], p[1], p[2]] += length / R


def lumped_port(mesh, mat, n, start, stop, R, excite=1.0):
    """Lumped port along axis n between nodes start and stop (node index
    triples equal off axis n, or spanning a transverse rectangle), with the
    conventions of the reference's LumpedPort (src/ports.mjs:220-330):
    the resistor R is folded into the port edges (R > 0), the soft source
    drives E = -dir*excite (V/m) on them, i.e. a voltage amplitude of
    -dir*excite*dl per edge, the voltage probe runs along the port at the
    transverse centre with weight -1, and the current probe is the loop on
    the dual lines around the port at its middle (normal n, weight dir; a
    middle that falls between two dual lines takes the lower one).

    Returns the excitation point lists (pos, dir, amp) and the two probe
    records (the gather format of efs_amd.ports)."""
    start, stop = [int(v) for v in start], [int(v) for v in stop]
    d = 1 if stop[n] >= start[n] else -1
    lo, hi = min(start[n], stop[n]), max(start[n], stop[n])
    if hi == lo:
        raise ValueError("lumped port needs a non-zero extent along its axis")
    a, b = (n + 1) % 3, (n + 2) % 3
    nx, ny, nz = mesh.shape
    rng = [range(min(start[m], stop[m]), max(start[m], stop[m]) + 1) for m in range(3)]
    rng[n] = range(lo, hi)
    pos, amp = [], []
    n_par = len(rng[a]) * len(rng[b])
    for i in rng[0]:
        for j in rng[1]:
            for k in rng[2]:
                p = (i, j, k)
                if R > 0:
                    # strings of hi-lo edges in series (each edge's share of
                    # R proportional to its length), n_par strings in parallel
                    L = mesh.lines[n][hi] - mesh.lines[n][lo]
                    mat.port_g[n, i, j, k] += L / (R * n_par * mesh.primal[n][p[n]])
                pos.append(i * ny * nz + j * nz + k)
                amp.append(-d * excite * mesh.primal[n][p[n]])
    mid = 0.5 * (mesh.lines[n][start[n]] + mesh.lines[n][stop[n]])
    duals = 0.5 * (mesh.lines[n][:-1] + mesh.lines[n][1:])
    kd = int(np.argmin(np.abs(duals - mid)))  # first minimum: ties to the lower line
    centre = [(start[m] + stop[m]) // 2 for m in range(3)]
    us, ue = list(centre), list(centre)
    us[n], ue[n] = start[n], stop[n]
    s, t = [0, 0, 0], [0, 0, 0]
    for m in (a, b):
        s[m], t[m] = min(start[m], stop[m]) - 1, max(start[m], stop[m])
    s[n] = t[n] = kd

    def inside(v, m):
        return bool(0 <= v <= mesh.n[m] - 2)
    probes = [
        {"name": "ut", "type": 0, "start": us, "stop": ue, "weight": -1.0},
        {"name": "it", "type": 1, "start": s, "stop": t, "normDir": n, "weight": float(d),
         "startInside": [inside(s[m], m) for m in range(3)], "stopInside": [inside(t[m], m) for m in range(3)]},
    ]
    return {"pos": np.array(pos, np.uint32), "dir": np.full(len(pos), n, np.uint32),
            "amp": np.array(amp, np.float32)}, probes


def _pad_cells(a):
    return np.pad(a, 1)


def build_coefficients(mesh, mat, dT):
    """vv, vi, ii, iv (3*N, NIJK) in FP64."""
    nx, ny, nz = mesh.shape
    vv = np.zeros((3, nx, ny, nz))
    vi = np.zeros((3, nx, ny, nz))
    ii = np.zeros((3, nx, ny, nz))
    iv = np.zeros((3, nx, ny, nz))
    epsP, kapP, muP = _pad_cells(mat.eps), _pad_cells(mat.kappa), _pad_cells(mat.mu)
    N = (nx, ny, nz)
    for n in range(3):
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        # --- E edges: area-weighted mean over the 4 cells around the edge
        eA = np.zeros(N)
        kA = np.zeros(N)
        for sP, wP in ((-1, mesh.half_lo[nP]), (0, mesh.half_hi[nP])):
            for sPP, wPP in ((-1, mesh.half_lo[nPP]), (0, mesh.half_hi[nPP])):
                sl = [None, None, None]
                sl[n] = slice(1, N[n] + 1)
                sl[nP] = slice(1 + sP, 1 + sP + N[nP])
                sl[nPP] = slice(1 + sPP, 1 + sPP + N[nPP])
                w = _bcast(wP, nP) * _bcast(wPP, nPP)
                eA += epsP[tuple(sl)] * w
                kA += kapP[tuple(sl)] * w
        dl = _bcast(mesh.primal[n], n)
        with np.errstate(divide="ignore", invalid="ignore"):
            Cc = EPS0 * eA / dl
            G = kA / dl + mat.port_g[n]
            g = dT * G / (2 * Cc)
            v1 = (1 - g) / (1 + g)
            v2 = (dT / Cc) / (1 + g)
        ok = (dl > 0) & (eA > 0) & ~mat.pec[n]
        vv[n] = np.where(ok, v1, 0.0)
        vi[n] = np.where(ok, v2, 0.0)
        # --- H edges: reluctance of the dual edge, sum of w_i / mu_i over the
        # cells it crosses (series), L = mu0 * A' / that sum
        rel = np.zeros(N)
        for s, w in ((-1, mesh.half_lo[n]), (0, mesh.half_hi[n])):
            sl = [None, None, None]
            sl[n] = slice(1 + s, 1 + s + N[n])
            sl[nP] = slice(1, N[nP] + 1)
            sl[nPP] = slice(1, N[nPP] + 1)
            mu_c = muP[tuple(sl)]
            wb = _bcast(w, n)
            with np.errstate(divide="ignore", invalid="ignore"):
                rel += np.where((wb > 0) & (mu_c > 0), wb / mu_c, 0.0)
        area = _bcast(mesh.primal[nP], nP) * _bcast(mesh.primal[nPP], nPP)
        with np.errstate(divide="ignore", invalid="ignore"):
            L = MUE0 * area / rel
            v3 = dT / L
        okh = (area > 0) & (rel > 0)
        ii[n] = np.where(okh, 1.0, 0.0)
        iv[n] = np.where(okh, v3, 0.0)
    return {"vv": vv, "vi": vi, "ii": ii, "iv": iv}


def pml_sigma_profile(width_cells, dl, R0=1e-6, g=2.5):
    s0 = -np.log(R0) * np.log(g) / (2 * dl * Z0 * (g ** width_cells - 1))
    return lambda d: np.where(d > 0, s0 * g ** (d / dl), 0.0)


def pml_regions_for_shell(mesh, widths):
    """Non-overlapping UPML boxes for widths (xmin,xmax,ymin,ymax,zmin,zmax):
    x slabs span the full y,z range, y slabs the x interior, z slabs the
    x and y interior."""
    nx, ny, nz = mesh.shape
    w = widths
    regs = []
    if w[0]:
        regs.append(([0, 0, 0], [w[0], ny, nz]))
    if w[1]:
        regs.append(([nx - w[1], 0, 0], [w[1], ny, nz]))
    x0, x1 = w[0], nx - w[1]
    if w[2]:
        regs.append(([x0, 0, 0], [x1 - x0, w[2], nz]))
    if w[3]:
        regs.append(([x0, ny - w[3], 0], [x1 - x0, w[3], nz]))
    y0, y1 = w[2], ny - w[3]
    if w[4]:
        regs.append(([x0, y0, 0], [x1 - x0, y1 - y0, w[4]]))
    if w[5]:
        regs.append(([x0, y0, nz - w[5]], [x1 - x0, y1 - y0, w[5]]))
    return regs


def build_upml(mesh, coeffs, widths, dT):
    """UPML regions (reference configurePML format) for a PEC-terminated
    shell; overwrites the main coefficients inside the regions with the flux
    update.  Returns the region list."""
    N = mesh.shape
    # 1-D sigma profiles per axis at node and edge-midpoint positions
    s_node, s_mid = [], []
    for m in range(3):
        lines = mesh.lines[m]
        prof = []
        wlo, whi = widths[2 * m], widths[2 * m + 1]
        if wlo:
            prof.append((lines[wlo], -1, pml_sigma_profile(wlo, lines[1] - lines[0])))
        if whi:
            prof.append((lines[N[m] - 1 - whi], +1, pml_sigma_profile(whi, lines[-1] - lines[-2])))

        def sig(pos):
            s = np.zeros_like(pos)
            for iface, side, f in prof:
                s = s + f((iface - pos) if side < 0 else (pos - iface))
            return s
        s_node.append(sig(lines))
        s_mid.append(sig(lines + mesh.primal[m] / 2))
    scale = dT / (2 * EPS0)
    regions = []
    for start, num in pml_regions_for_shell(mesh, widths):
        sl = tuple(slice(s, s + k) for s, k in zip(start, num))
        r = {"startPos": list(start), "numLines": list(num)}
        arrs = {k: np.zeros((3,) + tuple(num)) for k in ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn")}
        for n in range(3):
            nP, nPP = (n + 1) % 3, (n + 2) % 3
            for kind in ("e", "h"):
                # E_n sits on nodes across n and mid-edge along n; H_n the opposite
                def s_at(m):
                    prof = (s_mid[m] if m == n else s_node[m]) if kind == "e" else \
                           (s_node[m] if m == n else s_mid[m])
                    return _bcast(prof[sl[m]], m) * scale
                s_n, s_p, s_pp = s_at(n), s_at(nP), s_at(nPP)
                a_vv, a_fo, a_fn = ("vv", "vvfo", "vvfn") if kind == "e" else ("ii", "iifo", "iifn")
                m_vv, m_vi = ("vv", "vi") if kind == "e" else ("ii", "iv")
                arrs[a_vv][n] = (1 - s_pp) / (1 + s_pp)
                arrs[a_fo][n] = (1 - s_n) / (1 + s_pp)
                arrs[a_fn][n] = (1 + s_n) / (1 + s_pp)
                main_vi = coeffs[m_vi][n][sl]
                live = main_vi != 0
                coeffs[m_vv][n][sl] = np.where(live, (1 - s_p) / (1 + s_p), 0.0)
                coeffs[m_vi][n][sl] = np.where(live, main_vi / (1 + s_p), 0.0)
        for k, v in arrs.items():
            r[k] = v.astype(np.float32).ravel()
        regions.append(r)
    return regions


def pec_boundary(mat, faces=range(6)):
    """PEC on the given outer faces: tangential E on the face plane is zero."""
    N = mat.mesh.shape
    for face in faces:
        ax, top = face // 2, face % 2 == 1
        idx = N[ax] - 1 if top else 0
        for n in range(3):
            if n == ax:
                continue
            s = [slice(None)] * 3
            s[ax] = idx
            mat.pec[(n,) + tuple(s)] = True


def gaussian_pulse(f0, fc, dT):
    """openEMS-style Gaussian: cos(2 pi f0 (t - T0)) exp(-((t - T0) 2 pi fc / 3)^2),
    T0 = 9 / (2 pi fc), over 2 T0.  Sample 0 is zero and sample k >= 1 is the
    pulse at t = (k-1) dT: the engine reads sample 0 once the pulse has
    ended (excitation position past the end), so that must inject nothing."""
    T0 = 9.0 / (2 * np.pi * fc)
    n = int(2 * T0 / dT)
    t = np.arange(n) * dT
    pulse = np.cos(2 * np.pi * f0 * (t - T0)) * np.exp(-((t - T0) * 2 * np.pi * fc / 3) ** 2)
    return np.concatenate([[0.0], pulse]).astype(np.float32)


def mur_config(mesh, faces, dT):
    """First-order Mur on the given faces (0..5 = xmin,xmax,ymin,...), dual
    nyP/nyPP form: coefficient (c dT - dl)/(c dT + dl) per tangential E
    component (tests/test_webgpu.mjs:2360-2380)."""
    nx, ny, nz = mesh.shape
    N = nx * ny * nz
    dims = mesh.shape
    out = {k: [] for k in ("coeff_nyP", "coeff_nyPP", "normal_idx_nyP", "shifted_idx_nyP", "normal_idx_nyPP",
                           "shifted_idx_nyPP")}
    for face in faces:
        ny_ = face // 2
        top = face % 2 == 1
        line = dims[ny_] - 1 if top else 0
        shift = line - 1 if top else 1
        dl = abs(mesh.lines[ny_][shift] - mesh.lines[ny_][line])
        c = (C0 * dT - dl) / (C0 * dT + dl)
        nP, nPP = (ny_ + 1) % 3, (ny_ + 2) % 3
        for tag, comp in (("nyP", nP), ("nyPP", nPP)):
            a = np.arange(dims[nP])
            b = np.arange(dims[nPP])
            A, B = np.meshgrid(a, b, indexing="ij")
            pos = [None, None, None]
            pos[ny_] = np.full(A.shape, line)
            pos[nP], pos[nPP] = A, B
            sh = list(pos)
            sh[ny_] = np.full(A.shape, shift)
            lin_n = (pos[0] * ny * nz + pos[1] * nz + pos[2]).ravel()
            lin_s = (sh[0] * ny * nz + sh[1] * nz + sh[2]).ravel()
            out["normal_idx_" + tag].append(comp * N + lin_n)
            out["shifted_idx_" + tag].append(comp * N + lin_s)
            out["coeff_" + tag].append(np.full(lin_n.shape, c))
    return {k: (np.concatenate(v).astype(np.float32 if k.startswith("coeff") else np.uint32)) for k, v in out.items()}


def _face_edge_lengths(lines):
    """Midpoint-rule lengths of the lines of one face: half the distance to
    each neighbour inside the face (app/sim-worker.js:626-639)."""
    n = len(lines)
    out = np.zeros(n)
    if n > 1:
        out[0] = 0.5 * abs(lines[1] - lines[0])
        out[1:-1] = 0.5 * np.abs(lines[2:] - lines[:-2])
        out[-1] = 0.5 * abs(lines[-1] - lines[-2])
    return out


def nf2ff_box(mesh, lo, hi):
    """Six-face NF2FF surface at node index bounds lo..hi (inclusive), as the
    reference's worker builds it (app/sim-worker.js:429-495, 563-624): faces
    in the order x-, x+, y-, y+, z-, z+, points of a face in x, y, z loop
    order; point metadata (pos xyz, normalDir, normSign, area, 0, 0) with the
    area from the face's own midpoint-rule edge lengths; and the engine's
    primal/dual edge-length tables (app/sim-worker.js:772-785)."""
    pts, meta = [], []
    for ax in range(3):
        for side, sign in ((lo[ax], -1.0), (hi[ax], 1.0)):
            rng = [np.arange(lo[d], hi[d] + 1) for d in range(3)]
            rng[ax] = np.array([side])
            X, Y, Z = np.meshgrid(*rng, indexing="ij")
            p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.int64)
            pts.append(p)
            m = np.zeros((len(p), 8))
            for d in range(3):
                m[:, d] = mesh.lines[d][p[:, d]]
            a1, a2 = (ax + 1) % 3, (ax + 2) % 3
            e1 = _face_edge_lengths(mesh.lines[a1][rng[a1]])
            e2 = _face_edge_lengths(mesh.lines[a2][rng[a2]])
            m[:, 3] = ax
            m[:, 4] = sign
            m[:, 5] = e1[p[:, a1] - rng[a1][0]] * e2[p[:, a2] - rng[a2][0]]
            meta.append(m)
    pts = np.concatenate(pts).astype(np.uint32)
    meta = np.concatenate(meta).astype(np.float32)
    primal = [mesh.primal[d].astype(np.float32) for d in range(3)]
    for d in range(3):
        if len(primal[d]) > 1:
            primal[d][-1] = primal[d][-2]
    dual = [mesh.dual[d].astype(np.float32) for d in range(3)]
    return pts.ravel(), meta.ravel(), primal, dual


def to_f32(coeffs):
    return {k: np.ascontiguousarray(v, dtype=np.float32).ravel() for k, v in coeffs.items()}
