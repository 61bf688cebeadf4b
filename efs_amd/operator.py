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

    def timestep(self, factor=1.0, wp_max=0.0):
        """CFL limit of the smallest cell, times ``factor``.  With Drude or
        Lorentz poles (largest plasma angular frequency wp_max) the
        leapfrog update of field and polarisation current is stable when
        (dT / dT_cfl)^2 + (wp_max dT / 2)^2 <= 1, which this limit meets."""
        m = [np.min(d[:-1]) if len(d) > 1 else np.inf for d in self.primal]
        s = sum(1.0 / v ** 2 for v in m if np.isfinite(v))
        return factor / np.sqrt(C0 * C0 * s + wp_max * wp_max / 4)


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


def build_coefficients_threaded(mesh, mat, dT, threads=8, chunk=None):
    """build_coefficients over x-slabs on a thread pool (numpy releases the
    interpreter lock in its array loops).  Each slab is built as its own
    sub-grid with one node of halo on each side, so every node's arithmetic
    sees the same operands as in the whole-grid build: the result is the
    same numbers."""
    from concurrent.futures import ThreadPoolExecutor
    nx = mesh.n[0]
    if threads <= 1 or nx < 8:
        return build_coefficients(mesh, mat, dT)
    out = {k: np.zeros((3,) + mesh.shape) for k in ("vv", "vi", "ii", "iv")}
    chunk = chunk or max(2, -(-nx // (4 * threads)))

    def work(x0):
        x1 = min(nx, x0 + chunk)
        a, b = max(0, x0 - 1), min(nx, x1 + 1)  # nodes a .. b-1 (halo included)
        sub = Mesh(mesh.lines[0][a:b], mesh.lines[1], mesh.lines[2])
        sm = Materials.__new__(Materials)
        sm.mesh = sub
        sm.eps, sm.kappa, sm.mu = mat.eps[a:b - 1], mat.kappa[a:b - 1], mat.mu[a:b - 1]
        sm.pec, sm.port_g = mat.pec[:, a:b], mat.port_g[:, a:b]
        co = build_coefficients(sub, sm, dT)
        for k in out:
            out[k][:, x0:x1] = co[k][:, x0 - a:x1 - a]
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(0, nx, chunk)))
    return out


def pml_sigma_profile(width_cells, dl, R0=1e-6, g=2.5):
    s0 = -np.log(R0) * np.log(g) / (2 * dl * Z0 * (g ** width_cells - 1))
    return lambda d: np.where(d > 0, s0 * g ** (d / dl), 0.0)


def pml_regions_for_shell(mesh, widths):
    """Non-overlapping UPML boxes for widths (xmin,xmax,ymin,ymax,zmin,zmax),
    each PML_w slab w + 1 lines deep as openEMS sets its UPML range
    (Operator_Ext_UPML::Create_UPML: lower lines 0..w, upper lines
    N-1-w..N-1, the grading measured from line w resp. N-1-w): the upper
    slab's first line holds the half-cell-deep E and H components (E along
    the axis, H across it), which the grading gives a nonzero conductivity;
    the lower slab's last line has none, but openEMS computes its
    coefficients with the UPML formulas too.  x slabs span the full y,z
    range, y slabs the x interior, z slabs the x and y interior."""
    nx, ny, nz = mesh.shape
    w = [k + 1 if k else 0 for k in widths]
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
    """openEMS's Gaussian excitation: sample k = 0 .. int(2 T0 / dT) is
    cos(2 pi f0 (t - T0)) exp(-((t - T0) 2 pi fc / 3)^2) at t = k dT,
    T0 = 9 / (2 pi fc).  The engine (like openEMS's Apply2Voltages) reads
    sample 0 once the pulse has ended, so openEMS keeps injecting g(0) =
    cos(2 pi f0 T0) e^-9 after the pulse; the reference's native openEMS
    fixtures carry that constant (tests/fixtures/coax: the late-time mean of
    the port voltage and current), and these samples reproduce them."""
    T0 = 9.0 / (2 * np.pi * fc)
    n = int(2 * T0 / dT) + 1
    t = np.arange(n) * dT
    return (np.cos(2 * np.pi * f0 * (t - T0)) * np.exp(-((t - T0) * 2 * np.pi * fc / 3) ** 2)).astype(np.float32)


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


def _edge_share(mesh, cellv, n):
    """Sum over the (up to 4) cells around each E edge along n of a cell
    quantity times the cell's share of the edge's dual face -- the weights
    build_coefficients uses for eps (cells outside the grid count 0)."""
    N = mesh.shape
    P = _pad_cells(np.asarray(cellv, np.float64))
    nP, nPP = (n + 1) % 3, (n + 2) % 3
    out = np.zeros(N)
    for sP, wP in ((-1, mesh.half_lo[nP]), (0, mesh.half_hi[nP])):
        for sPP, wPP in ((-1, mesh.half_lo[nPP]), (0, mesh.half_hi[nPP])):
            sl = [None, None, None]
            sl[n] = slice(1, N[n] + 1)
            sl[nP] = slice(1 + sP, 1 + sP + N[nP])
            sl[nPP] = slice(1 + sPP, 1 + sPP + N[nPP])
            out += P[tuple(sl)] * (_bcast(wP, nP) * _bcast(wPP, nPP))
    return out


def ade_orders(mesh, coeffs, dT, poles):
    """Drude / Lorentz poles as ADE orders in the configureLorentzADE format
    (src/webgpu-engine.mjs:1358-1475), for the openEMS write-back form
    (field -= ADE state after each update).

    poles: list of {"mask": bool over the cells, "wp": plasma angular
    frequency, "w0": pole angular frequency (0 for Drude), "gamma": damping
    (1/s)}; eps_inf is the cells' eps in the main coefficients.

    Model (per E edge, I_p the polarisation current through the dual face,
    s = sum over the touching cells of eps0 wp^2 x dual-face share):
        dI_p/dt = -gamma I_p - w0^2 Q + s E,   dQ/dt = I_p
    with the central-difference ADE of Taflove & Hagness (ch. 9):
        I_p^{n+1/2} = k I_p^{n-1/2} + dT s E^n / (1 + gamma dT/2) - (Lorentz term)
        k = (1 - gamma dT/2) / (1 + gamma dT/2).
    In the engine's units the state is A = vi I_p (volts, what the write-back
    subtracts from the edge voltage V = E dl) and, for Lorentz, L = w0^2 Q dl / s:
        L   += v_Lor A,        v_Lor = dT w0^2 dl / (s vi)
        A    = v_int A + v_ext (V - L)   (Drude: L = 0)
        v_int = k,   v_ext = vi dT s / (dl (1 + gamma dT / 2)).
    Each order lists the edges any direction couples to (shared positions,
    zero coefficients where a direction has none); the H side is zero."""
    N = mesh.shape
    orders = []
    if poles and dT > mesh.timestep(wp_max=max(float(p["wp"]) for p in poles)) * (1 + 1e-12):
        raise ValueError("dT exceeds the stability limit with these poles (Mesh.timestep(wp_max=...))")
    for pole in poles:
        wp, w0, gam = float(pole["wp"]), float(pole.get("w0", 0.0) or 0.0), float(pole.get("gamma", 0.0) or 0.0)
        cellv = np.where(np.asarray(pole["mask"], bool), EPS0 * wp * wp, 0.0)
        k = (1 - gam * dT / 2) / (1 + gam * dT / 2)
        per = []
        live = np.zeros(N, bool)
        for n in range(3):
            s = _edge_share(mesh, cellv, n)
            vi = coeffs["vi"][n].reshape(N)
            dl = _bcast(mesh.primal[n], n)
            ok = (s > 0) & (vi != 0) & (dl > 0)
            per.append((s, vi, np.broadcast_to(dl, N), ok))
            live |= ok
        pos = np.flatnonzero(live.ravel())
        dirs = []
        for n, (s, vi, dl, ok) in enumerate(per):
            s_, vi_, dl_, ok_ = (a.ravel()[pos] for a in (s, vi, dl, ok))
            with np.errstate(divide="ignore", invalid="ignore"):
                v_ext = np.where(ok_, vi_ * dT * s_ / (dl_ * (1 + gam * dT / 2)), 0.0)
                v_lor = np.where(ok_, dT * w0 * w0 * dl_ / (s_ * vi_), 0.0)
            d = {"dir": n, "pos_idx": pos.astype(np.uint32),
                 "v_int_ADE": np.where(ok_, k, 0.0).astype(np.float32),
                 "v_ext_ADE": v_ext.astype(np.float32)}
            if w0 > 0:
                d["v_Lor_ADE"] = v_lor.astype(np.float32)
            dirs.append(d)
        orders.append({"numCells": int(len(pos)), "hasLorentz": w0 > 0, "directions": dirs})
    return {"orders": orders}


def _tf_mask(mesh, lo, hi, comp, kind):
    """Total-field membership of component comp of E ("e") or H ("h") on the
    grid: its position (an edge midpoint along comp for E, across comp for
    H, in doubled index units) lies in the closed box [lo, hi] of nodes."""
    axes = []
    for a in range(3):
        i = np.arange(mesh.n[a])
        half = (a == comp) if kind == "e" else (a != comp)
        c2 = 2 * i + (1 if half else 0)
        axes.append((c2 >= 2 * lo[a]) & (c2 <= 2 * hi[a]))
    return axes[0][:, None, None] & axes[1][None, :, None] & axes[2][None, None, :]


def tfsf_plane_wave(mesh, coeffs, lo, hi, dT, signal, prop_axis=0, prop_sign=1, e_axis=2, e0=1.0):
    """Total-field / scattered-field plane wave in the configureTFSF format
    (src/webgpu-engine.mjs:1527-1645; injection F[idx] += amp * s(numTS -
    delay_int + delay_frac), src/webgpu-fdtd.mjs:683-739).

    The wave travels along prop_sign * axis prop_axis with E along e_axis:
    E_inc = e0 s(t - tau) e_hat, H_inc = (k_hat x e_hat) e0 / Z0 s(t - tau),
    tau = distance past the entry face / c.  The total-field region is the
    closed node box [lo, hi]; every update whose stencil mixes a total-field
    and a scattered-field value gets one point: the neighbour's incident
    value (E dl on a primal edge, H dl' on a dual edge) times the update's
    curl coefficient and sign, added when the updated value is total and
    the neighbour scattered, subtracted in the opposite case.  Timing: the
    voltage update of step n reads currents at (n + 1/2) dT, the current
    update voltages at (n + 1) dT; delays are shifted by a whole number of
    steps so that none is negative.  The box must lie in vacuum, outside
    UPML."""
    N = mesh.shape
    ehat = np.zeros(3)
    ehat[e_axis] = 1.0
    khat = np.zeros(3)
    khat[prop_axis] = float(prop_sign)
    hhat = np.cross(khat, ehat)
    a = prop_axis
    ref = mesh.lines[a][lo[a]] if prop_sign > 0 else mesh.lines[a][hi[a]]
    node = mesh.lines[a]
    mid = node + mesh.primal[a] / 2
    tfE = [_tf_mask(mesh, lo, hi, n, "e") for n in range(3)]
    tfH = [_tf_mask(mesh, lo, hi, n, "h") for n in range(3)]
    lists = {"v": [], "i": []}
    Ntot = int(np.prod(N))

    def shifted(arr, axis, off):
        """arr at p + off e_axis (False / 0 outside the grid)."""
        out = np.zeros_like(arr)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        if off < 0:
            src[axis], dst[axis] = slice(0, N[axis] + off), slice(-off, None)
        else:
            src[axis], dst[axis] = slice(off, None), slice(0, N[axis] - off)
        out[tuple(dst)] = arr[tuple(src)]
        return out

    for kind in ("v", "i"):
        own = tfE if kind == "v" else tfH
        other = tfH if kind == "v" else tfE
        ohat = hhat / Z0 if kind == "v" else ehat
        step = -1 if kind == "v" else 1
        coef = coeffs["vi"] if kind == "v" else coeffs["iv"]
        for n in range(3):
            nP, nPP = (n + 1) % 3, (n + 2) % 3
            # (neighbour component, offset axis or None, sign) of the update of n
            terms = ((nPP, None, 1.0), (nPP, nP, -1.0), (nP, None, -1.0), (nP, nPP, 1.0))
            cn = coef[n].reshape(N)
            for m, ax, sgn in terms:
                if ohat[m] == 0.0:
                    continue
                nb = other[m] if ax is None else shifted(other[m], ax, step)
                valid = np.ones(N, bool)
                if ax is not None:
                    # at the grid edge the stencil reads the cell itself
                    edge = [slice(None)] * 3
                    edge[ax] = 0 if step < 0 else N[ax] - 1
                    valid[tuple(edge)] = False
                mix = (own[n] != nb) & valid & (cn != 0)
                idx = np.argwhere(mix)
                if len(idx) == 0:
                    continue
                side = np.where(own[n][mix], 1.0, -1.0)
                q = idx.copy()
                if ax is not None:
                    q[:, ax] += step
                # incident value of neighbour component m at q
                if kind == "v":
                    length = mesh.dual[m][q[:, m]]
                    halfa = m != a  # H_m sits at a midpoint along axes other than m
                else:
                    length = mesh.primal[m][q[:, m]]
                    halfa = m == a
                pos = mid[q[:, a]] if halfa else node[q[:, a]]
                tau = prop_sign * (pos - ref) / C0
                inc = e0 * ohat[m] * length
                amp = cn[mix] * sgn * side * inc
                lin = n * Ntot + (idx[:, 0] * N[1] + idx[:, 1]) * N[2] + idx[:, 2]
                D = tau / dT - (0.5 if kind == "v" else 1.0)
                lists[kind].append((lin, D, amp))
    Dall = np.concatenate([np.concatenate([t[1] for t in lists[k]]) for k in lists if lists[k]])
    shift = max(0.0, float(np.ceil(-Dall.min()))) if len(Dall) else 0.0
    out = {"signal": np.asarray(signal, np.float32), "period": 0}
    for kind, key in (("v", "voltagePoints"), ("i", "currentPoints")):
        if not lists[kind]:
            out[key] = {"field_idx": np.zeros(0, np.uint32), "delay_int": np.zeros(0, np.uint32),
                        "delay_frac": np.zeros(0, np.float32), "amp": np.zeros(0, np.float32)}
            continue
        lin = np.concatenate([t[0] for t in lists[kind]])
        D = np.concatenate([t[1] for t in lists[kind]]) + shift
        amp = np.concatenate([t[2] for t in lists[kind]])
        di = np.ceil(D - 1e-9)
        out[key] = {"field_idx": lin.astype(np.uint32), "delay_int": di.astype(np.uint32),
                    "delay_frac": np.clip(di - D, 0.0, 1.0).astype(np.float32), "amp": amp.astype(np.float32)}
    out["delay_shift"] = shift
    return out


def to_f32(coeffs):
    return {k: np.ascontiguousarray(v, dtype=np.float32).ravel() for k, v in coeffs.items()}
