"""openEMS XML front end: an openEMS simulation file (the document the
reference's script API writes, app/ems-api.mjs:847-975, and the fixtures'
generator writes, tests/generate_fixtures.py) turned into the engine's
configuration through the FP64 builder (efs_amd/operator.py).

Read:
  <FDTD NumberOfTimesteps endCriteria f_max>  with <Excitation Type="0"
      f0 fc/> (Gaussian) and <BoundaryCond xmin..zmax> (0/PEC, 2/MUR,
      3/PML_n with PML_xmin="n" or the value "PML_n")
  <RectilinearGrid DeltaUnit> with <XLines>/<YLines>/<ZLines>
  <Properties>: Metal, Material (<Property Epsilon Kappa Mue>),
      LumpedElement (Direction, R; C = L = 0), Excitation (Type 0, Excite,
      <Weight X Y Z> expressions), ProbeBox (Type 0 voltage / 1 current,
      Weight, NormDir), DumpBox nf2ff_E / nf2ff_H (<FD_Samples>)
  primitives: Box, Cylinder, CylindricalShell, Sphere, Curve
Everything else (PMC, periodic boundaries, other primitives or excitation
types, transformations, dispersive materials) raises NotImplementedError
naming what it met.

Voxelisation (coordinates in drawing units, DeltaUnit applied to the mesh):
  - a cell takes the Material of the highest-priority primitive holding its
    centre (later properties win ties);
  - an E edge is PEC when a Metal primitive holds the edge's centre;
  - a LumpedElement box holds strings of edges along Direction in series
    (each edge's share of R in proportion to its length), the strings in
    parallel: G = L / (R n_parallel dl) per edge;
  - an Excitation drives each edge along n whose centre a primitive holds
    with amplitude Excite[n] * weight_n(centre) * dl; a Curve drives the
    mesh path between its vertices (diagonal rounds, see _Grid.curve_path);
  - probe boxes snap to the nearest mesh line, ties to the lower index:
    primal lines for voltage probes, dual lines for current probes (the
    rule openEMS's own probe headers show in the reference's coax fixture,
    tests/fixtures/coax/probe_*.csv);
  - the NF2FF box spans the snapped extent of the nf2ff_E faces.
dT is the CFL limit of the smallest cell (Mesh.timestep).
"""
import ast
import math
import xml.etree.ElementTree as ET

import numpy as np

from . import operator as op
from . import ports

# Interpolation of E / H point probes (openEMS ProbeBox types 2 and 3).
FIELD_PROBE_INTERPOLATION = "none"

_FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
_SUPPORTED_PRIMS = ("Box", "Cylinder", "CylindricalShell", "Sphere", "Curve")


# ---------------------------------------------------------------- parsing
def _f(v, default=None):
    if v is None:
        return default
    return float(v)


def _point(el):
    return np.array([float(el.get("X")), float(el.get("Y")), float(el.get("Z"))])


def _lines(text):
    vals = [float(v) for v in (text or "").replace(";", ",").split(",") if v.strip()]
    return np.unique(np.array(vals, np.float64))


def _parse_primitive(el):
    kind = el.tag
    if kind not in _SUPPORTED_PRIMS:
        raise NotImplementedError("openEMS primitive %r is not supported" % kind)
    if el.find("Transformation") is not None:
        raise NotImplementedError("primitive transformations are not supported")
    p = {"kind": kind, "priority": int(float(el.get("Priority", "0")))}
    if kind in ("Box", "Cylinder", "CylindricalShell"):
        p["p1"], p["p2"] = _point(el.find("P1")), _point(el.find("P2"))
    if kind in ("Cylinder", "CylindricalShell", "Sphere"):
        p["radius"] = float(el.get("Radius"))
    if kind == "CylindricalShell":
        p["width"] = float(el.get("ShellWidth"))
    if kind == "Sphere":
        c = el.find("Center") if el.find("Center") is not None else el.find("P1")
        p["center"] = _point(c)
    if kind == "Curve":
        p["vertices"] = [_point(v) for v in el.findall("Vertex")]
    return p


def parse(text):
    """The simulation described by an openEMS XML document, as plain dicts."""
    root = ET.fromstring(text)
    if root.tag != "openEMS":
        raise ValueError("not an openEMS document (root <%s>)" % root.tag)
    fd = root.find("FDTD")
    if fd is None:
        raise ValueError("no <FDTD> element")
    exc = fd.find("Excitation")
    sim = {"nTS": int(float(fd.get("NumberOfTimesteps", "0"))), "endCriteria": _f(fd.get("endCriteria"), 0.0),
           "f_max": _f(fd.get("f_max"), 0.0)}
    if exc is None:
        raise ValueError("no FDTD <Excitation>")
    sim["excitation"] = {"type": int(float(exc.get("Type", "0"))), "f0": _f(exc.get("f0"), 0.0),
                         "fc": _f(exc.get("fc"), 0.0)}
    bc_el = fd.find("BoundaryCond")
    bcs = []
    for face in _FACES:
        v = (bc_el.get(face, "0") if bc_el is not None else "0").strip()
        if v.upper().startswith("PML_"):
            bcs.append(("PML", int(v[4:])))
        elif v in ("3", "PML"):
            n = bc_el.get("PML_" + face)
            bcs.append(("PML", int(n) if n is not None else 8))
        elif v in ("0", "PEC"):
            bcs.append(("PEC", 0))
        elif v in ("2", "MUR"):
            bcs.append(("MUR", 0))
        elif v in ("1", "PMC"):
            raise NotImplementedError("PMC boundaries are not supported")
        else:
            raise NotImplementedError("boundary condition %r on %s is not supported" % (v, face))
    sim["boundaries"] = bcs
    cs = root.find("ContinuousStructure")
    if cs is None:
        raise ValueError("no <ContinuousStructure>")
    grid = cs.find("RectilinearGrid")
    unit = float(grid.get("DeltaUnit", "1"))
    sim["unit"] = unit
    sim["lines"] = [_lines(grid.findtext(t)) for t in ("XLines", "YLines", "ZLines")]
    props = []
    pe = cs.find("Properties")
    for el in (list(pe) if pe is not None else []):
        prop = {"type": el.tag, "name": el.get("Name", ""), "attrs": dict(el.attrib)}
        if el.tag == "Material":
            pr = el.find("Property")
            a = dict(pr.attrib) if pr is not None else {}
            if el.get("Isotropy") == "0" or any("," in a.get(k, "") for k in ("Epsilon", "Kappa", "Mue")):
                raise NotImplementedError("anisotropic materials are not supported (%s)" % prop["name"])
            prop["epsilon"] = float(a.get("Epsilon", "1"))
            prop["kappa"] = float(a.get("Kappa", "0"))
            prop["mue"] = float(a.get("Mue", "1"))
            if float(a.get("Sigma", "0")) != 0:
                raise NotImplementedError("magnetic conductivity (Sigma) is not supported")
        elif el.tag == "Excitation":
            w = el.find("Weight")
            prop["weight"] = [w.get(k, "1") for k in ("X", "Y", "Z")] if w is not None else None
        elif el.tag == "DumpBox":
            prop["fd_samples"] = [float(v) for v in (el.findtext("FD_Samples") or "").split(",") if v.strip()]
        elif el.tag not in ("Metal", "LumpedElement", "ProbeBox"):
            raise NotImplementedError("openEMS property %r is not supported" % el.tag)
        prims = el.find("Primitives")
        prop["primitives"] = [_parse_primitive(p) for p in (list(prims) if prims is not None else [])]
        props.append(prop)
    sim["properties"] = props
    return sim


# --------------------------------------------------- weight expressions
_FUNCS = {"sqrt": np.sqrt, "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
          "atan": np.arctan, "atan2": np.arctan2, "exp": np.exp, "log": np.log, "log10": np.log10, "abs": np.abs,
          "pow": np.power, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh, "min": np.minimum,
          "max": np.maximum}
_OPS = {ast.Add: np.add, ast.Sub: np.subtract, ast.Mult: np.multiply, ast.Div: np.divide, ast.Pow: np.power}


def weight(expr, x, y, z):
    """Value of an openEMS weight expression (function-parser syntax: + - *
    / ^, the usual functions, variables x y z, rho, a (azimuth), r, t
    (polar angle)) at points in drawing units; non-finite values give 0.
    Only arithmetic is evaluated: the expression is parsed, never exec'd."""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    env = {"x": x, "y": y, "z": z, "rho": np.hypot(x, y), "a": np.arctan2(y, x),
           "r": np.sqrt(x * x + y * y + z * z), "pi": np.pi, "e": np.e}
    with np.errstate(divide="ignore", invalid="ignore"):
        env["t"] = np.arccos(np.where(env["r"] > 0, z / np.where(env["r"] > 0, env["r"], 1.0), 1.0))
    tree = ast.parse(str(expr).replace("^", "**"), mode="eval")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return np.float64(n.value)
        if isinstance(n, ast.Name) and n.id in env:
            return env[n.id]
        if isinstance(n, ast.BinOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS and not n.keywords:
            return _FUNCS[n.func.id](*[ev(a) for a in n.args])
        raise ValueError("unsupported element in weight expression %r" % (expr,))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        v = np.broadcast_to(ev(tree), x.shape).astype(np.float64)
    return np.where(np.isfinite(v), v, 0.0)


# ------------------------------------------------------------ geometry
def _inside(prim, p, tol):
    """Points p (..., 3) held by the primitive (closed, with tolerance)."""
    k = prim["kind"]
    if k == "Box":
        lo, hi = np.minimum(prim["p1"], prim["p2"]), np.maximum(prim["p1"], prim["p2"])
        return np.all((p >= lo - tol) & (p <= hi + tol), axis=-1)
    if k in ("Cylinder", "CylindricalShell"):
        a, b = prim["p1"], prim["p2"]
        L = np.linalg.norm(b - a)
        u = (b - a) / L
        d = p - a
        t = d @ u
        rad = np.linalg.norm(d - t[..., None] * u, axis=-1)
        along = (t >= -tol) & (t <= L + tol)
        if k == "Cylinder":
            return along & (rad <= prim["radius"] + tol)
        return along & (np.abs(rad - prim["radius"]) <= prim["width"] / 2 + tol)
    if k == "Sphere":
        return np.linalg.norm(p - prim["center"], axis=-1) <= prim["radius"] + tol
    raise NotImplementedError("primitive %r cannot hold volumes or edges" % k)


def snap(lines, v, dual=False):
    """Index of the nearest mesh line (dual: the midpoint between lines i and
    i + 1 has index i), ties to the lower index."""
    pos = 0.5 * (lines[:-1] + lines[1:]) if dual else lines
    d = np.abs(pos - v)
    return int(np.flatnonzero(d <= d.min() * (1 + 1e-12) + 1e-15 * max(1.0, abs(v)))[0])


class _Grid:
    def __init__(self, lines):
        self.lines = lines
        self.n = [len(a) for a in lines]
        span = max(float(a[-1] - a[0]) if len(a) > 1 else 1.0 for a in lines)
        self.tol = 1e-9 * span

    def edge_centres(self, n):
        """Centres of the E edges along n at every node (last node: none)."""
        c = [self.lines[m].copy() for m in range(3)]
        c[n] = np.concatenate([0.5 * (c[n][:-1] + c[n][1:]), [np.nan]])
        X, Y, Z = np.meshgrid(*c, indexing="ij")
        return np.stack([X, Y, Z], axis=-1)

    def cell_centres(self):
        c = [0.5 * (a[:-1] + a[1:]) if len(a) > 1 else a for a in self.lines]
        X, Y, Z = np.meshgrid(*c, indexing="ij")
        return np.stack([X, Y, Z], axis=-1)

    def curve_path(self, vertices):
        """Edges (n, i, j, k, sign) of the mesh path through the vertices:
        each leg snaps its ends to mesh nodes and walks one edge at a time
        from the start node to the end node, each step along the axis whose
        next node lies nearest the straight line between the leg's two
        vertices (ties to the lower axis), the grid walk openEMS uses for
        curves (Operator::FindPath).  A one-cell diagonal of the
        reference's cavity fixture thus drives its edges shortest cell side
        first, which reproduces openEMS's probe trace (tests/golden/
        cavity_openems.json)."""
        out = []
        for va, vb in zip(vertices[:-1], vertices[1:]):
            a = [snap(self.lines[m], va[m]) for m in range(3)]
            b = [snap(self.lines[m], vb[m]) for m in range(3)]
            p0 = np.asarray(va, np.float64)
            d = np.asarray(vb, np.float64) - p0
            dd = float(d @ d)

            def dist2(node):
                q = np.array([self.lines[m][node[m]] for m in range(3)]) - p0
                t = float(q @ d) / dd if dd > 0 else 0.0
                r = q - t * d
                return float(r @ r)
            cur = list(a)
            while cur != b:
                best, bm, bs = None, -1, 0
                for m in range(3):
                    if cur[m] == b[m]:
                        continue
                    step = 1 if b[m] > cur[m] else -1
                    nxt = list(cur)
                    nxt[m] += step
                    dv = dist2(nxt)
                    if best is None or dv < best:
                        best, bm, bs = dv, m, step
                e = list(cur)
                e[bm] = cur[bm] if bs > 0 else cur[bm] - 1
                out.append((bm, e[0], e[1], e[2], float(bs)))
                cur[bm] += bs
        return out


# --------------------------------------------------------------- build
def build(sim, k_block=64):
    """Engine configuration (the workload dictionary of efs_amd.workloads:
    gridSize, coefficients, mur or pmlRegions, excitation, probes with
    names, nf2ff + pointMeta + freqs when the file has NF2FF dump boxes,
    dT, mesh, nTS, f0, fc) of a parsed or textual openEMS simulation."""
    if isinstance(sim, str):
        sim = parse(sim)
    unit = sim["unit"]
    g = _Grid(sim["lines"])
    mesh = op.Mesh(*(a * unit for a in sim["lines"]))
    if min(mesh.n) < 2:
        raise ValueError("every axis needs at least two mesh lines")
    dT = mesh.timestep()
    mat = op.Materials(mesh)
    N = mesh.shape
    props = sim["properties"]
    # materials: cell centres, ascending priority (later properties win ties)
    mats = [(pr["priority"], k, prop) for k, prop in enumerate(props) if prop["type"] == "Material"
            for pr in prop["primitives"]]
    if mats:
        cc = g.cell_centres()
        for _, _, prop in sorted(mats, key=lambda t: (t[0], t[1])):
            for pr in prop["primitives"]:
                m = _inside(pr, cc, g.tol)
                mat.eps[m] = prop["epsilon"]
                mat.kappa[m] = prop["kappa"]
                mat.mu[m] = prop["mue"]
    # metal: E edges whose centre a Metal primitive holds
    for prop in props:
        if prop["type"] != "Metal":
            continue
        for n in range(3):
            ec = g.edge_centres(n)
            for pr in prop["primitives"]:
                if pr["kind"] == "Curve":
                    for (m, i, j, k, _) in g.curve_path(pr["vertices"]):
                        mat.pec[m, i, j, k] = True
                    continue
                mat.pec[n] |= _inside(pr, ec, g.tol)
    # lumped resistors
    for prop in props:
        if prop["type"] != "LumpedElement":
            continue
        a = prop["attrs"]
        if float(a.get("C", "0") or 0) != 0 or float(a.get("L", "0") or 0) != 0:
            raise NotImplementedError("lumped C / L are not supported (%s)" % prop["name"])
        R = float(a.get("R", "0") or 0)
        if R <= 0:
            continue
        n = int(float(a["Direction"]))
        ec = g.edge_centres(n)
        for pr in prop["primitives"]:
            m = _inside(pr, ec, g.tol)
            if not m.any():
                continue
            idx = np.argwhere(m)
            a1, a2 = (n + 1) % 3, (n + 2) % 3
            strings = {(int(p[a1]), int(p[a2])) for p in idx}
            lo, hi = idx[:, n].min(), idx[:, n].max() + 1
            L = mesh.lines[n][hi] - mesh.lines[n][lo]
            for p in idx:
                mat.port_g[n, p[0], p[1], p[2]] += L / (R * len(strings) * mesh.primal[n][p[n]])
    # boundaries: PEC faces, and UPML terminated by PEC
    bcs = sim["boundaries"]
    pml_w = [w if kind == "PML" else 0 for kind, w in bcs]
    op.pec_boundary(mat, [f for f in range(6) if bcs[f][0] != "MUR"])
    coeffs = op.build_coefficients(mesh, mat, dT)
    out = {"gridSize": list(N), "dT": dT, "mesh": mesh, "nTS": sim["nTS"], "unit": unit}
    if any(pml_w):
        out["pmlRegions"] = op.build_upml(mesh, coeffs, pml_w, dT)
        out["n_pml_cells"] = sum(int(np.prod(r["numLines"])) for r in out["pmlRegions"])
    mur_faces = [f for f, (kind, _) in enumerate(bcs) if kind == "MUR"]
    if mur_faces:
        out["mur"] = op.mur_config(mesh, mur_faces, dT)
    # excitation signal and points
    ex = sim["excitation"]
    if ex["type"] != 0:
        raise NotImplementedError("FDTD excitation type %d is not supported (Gaussian only)" % ex["type"])
    out["f0"], out["fc"] = ex["f0"], ex["fc"]
    signal = op.gaussian_pulse(ex["f0"], ex["fc"], dT)
    pos, dirs, amp = [], [], []
    for prop in props:
        if prop["type"] != "Excitation":
            continue
        etype = int(float(prop["attrs"].get("Type", "0")))
        if etype not in (0, 1):
            raise NotImplementedError("excitation type %s is not supported (soft or hard E only)" % etype)
        vec = [float(v) for v in prop["attrs"].get("Excite", "0,0,0").split(",")]
        for n in range(3):
            if vec[n] == 0:
                continue
            ec = g.edge_centres(n)
            for pr in prop["primitives"]:
                if pr["kind"] == "Curve":
                    sel = [(i, j, k) for (m, i, j, k, _) in g.curve_path(pr["vertices"]) if m == n]
                    idx = np.array(sel, np.int64).reshape(-1, 3)
                else:
                    idx = np.argwhere(_inside(pr, ec, g.tol))
                if len(idx) == 0:
                    continue
                c = ec[idx[:, 0], idx[:, 1], idx[:, 2]]
                w = np.ones(len(idx)) if prop["weight"] is None else \
                    weight(prop["weight"][n], c[:, 0], c[:, 1], c[:, 2])
                a = vec[n] * w * mesh.primal[n][idx[:, n]]
                if etype == 1:  # hard E: the edge's own update is switched off, the source sets it
                    coeffs["vv"][n, idx[:, 0], idx[:, 1], idx[:, 2]] = 0.0
                    coeffs["vi"][n, idx[:, 0], idx[:, 1], idx[:, 2]] = 0.0
                keep = a != 0
                pos.append(((idx[:, 0] * N[1] + idx[:, 1]) * N[2] + idx[:, 2])[keep])
                dirs.append(np.full(int(keep.sum()), n))
                amp.append(a[keep])
    out["coefficients"] = op.to_f32(coeffs)
    if not pos:
        raise ValueError("the simulation has no excitation")
    pos = np.concatenate(pos)
    out["excitation"] = {"signal": signal, "delay": np.zeros(len(pos), np.uint32),
                         "amp": np.concatenate(amp).astype(np.float32), "dir": np.concatenate(dirs).astype(np.uint32),
                         "pos": pos.astype(np.uint32), "period": 0}
    # probes
    probes = []
    for prop in props:
        if prop["type"] != "ProbeBox":
            continue
        a = prop["attrs"]
        t = int(float(a.get("Type", "0")))
        wgt = float(a.get("Weight", "1"))
        for pr in prop["primitives"]:
            if pr["kind"] != "Box":
                raise NotImplementedError("probe primitive %r is not supported" % pr["kind"])
            p1, p2 = pr["p1"], pr["p2"]
            if t == 0:
                s = [snap(g.lines[m], p1[m]) for m in range(3)]
                e = [snap(g.lines[m], p2[m]) for m in range(3)]
                probes.append({"name": prop["name"], "type": 0, "start": s, "stop": e, "weight": wgt})
            elif t == 1:
                nd = int(float(a.get("NormDir", "-1")))
                if nd < 0:
                    flat = [m for m in range(3) if abs(p1[m] - p2[m]) <= g.tol]
                    if len(flat) != 1:
                        raise ValueError("current probe %s needs NormDir or one flat axis" % prop["name"])
                    nd = flat[0]
                lo = [snap(g.lines[m], min(p1[m], p2[m]), dual=True) for m in range(3)]
                hi = [snap(g.lines[m], max(p1[m], p2[m]), dual=True) for m in range(3)]
                hi[nd] = lo[nd] = snap(g.lines[nd], 0.5 * (p1[nd] + p2[nd]), dual=True)

                def inside(v, m):
                    return bool(0 <= v <= N[m] - 2)
                probes.append({"name": prop["name"], "type": 1, "start": lo, "stop": hi, "normDir": nd,
                               "weight": wgt, "startInside": [inside(lo[m], m) for m in range(3)],
                               "stopInside": [inside(hi[m], m) for m in range(3)]})
            elif t in (2, 3):  # E / H field at the node nearest P1
                node = [snap(g.lines[m], p1[m]) for m in range(3)]
                probes.append({"name": prop["name"], "type": t, "start": node, "stop": node, "weight": wgt,
                               "taps": ports.field_taps(mesh, t, node, FIELD_PROBE_INTERPOLATION)})
            else:
                raise NotImplementedError("probe type %d is not supported" % t)
    out["probes"] = probes
    # NF2FF surface from the nf2ff_E dump faces
    nf = [p for p in props if p["type"] == "DumpBox" and p["name"] == "nf2ff_E"]
    if nf:
        pts = np.array([q for pr in nf[0]["primitives"] for q in (pr["p1"], pr["p2"])])
        lo = [snap(g.lines[m], pts[:, m].min()) for m in range(3)]
        hi = [snap(g.lines[m], pts[:, m].max()) for m in range(3)]
        freqs = np.array(nf[0]["fd_samples"] or [sim["f_max"] / 2], np.float64)
        sp, meta, primal, dual = op.nf2ff_box(mesh, lo, hi)
        out["nf2ff"] = {"surfaceIndices": sp, "numPoints": len(sp) // 3, "omega": list(2 * math.pi * freqs),
                        "dT": dT, "primalEdgeLens": primal, "dualEdgeLens": dual, "kBlock": k_block}
        out["pointMeta"] = meta
        out["freqs"] = freqs
    return out


def load(path, k_block=64):
    """build() of an openEMS XML file."""
    with open(path) as f:
        return build(f.read(), k_block=k_block)


def run(path, steps=None, storage="fp32", out=None, freqs=None):
    """Run an openEMS XML file on the HIP engine: build it, step nTS (or
    `steps`) steps, and return the probe traces, port S11 when the file
    has a ut/it probe pair (names ending in ut<k> / it<k>), and Dmax at
    each NF2FF frequency.  With `out`, the result is also written there as
    JSON."""
    import json
    from . import farfield, ports
    from .engine import HIPEngine
    w = load(path)
    K = int(steps or w["nTS"])
    g = HIPEngine()
    if not g.initGPU():
        raise RuntimeError("no HIP device")
    g.init(w["gridSize"], w["coefficients"], storage=storage)
    if "pmlRegions" in w:
        g.configurePML(w["pmlRegions"])
    if "mur" in w:
        g.configureMur(w["mur"])
    g.configureExcitation(w["excitation"])
    idx, slices = ports.build_gather_indices(w["probes"], w["gridSize"])
    if len(idx):
        g.configureProbeGatherBuffered(idx, K, 1)
    if "nf2ff" in w:
        g.configureNF2FFAccumulation(w["nf2ff"])
    g.iterate(K)
    g.sync()
    res = {"gridSize": w["gridSize"], "dT": w["dT"], "steps": K, "probes": {}}
    traces = {}
    if len(idx):
        r = g.readProbeGatherBuffered()
        tr = ports.probe_traces(r["data"], r["numSamples"], r["stride"], w["probes"], slices, 1, w["dT"])
        for p, (t, v) in zip(w["probes"], tr):
            traces[p["name"]] = (t, v)
            res["probes"][p["name"]] = {"time_s": t.tolist(), "value": v.tolist()}
    f = np.asarray(freqs if freqs is not None else np.linspace(max(w["f0"] - w["fc"], 0.0) or w["fc"] / 100,
                                                                 w["f0"] + w["fc"], 201), np.float64)
    res["s11"] = {}
    for name, (t, v) in traces.items():
        if "ut" in name:
            other = name.replace("ut", "it", 1)
            if other in traces:
                s = ports.s11_at((t, v), traces[other], f)
                res["s11"][name.replace("ut", "", 1) or "port"] = {"freq_Hz": f.tolist(),
                                                                    "s11_dB": s["s11_dB"].tolist()}
    if "nf2ff" in w:
        theta, phi = farfield.sphere_grid()
        acc = g.readNF2FFAccumulation()
        res["Dmax_dBi"] = []
        for k, fr in enumerate(w["freqs"]):
            ff = g.computeNF2FFfarField({"pointMeta": w["pointMeta"], "theta": theta, "phi": phi,
                                         "center": [0, 0, 0], "frequency": float(fr), "radius": 1.0,
                                         "freqIndex": k})
            p_rad = farfield.rad_power(np.asarray(acc["accumE"][k]), np.asarray(acc["accumH"][k]),
                                       w["pointMeta"], w["nf2ff"]["numPoints"])
            res["Dmax_dBi"].append(farfield.directivity(ff, len(theta), len(phi), p_rad)["DmaxdBi"])
    g.destroy()
    if out:
        with open(out, "w") as fh:
            json.dump(res, fh)
    return res


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Run an openEMS XML simulation on the MI355X engine.")
    ap.add_argument("xml")
    ap.add_argument("--steps", type=int, default=None, help="time steps (default: NumberOfTimesteps)")
    ap.add_argument("--storage", choices=["fp32", "fp16"], default="fp32")
    ap.add_argument("--out", default=None, help="write probe traces, S11 and Dmax as JSON")
    a = ap.parse_args(argv)
    res = run(a.xml, a.steps, a.storage, a.out)
    for name, s in res["s11"].items():
        k = int(np.argmin(s["s11_dB"]))
        print("port %s: min S11 %.2f dB at %.4g GHz" % (name, s["s11_dB"][k], s["freq_Hz"][k] / 1e9))
    for k, d in enumerate(res.get("Dmax_dBi", [])):
        print("Dmax %.2f dBi (NF2FF frequency %d)" % (d, k))


if __name__ == "__main__":
    main()
