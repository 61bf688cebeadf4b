"""Seeded synthetic cases shared by the parity tests: the same dictionary
configures the oracle (CPUFDTDEngine method names) and the HIP engine
(WebGPUEngine method names)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def f32(rng, n, lo, hi):
    return rng.uniform(lo, hi, n).astype(np.float32)


def random_coeffs(rng, N, quantize=0):
    """Uniform random coefficients; with quantize=k each array takes only k
    distinct values (the case the engine's coefficient tables compress)."""
    t = 3 * N

    def arr(lo, hi):
        if not quantize:
            return f32(rng, t, lo, hi)
        levels = f32(rng, quantize, lo, hi)
        return levels[rng.integers(0, quantize, t)]
    return {"vv": arr(0.9, 1.0), "vi": arr(0.05, 0.3), "ii": arr(0.9, 1.0), "iv": arr(0.05, 0.3)}


def shell_regions(dims, w):
    """Six non-overlapping UPML boxes covering a boundary shell of width w
    (x slabs full, y slabs inside x, z slabs inside x and y) -- the layout
    openEMS uses for PML on all six faces."""
    nx, ny, nz = dims
    regs = []
    wx, wy, wz = (min(w, d // 2) for d in dims)
    if wx:
        regs += [([0, 0, 0], [wx, ny, nz]), ([nx - wx, 0, 0], [wx, ny, nz])]
    ix0, ix1 = wx, nx - wx
    if wy and ix1 > ix0:
        regs += [([ix0, 0, 0], [ix1 - ix0, wy, nz]), ([ix0, ny - wy, 0], [ix1 - ix0, wy, nz])]
    iy0, iy1 = wy, ny - wy
    if wz and ix1 > ix0 and iy1 > iy0:
        regs += [([ix0, iy0, 0], [ix1 - ix0, iy1 - iy0, wz]), ([ix0, iy0, nz - wz], [ix1 - ix0, iy1 - iy0, wz])]
    return regs


def random_pml(rng, regions, quantize=0):
    """Random UPML coefficients; with quantize=k each region array takes
    only k distinct values (the case the engine's UPML tables compress)."""
    out = []
    for start, num in regions:
        t = 3 * int(np.prod(num))

        def arr(lo, hi):
            if not quantize:
                return f32(rng, t, lo, hi)
            return f32(rng, quantize, lo, hi)[rng.integers(0, quantize, t)]
        out.append({"startPos": list(start), "numLines": list(num),
                    "vv": arr(0.6, 1.0), "vvfo": arr(0.0, 0.3), "vvfn": arr(0.8, 1.4),
                    "ii": arr(0.6, 1.0), "iifo": arr(0.0, 0.3), "iifn": arr(0.8, 1.4)})
    return out


def lin(dims, x, y, z):
    return x * dims[1] * dims[2] + y * dims[2] + z


def random_case(seed, dims, pml_width=0, exc=True, mur=False, ade=False, tfsf=False, rlc=False, fields=True,
                quantize=0, pml_quantize=0, pbc=None):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    N = nx * ny * nz
    case = {"dims": list(dims), "coeffs": random_coeffs(rng, N, quantize)}
    if fields:
        case["volt0"] = f32(rng, 3 * N, -0.5, 0.5)
        case["curr0"] = f32(rng, 3 * N, -0.5, 0.5)
    if pml_width:
        case["pml"] = random_pml(rng, shell_regions(dims, pml_width), pml_quantize)
    if exc:
        k = 4
        case["excitation"] = {"signal": f32(rng, 17, -1, 1), "delay": rng.integers(0, 6, k).astype(np.uint32),
                              "amp": f32(rng, k, -1, 1), "dir": rng.integers(0, 3, k).astype(np.uint32),
                              "pos": rng.integers(0, N, k).astype(np.uint32), "period": 0}
    if mur:
        n_idx, s_idx, co = [], [], []
        for y in range(ny):
            for z in range(nz):
                for n in (1, 2):
                    n_idx.append(n * N + lin(dims, nx - 1, y, z))
                    s_idx.append(n * N + lin(dims, nx - 2, y, z))
                    co.append(rng.uniform(0.2, 0.5))
        case["mur"] = {"coeff": np.array(co, np.float32), "normal_idx": np.array(n_idx, np.uint32),
                       "shifted_idx": np.array(s_idx, np.uint32)}
    if ade:
        m = 12
        pos = rng.integers(0, N, m).astype(np.uint32)

        def d(dr):
            return {"dir": dr, "pos_idx": pos, "v_int_ADE": f32(rng, m, 0.5, 0.95), "v_ext_ADE": f32(rng, m, 0.05, 0.3),
                    "v_Lor_ADE": f32(rng, m, 0, 0.2), "i_int_ADE": f32(rng, m, 0.5, 0.95),
                    "i_ext_ADE": f32(rng, m, 0.05, 0.3), "i_Lor_ADE": f32(rng, m, 0, 0.2)}
        case["ade"] = {"orders": [{"numCells": m, "hasLorentz": False, "directions": [d(0), d(2)]},
                                  {"numCells": m, "hasLorentz": True, "directions": [d(1)]}]}
    if tfsf:
        def pts(k):
            return [{"field_idx": int(rng.integers(0, 3 * N)), "delay_int": int(rng.integers(0, 4)),
                     "delay_frac": float(np.float32(rng.uniform())), "amp": float(np.float32(rng.uniform(0.5, 1.5)))}
                    for _ in range(k)]
        case["tfsf"] = {"signal": f32(rng, 14, -1, 1), "period": 0, "voltagePoints": pts(5), "currentPoints": pts(4)}
    if rlc:
        def r():
            return float(np.float32(rng.uniform()))
        case["rlc"] = {"elements": [
            {"field_idx": int(rng.integers(0, N)), "direction": 2, "type_flag": 0, "i2v": r(), "ilv": r(),
             "vvd": 0, "vv2": 0, "vj1": 0, "vj2": 0, "ib0": 0, "b1": 0, "b2": 0},
            {"field_idx": int(rng.integers(0, N)), "direction": 1, "type_flag": 1, "i2v": 0, "ilv": 0,
             "vvd": 0.5 + 0.4 * r(), "vv2": 0.2 * r(), "vj1": 0.1 * r(), "vj2": 0.1 * r(), "ib0": 0.3 * r(),
             "b1": 0.2 * r(), "b2": 0.1 * r()}]}
    if pbc:  # [(axis, phase), ...]
        case["pbc"] = {"axes": [{"axis": a, "phase": p} for a, p in pbc]}
    return case


def configure(eng, case, ade_apply=False):
    """Apply a case to either engine (oracle or HIP)."""
    if "volt0" in case:
        eng.uploadFields(case["volt0"], case["curr0"])
    if "pml" in case:
        eng.configurePML(case["pml"])
    if "excitation" in case:
        eng.configureExcitation(case["excitation"])
    if "mur" in case:
        eng.configureMur(case["mur"])
    if "ade" in case:
        if hasattr(eng, "configureLorentzADE"):
            eng.configureLorentzADE(case["ade"], apply=ade_apply)
        else:
            eng.configureLorentz(case["ade"], apply=ade_apply)
    if "tfsf" in case:
        eng.configureTFSF(case["tfsf"])
    if "rlc" in case:
        eng.configureRLC(case["rlc"])
    if "pbc" in case:
        eng.configurePBC(case["pbc"])


def golden_cases():
    with open(os.path.join(HERE, "golden", "reference_step_vectors.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
