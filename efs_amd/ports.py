"""Probe -> port -> S11 chain over the engine's probe gather.

FP64 numpy restatement of the host-side post-processing that turns the
gathered edge values into port voltage/current traces and then into port
quantities:

  build_gather_indices     app/sim-worker.js:225-283  (gather index list,
                           one slice per probe; voltage probes sum edges
                           along the probe line, current probes walk the
                           Ampere loop on dual lines, per normal direction)
  compute_probe_integrals  app/sim-worker.js:285-301
  probe_traces             app/sim-worker.js:926-937 (gather sample s is
                           step (s+1)*interval; current probes sit half a
                           step later)
  dft_time2freq            src/analysis.mjs:43-61 (openEMS DFT_time2freq:
                           rectangle rule with the local sample spacing)
  calc_port                src/ports.mjs:154-205 (incident/reflected waves
                           and powers against Z_ref)
  compute_s11              scripts/lib/post-process.mjs:58-124 (Z = U/I,
                           Gamma = (Z - Z0)/(Z + Z0), in dB)

A probe is a dict {type: 0 voltage | 1 current, start, stop (node or dual
line indices), weight, and for current probes normDir, startInside,
stopInside} -- the fields of the reference's probe records
(getProbeInfo, src/embind_api.cpp:921-980).  Field probes (openEMS ProbeBox
types 2 and 3: E and H at a mesh node) carry their interpolation as taps
[[component, edge component, x, y, z, weight], ...] over volt (type 2) or
curr (type 3); their trace is the field vector per sample (field_taps).
"""
import numpy as np


def build_gather_indices(probes, grid_size):
    """Flat gather indices (NIJK, n*N + x*Ny*Nz + y*Nz + z) and one slice
    record per probe."""
    nx, ny, nz = (int(v) for v in grid_size)
    N = nx * ny * nz

    def lin(n, xyz):
        return n * N + int(xyz[0]) * ny * nz + int(xyz[1]) * nz + int(xyz[2])

    indices, slices = [], []
    for p in probes:
        off = len(indices)
        s, t = [int(v) for v in p["start"]], [int(v) for v in p["stop"]]
        if p["type"] in (2, 3):
            comp, wts = [], []
            for c, n, x, y, z, wt in p["taps"]:
                indices.append(lin(int(n), (x, y, z)))
                comp.append(int(c))
                wts.append(float(wt))
            slices.append({"offset": off, "count": len(indices) - off,
                           "type": "efield" if p["type"] == 2 else "hfield",
                           "comp": np.array(comp, np.int64), "w": np.array(wts)})
            continue
        if p["type"] == 0:
            signs = []
            for n in range(3):
                lo, hi = min(s[n], t[n]), max(s[n], t[n])
                if lo == hi:
                    continue
                sgn = 1.0 if s[n] < t[n] else -1.0
                for k in range(lo, hi):
                    xyz = list(s)
                    xyz[n] = k
                    indices.append(lin(n, xyz))
                    signs.append(sgn)
            slices.append({"offset": off, "count": len(indices) - off, "type": "voltage",
                           "signs": np.array(signs)})
            continue
        si = [bool(v) for v in p.get("startInside", (True, True, True))]
        so = [bool(v) for v in p.get("stopInside", (True, True, True))]
        edges = []
        # Ampere loop around normDir, counter-clockwise seen from +normDir:
        # four legs of (component, fixed corner, sign, keep); a leg runs its
        # component from s+1 to t and is kept only when the dual lines it
        # uses lie inside the grid.
        for comp, corner, sign, keep in _current_legs(int(p["normDir"]), s, t, si, so):
            if not keep:
                continue
            for k in range(s[comp] + 1, t[comp] + 1):
                xyz = list(corner)
                xyz[comp] = k
                edges.append((comp, xyz, sign))
        for n, xyz, sign in edges:
            indices.append(lin(n, xyz))
        slices.append({"offset": off, "count": len(indices) - off, "type": "current",
                       "signs": np.array([e[2] for e in edges])})
    return np.array(indices, dtype=np.uint32), slices


def _current_legs(d, s, t, si, so):
    """The four legs of a current probe's loop for normal direction d, as
    the reference's three cases list them (app/sim-worker.js:244-272)."""
    if d == 0:
        return ((1, (t[0], 0, s[2]), 1.0, so[0] and si[2]), (2, (t[0], t[1], 0), 1.0, so[0] and so[1]),
                (1, (s[0], 0, t[2]), -1.0, si[0] and so[2]), (2, (s[0], s[1], 0), -1.0, si[0] and si[1]))
    if d == 1:
        return ((2, (s[0], s[1], 0), 1.0, si[0] and si[1]), (0, (0, t[1], t[2]), 1.0, so[1] and so[2]),
                (2, (t[0], t[1], 0), -1.0, so[0] and so[1]), (0, (0, s[1], s[2]), -1.0, si[1] and si[2]))
    if d == 2:
        return ((0, (0, s[1], s[2]), 1.0, si[1] and si[2]), (1, (t[0], 0, s[2]), 1.0, so[0] and si[2]),
                (0, (0, t[1], t[2]), -1.0, so[1] and so[2]), (1, (s[0], 0, t[2]), -1.0, si[0] and so[2]))
    raise ValueError("current probe normDir must be 0, 1 or 2 (got %r)" % (d,))


def compute_probe_integrals(gathered, probes, slices):
    """One value per probe from one gather sample (interleaved volt, curr
    pairs per index): voltage probes sum volt, current probes sum curr, with
    the edge signs, times the probe weight; field probes give their
    magnitude (field_values has the vector)."""
    g = np.asarray(gathered, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(len(probes))
    for k, (p, sl) in enumerate(zip(probes, slices)):
        if sl["type"] in ("efield", "hfield"):
            out[k] = float(np.linalg.norm(field_values(g, p, sl)))
            continue
        col = 0 if sl["type"] == "voltage" else 1
        vals = g[sl["offset"]:sl["offset"] + sl["count"], col]
        out[k] = float(np.sum(vals * sl["signs"])) * float(p.get("weight", 1.0))
    return out


def field_values(g, p, sl):
    """The field vector of a field probe from one gather sample (g: the
    (volt, curr) pairs per gathered index)."""
    col = 0 if sl["type"] == "efield" else 1
    vals = g[sl["offset"]:sl["offset"] + sl["count"], col] * sl["w"]
    return np.bincount(sl["comp"], weights=vals, minlength=3)[:3] * float(p.get("weight", 1.0))


def probe_traces(data, num_samples, stride, probes, slices, sample_interval, dT):
    """(time, value) per probe from a buffered gather readback; a field
    probe's value is its [num_samples, 3] field vector, sampled on the
    primal (E) or dual (H, half a step later) time grid."""
    g = np.asarray(data, dtype=np.float64)[:num_samples * stride].reshape(num_samples, stride // 2, 2)
    steps = (np.arange(num_samples) + 1) * sample_interval
    out = []
    for p, sl in zip(probes, slices):
        half = 0.5 if p["type"] in (1, 3) else 0.0
        wt = float(p.get("weight", 1.0))
        if sl["type"] in ("efield", "hfield"):
            col = 0 if sl["type"] == "efield" else 1
            vals = g[:, sl["offset"]:sl["offset"] + sl["count"], col] * sl["w"]
            v = np.stack([vals[:, sl["comp"] == c].sum(axis=1) for c in range(3)], axis=1) * wt
        else:
            col = 0 if sl["type"] == "voltage" else 1
            v = (g[:, sl["offset"]:sl["offset"] + sl["count"], col] * sl["signs"]).sum(axis=1) * wt
        out.append(((steps + half) * dT, v))
    return out


def field_taps(mesh, kind, node, interpolation="node"):
    """Taps of an E (kind 2) or H (kind 3) point probe at mesh node `node`
    as openEMS's field probes read them (Engine_Interface_FDTD: the E field
    of an edge is volt / edge length, the H field curr / dual edge length).
    interpolation "node" (openEMS NODE_INTERPOLATE): E_n is the mean of the
    two n-edges meeting at the node, weighted by the other edge's length
    (the lower-boundary node takes its own edge, the last node 0); H_n the
    mean of the four H_n around the node in the plane across n (0 on the
    grid's outer planes).  "none": the field of the edge at the node."""
    N = mesh.shape
    p = [int(v) for v in node]
    taps = []
    for n in range(3):
        if kind == 2:
            d = mesh.primal[n][p[n]]
            if interpolation == "none":
                if d > 0:
                    taps.append([n, n, *p, 1.0 / d])
                continue
            if p[n] == N[n] - 1 or d == 0:
                continue
            if p[n] == 0:
                taps.append([n, n, *p, 1.0 / d])
                continue
            q = list(p)
            q[n] -= 1
            dd = mesh.primal[n][q[n]]
            r = d / (d + dd)
            taps.append([n, n, *p, (1.0 - r) / d])
            taps.append([n, n, *q, r / dd])
        else:
            nP, nPP = (n + 1) % 3, (n + 2) % 3
            if interpolation == "none":
                d = mesh.dual[n][p[n]]
                if d > 0:
                    taps.append([n, n, *p, 1.0 / d])
                continue
            if any(p[m] == N[m] - 1 for m in range(3)) or p[nP] == 0 or p[nPP] == 0:
                continue
            for a, b in ((0, 0), (1, 0), (1, 1), (0, 1)):
                q = list(p)
                q[nP] -= a
                q[nPP] -= b
                taps.append([n, n, *q, 0.25 / mesh.dual[n][q[n]]])
    return taps


def dft_time2freq(time, values, freq, signal_type="pulse"):
    """sum_n v_n exp(-j w t_n) dt_n with dt_n the spacing to the previous
    sample (the first uses t1 - t0); 'periodic' drops the dt weight."""
    t = np.asarray(time, dtype=np.float64)
    v = np.asarray(values, dtype=np.float64)
    f = np.atleast_1d(np.asarray(freq, dtype=np.float64))
    if signal_type == "periodic":
        w = np.ones_like(t)
    else:
        w = np.empty_like(t)
        if len(t) > 1:
            w[1:] = np.diff(t)
            w[0] = t[1] - t[0]
        elif len(t):
            w[0] = 1.0
    ph = 2 * np.pi * f[:, None] * t[None, :]
    vw = (v * w)[None, :]
    return np.sum(vw * np.cos(ph), axis=1), -np.sum(vw * np.sin(ph), axis=1)


def spectrum(trace, freq):
    """Complex DFT of a (time, value) trace at the given frequencies."""
    re, im = dft_time2freq(trace[0], trace[1], freq)
    return re + 1j * im


def port_s11(ut, it, freq, z_ref=50.0):
    """S11 = uf_ref / uf_inc of a port from its voltage and current traces."""
    p = calc_port(spectrum(ut, freq), spectrum(it, freq), z_ref)
    return p["uf_ref"] / p["uf_inc"]


def calc_port(uf_tot, if_tot, z_ref):
    """Incident/reflected voltage and current waves and powers of a port
    from its total frequency-domain voltage and current (complex arrays)."""
    if z_ref is None or not z_ref > 0:
        raise ValueError("Port Z_ref must be positive (got %r). Use R > 0 for lumped ports." % (z_ref,))
    u = np.asarray(uf_tot, dtype=np.complex128)
    i = np.asarray(if_tot, dtype=np.complex128)
    uf_inc = 0.5 * (u + i * z_ref)
    if_inc = 0.5 * (i + u / z_ref)
    uf_ref = u - uf_inc
    if_ref = if_inc - i

    def p(a, b):
        return 0.5 * (a.real * b.real + a.imag * b.imag)
    return {"uf_inc": uf_inc, "uf_ref": uf_ref, "if_inc": if_inc, "if_ref": if_ref,
            "P_inc": p(uf_inc, if_inc), "P_ref": p(uf_ref, if_ref), "P_acc": p(u, i)}


def compute_s11(ut, it, f_max=0.0, nf=201, z0=50.0):
    """S11 (dB) and input impedance on nf points from 0 to f_max (0: the
    Nyquist rate of the voltage trace).  ut, it: (time, values)."""
    tu = np.asarray(ut[0], dtype=np.float64)
    if len(tu) < 10 or len(it[0]) < 10:
        return None
    dt = tu[1] - tu[0]
    fmax = f_max if f_max > 0 else 0.5 / dt
    freq = np.arange(nf) / (nf - 1) * fmax
    return s11_at(ut, it, freq, z0)


def s11_at(ut, it, freq, z0=50.0):
    """S11 and Z at the given frequencies (the per-frequency body of
    compute_s11): Z = U/I, Gamma = (Z - z0)/(Z + z0), |Gamma| in dB floored
    at 1e-15, and 0 dB / Z = 0 where |I|^2 < 1e-30."""
    freq = np.asarray(freq, dtype=np.float64)
    ur, ui = dft_time2freq(ut[0], ut[1], freq)
    ir, ii = dft_time2freq(it[0], it[1], freq)
    U, I = ur + 1j * ui, ir + 1j * ii
    ok = (I.real ** 2 + I.imag ** 2) >= 1e-30
    Z = np.where(ok, U / np.where(ok, I, 1.0), 0.0)
    gamma = (Z - z0) / (Z + z0)
    s11 = np.where(ok, 20 * np.log10(np.maximum(np.abs(gamma), 1e-15)), 0.0)
    return {"freq": freq, "s11_dB": s11, "Z": Z}
