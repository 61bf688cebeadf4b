"""Synthetic workloads built with the FP64 operator builder.

``stress_grid`` is the plain-FP32 stress case: an n^3 grid of 1 mm cells
with 8-cell UPML on all six faces, 40 seeded dielectric boxes (eps_r in
U(1,10), kappa in U(0,0.05) S/m), 10 PEC plates, a 50-ohm lumped port fed
by a Gaussian pulse (f0 = 5 GHz, fc = 4 GHz), field energy every 100 steps
and a six-face NF2FF box at 16 frequencies from 1 to 9 GHz.  ``probes``
holds the port's voltage and current probes (efs_amd.ports format).  n = 256 is the
full-size workload; smaller n keep the same structure for tests.
"""
import numpy as np

from . import meshing, ports
from . import operator as op

STRESS_SEED = 20261203 + 2


def stress_grid(n=256, seed=STRESS_SEED, pml=8, nfreq=16, n_boxes=40, n_plates=10, k_block=64):
    mesh = op.uniform_mesh((n, n, n), 1e-3)
    dT = mesh.timestep()
    mat = op.Materials(mesh)
    rng = np.random.default_rng(seed)
    lo_lim, hi_lim = pml + 1, n - pml - 1
    span = hi_lim - lo_lim
    c = n // 2
    for _ in range(n_boxes):
        size = rng.integers(max(span // 16, 1), max(span // 4, 2), 3)
        lo = np.array([rng.integers(lo_lim, max(hi_lim - s, lo_lim + 1)) for s in size])
        mat.box(lo, lo + size, eps=rng.uniform(1.0, 10.0), kappa=rng.uniform(0.0, 0.05))
    plates = 0
    while plates < n_plates:
        ax = int(rng.integers(0, 3))
        size = rng.integers(max(span // 10, 1), max(span // 4, 2), 3)
        lo = np.array([rng.integers(lo_lim, max(hi_lim - s, lo_lim + 1)) for s in size])
        hi = lo + size
        lo[ax] = hi[ax] = int(rng.integers(lo_lim, hi_lim))
        if all(lo[m] - 2 <= c <= hi[m] + 2 for m in range(3)):
            continue  # keep the port clear of metal
        mat.pec_box(lo, hi)
        plates += 1
    op.pec_boundary(mat)
    # lumped port: 2 z-edges in series at the centre, 50 ohm, 0.5 V per edge
    port, probes = op.lumped_port(mesh, mat, 2, (c, c, c - 1), (c, c, c + 1), 50.0, excite=500.0)
    coeffs = op.build_coefficients(mesh, mat, dT)
    regions = op.build_upml(mesh, coeffs, [pml] * 6, dT)
    signal = op.gaussian_pulse(5e9, 4e9, dT)
    excitation = {"signal": signal, "delay": np.zeros(len(port["pos"]), np.uint32), "amp": port["amp"],
                  "dir": port["dir"], "pos": port["pos"], "period": 0}
    pts, meta, primal, dual = op.nf2ff_box(mesh, [pml + 1] * 3, [n - pml - 2] * 3)
    freqs = np.linspace(1e9, 9e9, nfreq)
    nf2ff = {"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": list(2 * np.pi * freqs), "dT": dT,
             "primalEdgeLens": primal, "dualEdgeLens": dual, "kBlock": k_block}
    n_pml = sum(int(np.prod(r["numLines"])) for r in regions)
    return {"gridSize": [n, n, n], "coefficients": op.to_f32(coeffs), "pmlRegions": regions,
            "excitation": excitation, "nf2ff": nf2ff, "pointMeta": meta, "freqs": freqs, "dT": dT,
            "seed": seed, "n_pml_cells": n_pml, "pml": pml, "probes": probes}


def configure_engine(eng, w, nf2ff=True, storage=None):
    """Configure a WebGPUEngine-style engine with a workload dictionary
    (storage: field storage of the HIP engine, "fp32" or "fp16")."""
    if storage is None:
        eng.init(w["gridSize"], w["coefficients"])
    else:
        eng.init(w["gridSize"], w["coefficients"], storage=storage)
    eng.configurePML(w["pmlRegions"])
    eng.configureExcitation(w["excitation"])
    if nf2ff:
        eng.configureNF2FFAccumulation(w["nf2ff"])


PATCH_SEED = 20261203 + 1


def _at(lines, v):
    """Index of the mesh line at coordinate v (lines placed there exactly)."""
    k = int(np.argmin(np.abs(lines - v)))
    if abs(lines[k] - v) > 1e-9 * max(1.0, abs(v)):
        raise ValueError("no mesh line at %g" % v)
    return k


def patch_antenna(patch_w=32.0, patch_l=40.0, f0=3e9, fc=3e9, eps_r=3.38, sub_h=1.524, sub_w=60.0, sub_l=60.0,
                  feed_x=-6.0, feed_r=50.0, sim_box=(200.0, 200.0, 150.0), nf2ff_freq=2.4e9, k_block=64,
                  n_ts=30000):
    """Microstrip patch antenna of the reference's script example
    (app/examples.mjs:17-103, the openEMS Simple_Patch_Antenna tutorial):
    a patch_w x patch_l mm PEC patch on a sub_h mm substrate (eps_r,
    loss tangent 1e-3 at 2.45 GHz) over a sub_w x sub_l mm ground plane,
    fed by a 50-ohm lumped port at x = feed_x, Gaussian excitation (f0, fc),
    first-order Mur on all six faces and an NF2FF box 3 lines inside the
    boundary at nf2ff_freq.  Lengths in mm.  The mesh is the example's:
    box, patch, ground and substrate lines smoothed to lambda/20 at f0+fc
    with ratio 1.4, then the feed lines added exactly (86 x 87 x 71 for the
    defaults).  dT is the CFL limit of the smallest cell."""
    unit = 1e-3
    res = op.C0 / (f0 + fc) / unit / 20
    hw, hl, hpw, hpl = sub_w / 2, sub_l / 2, patch_w / 2, patch_l / 2
    xs = meshing.grid_lines([-sim_box[0] / 2, sim_box[0] / 2, -hpw, hpw, -hw, hw], res, 1.4, extra=[feed_x, 0.0])
    ys = meshing.grid_lines([-sim_box[1] / 2, sim_box[1] / 2, -hpl, hpl, -hl, hl], res, 1.4, extra=[0.0])
    zs = meshing.grid_lines([-sim_box[2] / 3, sim_box[2] * 2 / 3] + [sub_h * i / 4 for i in range(5)], res, 1.4)
    mesh = op.Mesh(xs * unit, ys * unit, zs * unit)
    dT = mesh.timestep()
    mat = op.Materials(mesh)
    kappa = 1e-3 * 2 * np.pi * 2.45e9 * op.EPS0 * eps_r
    ix = lambda v: _at(xs, v)  # noqa: E731
    iy = lambda v: _at(ys, v)  # noqa: E731
    iz = lambda v: _at(zs, v)  # noqa: E731
    mat.box((ix(-hw), iy(-hl), iz(0.0)), (ix(hw), iy(hl), iz(sub_h)), eps=eps_r, kappa=kappa)
    mat.pec_box((ix(-hw), iy(-hl), iz(0.0)), (ix(hw), iy(hl), iz(0.0)))
    mat.pec_box((ix(-hpw), iy(-hpl), iz(sub_h)), (ix(hpw), iy(hpl), iz(sub_h)))
    port, probes = op.lumped_port(mesh, mat, 2, (ix(feed_x), iy(0.0), iz(0.0)), (ix(feed_x), iy(0.0), iz(sub_h)),
                                  feed_r, excite=1.0)
    coeffs = op.build_coefficients(mesh, mat, dT)
    signal = op.gaussian_pulse(f0, fc, dT)
    excitation = {"signal": signal, "delay": np.zeros(len(port["pos"]), np.uint32), "amp": port["amp"],
                  "dir": port["dir"], "pos": port["pos"], "period": 0}
    mur = op.mur_config(mesh, range(6), dT)
    n = mesh.shape
    ins = 3
    pts, meta, primal, dual = op.nf2ff_box(mesh, [ins] * 3, [k - 1 - ins for k in n])
    nf2ff = {"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": [2 * np.pi * nf2ff_freq], "dT": dT,
             "primalEdgeLens": primal, "dualEdgeLens": dual, "kBlock": k_block}
    return {"gridSize": list(n), "coefficients": op.to_f32(coeffs), "mur": mur, "excitation": excitation,
            "nf2ff": nf2ff, "pointMeta": meta, "freqs": np.array([nf2ff_freq]), "dT": dT, "probes": probes,
            "mesh": mesh, "lines_mm": [xs, ys, zs], "nTS": n_ts, "f0": f0, "fc": fc, "patch": [patch_w, patch_l]}


def configure_patch(eng, w, nf2ff=True, gather_interval=1, max_samples=None, storage=None):
    """Configure an engine with a patch workload: coefficients (engines
    with an init method; the CPU oracle takes them at construction), Mur,
    excitation, the port probes' buffered gather and optionally the NF2FF
    accumulation.  Returns the gather slices."""
    if hasattr(eng, "init"):
        if storage is None:
            eng.init(w["gridSize"], w["coefficients"])
        else:
            eng.init(w["gridSize"], w["coefficients"], storage=storage)
    eng.configureMur(w["mur"])
    eng.configureExcitation(w["excitation"])
    idx, slices = ports.build_gather_indices(w["probes"], w["gridSize"])
    n = max_samples if max_samples is not None else w["nTS"] // gather_interval
    eng.configureProbeGatherBuffered(idx, n, gather_interval)
    if nf2ff:
        eng.configureNF2FFAccumulation(w["nf2ff"])
    return slices


BATCH_SEED = 20261203 + 4


def patch_sweep_params(n=64, seed=BATCH_SEED, w_lo=20.0, w_hi=50.0):
    """Patch widths of a design sweep (uniform in [w_lo, w_hi] mm from the
    seed); the other lengths scale with the width as in the 32 mm design
    (patch length 1.25 W, substrate 1.875 W square, feed at -0.1875 W) and
    the simulation box grows with it, which gives graded meshes from about
    60 x 60 x 50 (W = 20) to 160 x 150 x 90 cells (W = 50)."""
    rng = np.random.default_rng(seed)
    out = []
    for W in rng.uniform(w_lo, w_hi, n):
        W = float(W)
        out.append({"patch_w": W, "patch_l": 1.25 * W, "sub_w": 1.875 * W, "sub_l": 1.875 * W,
                    "feed_x": -0.1875 * W, "sim_box": (8.33 * W - 31.7, 7.43 * W - 10.7, 3.35 * W + 30.5)})
    return out


def add_patch(batch, w):
    """Add a patch workload to a BatchEngine as one instance: coefficients,
    Mur, excitation and the port probes.  Returns (instance id, slices)."""
    i = batch.add(w["gridSize"], w["coefficients"])
    batch.configureMur(i, w["mur"])
    batch.configureExcitation(i, w["excitation"])
    idx, slices = ports.build_gather_indices(w["probes"], w["gridSize"])
    batch.configureProbeGather(i, idx)
    return i, slices


DISPERSIVE_SEED = 20261203 + 3


DRUDE = {"eps_inf": 1.0, "wp": 2 * np.pi * 12e9, "w0": 0.0, "gamma": 1e9}
LORENTZ = {"eps_inf": 2.0, "wp": 2 * np.pi * 6e9 * np.sqrt(3.0), "w0": 2 * np.pi * 6e9, "gamma": 1e9}


def dispersive_scatter(n=192, seed=DISPERSIVE_SEED, pml=8, sphere="drude", slab=True, f0=6e9, fc=5e9,
                       rcs_freqs=(3e9, 5e9, 7e9, 9e9), n_angles=64, n_ts=15000, k_block=64, radius=None):
    """Dispersive scatterers under a plane wave: an n^3 grid of 1 mm cells
    with UPML on all six faces, a Drude sphere (plasma frequency 2 pi 12 GHz,
    damping 1e9 1/s; sphere="lorentz" makes it of the slab's material, None
    leaves it out) whose centre and radius come from the seed, next to a
    single-pole Lorentz slab (eps_inf 2, delta eps 3, pole 6 GHz, damping
    1e9 1/s; the slab is 1/10 of the grid thick along x), both in ADE form
    with the openEMS write-back; a total-field / scattered-field plane wave
    along +x with E along z (Gaussian f0, fc); and an NF2FF box in the
    scattered-field shell for the bistatic RCS in the xy plane (n_angles
    angles from phi = 0, the forward direction) at rcs_freqs.
    Layout along each axis: UPML [0, pml), NF2FF box at pml + 4, total-field
    box from pml + 8 to n - 1 - (pml + 8).  dT is the CFL limit reduced for
    the poles (Mesh.timestep(wp_max=...))."""
    mesh = op.uniform_mesh((n, n, n), 1e-3)
    mats = ([DRUDE if sphere == "drude" else LORENTZ] if sphere else []) + ([LORENTZ] if slab else [])
    dT = mesh.timestep(wp_max=max([m["wp"] for m in mats] + [0.0]))
    mat = op.Materials(mesh)
    rng = np.random.default_rng(seed)
    nf_lo, tf_lo = pml + 4, pml + 8
    tf_hi = n - 1 - tf_lo
    span = tf_hi - tf_lo
    cells = np.indices(mat.eps.shape, dtype=np.float64) + 0.5  # cell centres in cell units
    poles = []
    info = {}
    if sphere:
        r = radius if radius is not None else float(rng.uniform(0.10, 0.14) * span)
        cx = tf_lo + span * float(rng.uniform(0.28, 0.34))
        cy, cz = (tf_lo + span * float(rng.uniform(0.45, 0.55)) for _ in range(2))
        m = (cells[0] - cx) ** 2 + (cells[1] - cy) ** 2 + (cells[2] - cz) ** 2 <= r * r
        p = DRUDE if sphere == "drude" else LORENTZ
        mat.eps[m] = p["eps_inf"]
        poles.append({"mask": m, **{k: p[k] for k in ("wp", "w0", "gamma")}})
        info["sphere"] = {"centre": [cx, cy, cz], "radius": r, "cells": int(m.sum()), "material": dict(p)}
    if slab:
        t = max(2, n // 10)
        x0 = tf_lo + int(0.62 * span)
        m = np.zeros(mat.eps.shape, bool)
        m[x0:x0 + t, tf_lo + span // 5:tf_hi - span // 5, tf_lo + span // 5:tf_hi - span // 5] = True
        mat.eps[m] = LORENTZ["eps_inf"]
        poles.append({"mask": m, **{k: LORENTZ[k] for k in ("wp", "w0", "gamma")}})
        info["slab"] = {"x": [x0, x0 + t], "cells": int(m.sum())}
    op.pec_boundary(mat)
    coeffs = op.build_coefficients(mesh, mat, dT)
    ade = op.ade_orders(mesh, coeffs, dT, poles) if poles else None
    regions = op.build_upml(mesh, coeffs, [pml] * 6, dT)
    signal = op.gaussian_pulse(f0, fc, dT)
    tfsf = op.tfsf_plane_wave(mesh, coeffs, [tf_lo] * 3, [tf_hi] * 3, dT, signal, prop_axis=0, prop_sign=1,
                              e_axis=2)
    pts, meta, primal, dual = op.nf2ff_box(mesh, [nf_lo] * 3, [n - 1 - nf_lo] * 3)
    freqs = np.asarray(rcs_freqs, np.float64)
    nf2ff = {"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": list(2 * np.pi * freqs), "dT": dT,
             "primalEdgeLens": primal, "dualEdgeLens": dual, "kBlock": k_block}
    phi = 2 * np.pi * np.arange(n_angles) / n_angles
    n_pml = sum(int(np.prod(r["numLines"])) for r in regions)
    centre = [float(mesh.lines[d][n // 2]) for d in range(3)]
    return {"gridSize": [n, n, n], "coefficients": op.to_f32(coeffs), "pmlRegions": regions, "ade": ade,
            "tfsf": tfsf, "nf2ff": nf2ff, "pointMeta": meta, "freqs": freqs, "dT": dT, "signal": signal,
            "theta": np.array([np.pi / 2]), "phi": phi, "centre": centre, "n_pml_cells": n_pml, "nTS": n_ts,
            "mesh": mesh, "seed": seed, "info": info}


def configure_dispersive(eng, w, nf2ff=True, storage=None):
    """Configure an engine with a dispersive_scatter workload (engines with
    an init method; the CPU oracle takes the coefficients at construction)."""
    if hasattr(eng, "init"):
        if storage is None:
            eng.init(w["gridSize"], w["coefficients"])
        else:
            eng.init(w["gridSize"], w["coefficients"], storage=storage)
    eng.configurePML(w["pmlRegions"])
    if w["ade"] is not None:
        if hasattr(eng, "configureLorentzADE"):
            eng.configureLorentzADE(w["ade"], apply=True)
        else:
            eng.configureLorentz(w["ade"], apply=True)
    eng.configureTFSF(w["tfsf"])
    if nf2ff:
        eng.configureNF2FFAccumulation(w["nf2ff"])


LARGE_SEED = 20261203 + 5


def large_array(dims=(800, 800, 600), cell=0.5e-3, seed=LARGE_SEED, pml=8, n_boxes=200, pec_fraction=0.25,
                ports=(4, 4), port_pitch=40, arm=14, f0=10e9, fc=8e9, nfreq=32, k_block=64, n_ts=8000, threads=16,
                progress=None):
    """Large mixed-precision workload: a dims grid of `cell` cubes (800 x 800
    x 600 x 0.5 mm = 384 M cells by default) with UPML on all six faces,
    n_boxes seeded boxes (a pec_fraction of them PEC plates, the others
    dielectric with eps_r in U(1, 10) and kappa in U(0, 0.05) S/m), and a
    ports[0] x ports[1] array of centre-fed z dipoles (thin PEC arms of
    `arm` cells each side of a one-edge 50-ohm lumped port: 14.5 mm long, a
    half wave near 10 GHz at 0.5 mm cells; port_pitch cells apart, centred,
    all driven by one Gaussian pulse f0, fc), the ports' voltage and current
    probes, and a six-face NF2FF box at nfreq
    frequencies over the pulse's band (f0 - fc .. f0 + fc).  The boxes keep
    clear of the port array.  Coefficients come from the threaded slab
    builder (the same numbers as the whole-grid one)."""
    say = progress or (lambda *a: None)
    mesh = op.uniform_mesh(dims, cell)
    dT = mesh.timestep()
    mat = op.Materials(mesh)
    rng = np.random.default_rng(seed)
    n = np.array(dims)
    lo_lim, hi_lim = np.full(3, pml + 2), n - pml - 2
    span = hi_lim - lo_lim
    c = n // 2
    pa = np.array([port_pitch * (ports[0] - 1) / 2, port_pitch * (ports[1] - 1) / 2, arm + 2])
    keep_lo, keep_hi = c - pa - 4, c + pa + 4
    placed, pec = 0, 0
    n_pec = int(round(pec_fraction * n_boxes))
    while placed < n_boxes:
        size = np.maximum(rng.integers(span // 40, span // 8 + 1), 1)
        lo = np.array([rng.integers(lo_lim[m], max(hi_lim[m] - size[m], lo_lim[m] + 1)) for m in range(3)])
        hi = lo + size
        if all(lo[m] <= keep_hi[m] and hi[m] >= keep_lo[m] for m in range(3)):
            continue  # keep the port array clear
        if pec < n_pec:
            ax = int(rng.integers(0, 3))
            lo[ax] = hi[ax] = int(rng.integers(lo[ax], hi[ax] + 1))
            mat.pec_box(lo, hi)
            pec += 1
        else:
            mat.box(lo, hi, eps=rng.uniform(1.0, 10.0), kappa=rng.uniform(0.0, 0.05))
        placed += 1
    op.pec_boundary(mat)
    exc = {"pos": [], "dir": [], "amp": []}
    probes = []
    for i in range(ports[0]):
        for j in range(ports[1]):
            px = int(c[0] + (i - (ports[0] - 1) / 2) * port_pitch)
            py = int(c[1] + (j - (ports[1] - 1) / 2) * port_pitch)
            mat.pec_box((px, py, c[2] + 1), (px, py, c[2] + 1 + arm))
            mat.pec_box((px, py, c[2] - arm), (px, py, c[2]))
            port, pr = op.lumped_port(mesh, mat, 2, (px, py, c[2]), (px, py, c[2] + 1), 50.0, excite=1.0)
            for k in exc:
                exc[k].append(port[k])
            for p in pr:
                p["name"] = "%s_%d_%d" % (p["name"], i, j)
            probes += pr
    say("geometry done")
    coeffs = op.build_coefficients_threaded(mesh, mat, dT, threads=threads)
    del mat
    say("coefficients done")
    regions = op.build_upml(mesh, coeffs, [pml] * 6, dT)
    say("upml done")
    coeffs = op.to_f32(coeffs)
    signal = op.gaussian_pulse(f0, fc, dT)
    pos = np.concatenate(exc["pos"])
    excitation = {"signal": signal, "delay": np.zeros(len(pos), np.uint32),
                  "amp": np.concatenate(exc["amp"]).astype(np.float32), "dir": np.concatenate(exc["dir"]),
                  "pos": pos, "period": 0}
    pts, meta, primal, dual = op.nf2ff_box(mesh, [pml + 2] * 3, [k - pml - 3 for k in dims])
    freqs = np.linspace(f0 - fc, f0 + fc, nfreq)
    nf2ff = {"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": list(2 * np.pi * freqs), "dT": dT,
             "primalEdgeLens": primal, "dualEdgeLens": dual, "kBlock": k_block}
    n_pml = sum(int(np.prod(r["numLines"])) for r in regions)
    say("workload done")
    return {"gridSize": list(dims), "coefficients": coeffs, "pmlRegions": regions, "excitation": excitation,
            "probes": probes, "nf2ff": nf2ff, "pointMeta": meta, "freqs": freqs, "dT": dT, "nTS": n_ts,
            "n_pml_cells": n_pml, "ports": ports, "f0": f0, "fc": fc, "seed": seed}


def configure_large(eng, w, nf2ff=True, storage=None, max_samples=None):
    """Configure an engine with a large_array workload (coefficients through
    init for engines that have it), the port probes' buffered gather and
    optionally the NF2FF accumulation; returns the gather slices."""
    if hasattr(eng, "init"):
        if storage is None:
            eng.init(w["gridSize"], w["coefficients"])
        else:
            eng.init(w["gridSize"], w["coefficients"], storage=storage)
    eng.configurePML(w["pmlRegions"])
    eng.configureExcitation(w["excitation"])
    idx, slices = ports.build_gather_indices(w["probes"], w["gridSize"])
    eng.configureProbeGatherBuffered(idx, max_samples if max_samples is not None else w["nTS"], 1)
    if nf2ff:
        eng.configureNF2FFAccumulation(w["nf2ff"])
    return slices
