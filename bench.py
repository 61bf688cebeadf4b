"""Benchmark: FP32 FDTD time-stepping throughput on one MI355X.

Workload (default): plain-FP32 stress grid, 256x256x256 cells of 1 mm, 8-cell
UPML on all six faces, 40 seeded dielectric boxes and 10 PEC plates, a 50-ohm
lumped port driven by a Gaussian pulse (f0 5 GHz, fc 4 GHz), field energy
every 100 steps and a six-face NF2FF box (339,864 surface points) at 16
frequencies in 1-9 GHz.  A "step" is one complete FDTD timestep of that
workload (E sweep, H sweep, UPML, excitation, NF2FF sampling and its DFT fold,
and the periodic energy reduction), with all inputs resident in HBM.

Prints one JSON line.  MCells/s = Nx*Ny*Nz*steps / wall seconds / 1e6.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "MCells/s & %HBM roofline: plain FP32 vs batched/mixed at fixed S11 error"
HBM_PEAK_GBS = 8000.0


def coef_bytes(storage, indexed):
    """Bytes per coefficient in the algorithmic count: 4 (FP32), and for
    FP16 fields with indexed coefficients (byte or 16-bit indices into
    tables, at most 2 B per coefficient) the FP16-coefficient figure, 2."""
    return 2.0 if storage == "fp16" and indexed else 4.0


def sweep_bytes(n_cells, n_pml, storage="fp32", flags=()):
    """Algorithmic HBM bytes of one E (or H) sweep launch.
    Per cell: read V, vv, vi, H (3 components each) and write V = 15 values:
    60 B with FP32 fields; with FP16 fields 9 field values x 2 B + 6
    coefficients = 42 B (FP32 coefficient arrays) or 30 B (indexed
    coefficients, `flags` holding coef_idx or coef_idx16).  Per UPML cell in
    addition: flux read + write and three UPML coefficients for 3 components
    = 60 B."""
    indexed = bool({"coef_idx", "coef_idx16"} & set(flags))
    per_cell = 18.0 + 6 * coef_bytes(storage, indexed) if storage == "fp16" else 60.0
    return per_cell * n_cells + 60.0 * n_pml


def byte_model(storage, flags=()):
    indexed = bool({"coef_idx", "coef_idx16"} & set(flags))
    if storage != "fp16":
        return "60 B/cell (FP32 fields and coefficients) + 60 B/UPML cell"
    return ("30 B/cell (FP16 fields, indexed coefficients counted at 2 B) + 60 B/UPML cell" if indexed
            else "42 B/cell (FP16 fields, FP32 coefficients) + 60 B/UPML cell")


TRAFFIC_FILES = {"fp32": os.path.join(ROOT, "profiles", "fused_sweeps_256cube_stress_traffic.json"),
                 "fp16": os.path.join(ROOT, "profiles", "fused_sweeps_fp16_256cube_stress_traffic.json")}


BATCH_TRAFFIC_FILES = {"fp16": os.path.join(ROOT, "profiles", "batched_sweeps_fp16_patch_design_sweep_traffic.json")}


def sweep_kernel_symbol(flags, side):
    """Name of the kernel that runs the E or H sweep for the path flags."""
    if "vec4" not in flags:
        return f"efs::sweep_{side}1<true>"
    cc = "c" if "coef_idx" in flags else "x" if "coef_idx16" in flags else ""
    if "fused" in flags:
        suffix = "f" + cc + ("p" if "upml_idx" in flags else "")
    elif "upml4" in flags:
        suffix = "m" + cc
    else:
        suffix = cc
    # grids with side work run the twin kernel (NAME_s)
    return (f"efs::sweep_{side}4{suffix}" + ("_h" if "f16" in flags else "") +
            ("_s" if "side_work" in flags else ""))


def measured_traffic(kernel, size, storage="fp32"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement of this workload (None if absent)."""
    path = TRAFFIC_FILES.get(storage)
    if size != 256 or not path or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get(kernel, {}).get("hbm_bytes_per_launch")


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


FP64_NOTE = ("the same engine with FP64 field storage (double fields, UPML flux, NF2FF buffer and accumulators on "
             "the 1-wide sweeps; held to the FP64 oracle at 1e-12 by the GPU tests), same case and steps")


def precision_stress(w, storage, steps):
    """The stress workload at full size on `storage` and on the FP64 storage
    form of the same engine: relative L2 of the fields and the NF2FF
    spectra, relative energy error, after `steps` steps."""
    from efs_amd import workloads
    from efs_amd.engine import HIPEngine
    out = {}
    for st in (storage, "fp64"):
        g = HIPEngine()
        g.initGPU()
        workloads.configure_engine(g, w, storage=st)
        g.iterate(steps)
        f = g.getFields(fp64=st == "fp64")
        acc = g.readNF2FFAccumulation(fp64=st == "fp64")
        out[st] = (f["volt"], f["curr"], g.energy(), acc)
        g.destroy()
    (v, c, e, a), (v64, c64, e64, a64) = out[storage], out["fp64"]
    nf = max(rel_l2(np.concatenate([a["accumE"][k], a["accumH"][k]]),
                    np.concatenate([a64["accumE"][k], a64["accumH"][k]])) for k in range(len(a64["accumE"])))
    return {"against": FP64_NOTE, "grid": w["gridSize"], "steps": steps, "volt_rel_l2": rel_l2(v, v64),
            "curr_rel_l2": rel_l2(c, c64), "energy_rel": abs(e - e64) / abs(e64), "nf2ff_rel_l2_max_freq": nf}


def host_threads():
    """CPU threads this process may use: the CPUs it is allowed to run on,
    capped by OMP_NUM_THREADS when that is set (a GPU host's per-GPU CPU
    share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else n


def host_info():
    """CPU model and counts of the host the CPU baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = None
    return {"cpu_model": model, "cpus_online": os.cpu_count(), "cpus_allowed": allowed,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(w, seconds_target=15.0, what=None):
    """Time the CPU oracle (every thread this process may use, f32 variant)
    on a bounded sample of the same workload: the same grid, coefficients,
    UPML (or Mur) and excitation; NF2FF and energy are not part of the
    sample."""
    from oracle import oracle as orc
    threads = host_threads()
    orc.Oracle.set_threads(threads)
    o = orc.Oracle(w["gridSize"], w["coefficients"])
    if "pmlRegions" in w:
        o.configurePML(w["pmlRegions"])
    if "mur" in w:
        o.configureMur(w["mur"])
    if "excitation" in w:
        o.configureExcitation(w["excitation"])
    if w.get("ade") is not None:
        o.configureLorentz(w["ade"], apply=True)
    if "tfsf" in w:
        o.configureTFSF(w["tfsf"])
    o.iterate(1)  # first touch
    n = int(np.prod(w["gridSize"]))
    steps, elapsed = 0, 0.0
    chunk = 2
    while elapsed < seconds_target:
        t0 = time.perf_counter()
        o.iterate(chunk)
        elapsed += time.perf_counter() - t0
        steps += chunk
        chunk = min(chunk * 2, 64)
    del o
    return {"value": n * steps / elapsed / 1e6, "unit": "MCells/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of " + (what or f"the same {w['gridSize'][0]}^3 UPML+excitation workload") +
                      f" ({elapsed:.1f} s) with the C oracle (float storage, double arithmetic, OpenMP); "
                      f"NF2FF and energy excluded; the oracle is the parity checker, a plain restatement of the "
                      f"reference's CPU engine, not a tuned CPU engine",
            "host": host_info()}




def build_patch_batch(n_inst, storage, max_samples, progress=False):
    """The design-sweep workload: n_inst patch antennas (widths from the
    seed) as instances of one BatchEngine; returns (batch, workloads,
    slices)."""
    from efs_amd import batch, workloads
    b = batch.BatchEngine(storage=storage)
    ws, sl = [], []
    for k, p in enumerate(workloads.patch_sweep_params(n_inst)):
        w = workloads.patch_antenna(**p, n_ts=20000)
        sl.append(workloads.add_patch(b, w)[1])
        w = {key: w[key] for key in ("gridSize", "dT", "probes", "patch", "excitation", "mur", "coefficients")}
        ws.append(w)
        if progress and k % 16 == 15:
            print(f"# built {k + 1}/{n_inst} instances", file=sys.stderr, flush=True)
    b.finalize(1, max_samples)
    return b, ws, sl


def time_batch(b, warmup, K):
    b.iterate(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.iterate(K)
    b.sync()
    return time.perf_counter() - t0


def s11_error(sample, storage, steps, freqs):
    """S11 of `sample` patch instances from the batched engine (storage)
    and from the FP64 storage form of the single-grid engine, over `steps`
    steps: max |dS11| in dB over freqs, and the relative L2 of the port
    traces."""
    from efs_amd import batch, ports, workloads
    from efs_amd.engine import HIPEngine
    b = batch.BatchEngine(storage=storage)
    sl = [workloads.add_patch(b, w)[1] for w in sample]
    b.finalize(1, steps)
    b.iterate(steps)
    out = []
    for i, w in enumerate(sample):
        r = b.readProbeGatherBuffered(i)
        tb = ports.probe_traces(r["data"], r["numSamples"], r["stride"], w["probes"], sl[i], 1, w["dT"])
        g = HIPEngine()
        g.initGPU()
        so = workloads.configure_patch(g, w, nf2ff=False, max_samples=steps, storage="fp64")
        g.iterate(steps)
        ro = g.readProbeGatherBuffered()
        g.destroy()
        to = ports.probe_traces(ro["data"], ro["numSamples"], ro["stride"], w["probes"], so, 1, w["dT"])
        cb, c64 = ports.port_s11(tb[0], tb[1], freqs), ports.port_s11(to[0], to[1], freqs)
        sb, s64 = 20 * np.log10(np.abs(cb)), 20 * np.log10(np.abs(c64))
        k, kb = int(np.argmin(s64)), int(np.argmin(sb))
        above = s64 >= -20.0  # away from the resonance null, where dB differences are ill-conditioned
        out.append({"grid": w["gridSize"], "cells": int(np.prod(w["gridSize"])),
                    "patch_w_mm": round(w["patch"][0], 2),
                    "s11_max_abs_err_dB": float(np.max(np.abs(sb - s64))),
                    "s11_max_abs_err_dB_where_above_m20dB": float(np.max(np.abs(sb - s64)[above])),
                    "s11_max_abs_err_linear": float(np.max(np.abs(cb - c64))),
                    "s11_min_dB_fp64": float(s64[k]), "s11_min_dB": float(sb[kb]),
                    "f_res_GHz": float(freqs[k] / 1e9), "f_res_GHz_this": float(freqs[kb] / 1e9),
                    "ut_rel_l2": rel_l2(tb[0][1], to[0][1]), "it_rel_l2": rel_l2(tb[1][1], to[1][1])})
    b.destroy()
    return out


def plain_sequential(ws, steps):
    """The plain single-grid engine over the same instances, one after the
    other (its aggregate MCells/s for the same design sweep)."""
    from efs_amd import workloads
    from efs_amd.engine import HIPEngine
    cells, secs = 0, 0.0
    for w in ws:
        g = HIPEngine()
        g.initGPU()
        workloads.configure_patch(g, w, nf2ff=False, max_samples=steps + 20)
        g.iterate(20)
        g.sync()
        t0 = time.perf_counter()
        g.iterate(steps)
        g.sync()
        secs += time.perf_counter() - t0
        cells += int(np.prod(w["gridSize"])) * steps
        g.destroy()
    return cells / secs / 1e6


def run_batch(args):
    n_inst, K, W = args.instances, args.steps, args.warmup
    storage = args.storage
    t_build = time.perf_counter()
    b, ws, _ = build_patch_batch(n_inst, storage, K + W + 64, progress=True)
    t_build = time.perf_counter() - t_build
    if args.batch_shape:  # a fixed launch shape instead of the tuner (profiling runs)
        m, tpb, zc, wv, blk = args.batch_shape.split(",")
        b.setShape(m, int(tpb), int(zc), int(wv), int(blk))
        shape = b.shape()
    else:
        shape = b.tune(steps=32)
    info = b.info()
    cells = info["cells"]
    indexed = [b.coefIndexed(i) for i in range(len(ws))]
    elapsed = time_batch(b, W, K)
    value = cells * K / elapsed / 1e6
    # sweep durations (direct launches with events); the persistent form
    # has no sweep launches of its own: its sweeps are timed in the launch
    # form at the same block count
    persistent = shape["mode"] == "persistent"
    if persistent:
        b.setShape("launch", 256, 4, 4, shape["blocks"])
    b.profile(True)
    b.iterate(args.profile_steps)
    p = b.profileRead()
    b.profile(False)
    b.destroy()
    avg_e, avg_h = p["e_ms"] / max(p["e_n"], 1), p["h_ms"] / max(p["h_n"], 1)
    # per cell and sweep: FP32 60 B; FP16 fields with 16-bit coefficient
    # indices count coefficients at 2 B (30 B), FP32 coefficient arrays 42 B
    per_inst = [int(np.prod(w["gridSize"])) for w in ws]  # real cells (row padding is not algorithmic work)
    by = sum(c * (18.0 + 6 * coef_bytes(storage, ix) if storage == "fp16" else 60.0)
             for c, ix in zip(per_inst, indexed))
    dom_name, dom_ms = ("sweep_h", avg_h) if avg_h >= avg_e else ("sweep_e", avg_e)
    dom_gbs = by / (dom_ms * 1e-3) / 1e9
    # measured HBM bytes of the dominant sweep kernel (committed rocprofv3
    # FETCH/WRITE summary of this workload and shape), if there is one
    prof_shape = shape if not persistent else {"tile_threads": 256, "z_cells": 4, "waves": 4}
    dom_sym = (f"efs::batch::bsweep<{'true' if dom_name == 'sweep_h' else 'false'}, "
               f"{'f16' if storage == 'fp16' else 'float'}, {prof_shape['waves']}, {prof_shape['tile_threads']}, "
               f"{prof_shape['z_cells']}>")
    traffic, traffic_src = None, None
    tpath = BATCH_TRAFFIC_FILES.get(storage)
    if tpath and os.path.exists(tpath) and n_inst == 64:
        with open(tpath) as f:
            traffic = json.load(f).get(dom_sym, {}).get("hbm_bytes_per_launch")
        traffic_src = os.path.relpath(tpath, ROOT) + f" [{dom_sym}]" if traffic else None
    extra = {}
    print("# plain sequential baseline", file=sys.stderr, flush=True)
    plain = plain_sequential(ws, args.plain_steps)
    extra["plain_sequential_value"] = round(plain, 1)
    extra["speedup_vs_plain"] = round(value / plain, 3)
    if storage == "fp16" and not args.no_compare:
        print("# FP32 batch for comparison", file=sys.stderr, flush=True)
        b32, _, _ = build_patch_batch(n_inst, "fp32", K + W + 64)
        b32.tune(steps=8)
        v32 = b32.info()["cells"] * K / time_batch(b32, W, K) / 1e6
        b32.destroy()
        extra["batched_fp32_value"] = round(v32, 1)
        extra["speedup_vs_batched_fp32"] = round(value / v32, 3)
    if not args.no_compare:
        print("# S11 against the FP64 storage form", file=sys.stderr, flush=True)
        order = np.argsort([np.prod(w["gridSize"]) for w in ws], kind="stable")
        pick = [order[0], order[len(order) // 2], order[-1]]  # smallest, median, largest
        sample = [ws[i] for i in pick]
        freqs = np.linspace(1e9, 6e9, 501)
        extra["precision"] = {"against": FP64_NOTE.replace("same case", "single-grid engine, same instances"),
                              "instances_picked": "smallest, median and largest grid",
                              "steps": args.s11_steps, "band_GHz": [1.0, 6.0],
                              "instances": s11_error(sample, storage, args.s11_steps, freqs)}
        for key in ("s11_max_abs_err_dB", "s11_max_abs_err_dB_where_above_m20dB", "s11_max_abs_err_linear"):
            extra["precision"][key] = max(x[key] for x in extra["precision"]["instances"])
    base = None
    if not args.no_cpu_baseline:
        w0 = ws[int(np.argmin([np.prod(w["gridSize"]) for w in ws]))]
        base = cpu_baseline(w0, args.cpu_seconds,
                            what=f"the smallest instance ({'x'.join(map(str, w0['gridSize']))}, Mur + port)")
    sizes = [int(np.prod(w["gridSize"])) for w in ws]
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "MCells/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(elapsed / K * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if storage == "fp32" else "f16 storage, f32 compute", "data": "synthetic",
        "config": {"workload": f"batched design sweep: {n_inst} microstrip patch antennas (width 20-50 mm from a "
                               f"seed, graded meshes), Mur on 6 faces, 50-ohm lumped port, Gaussian f0=fc=3 GHz, "
                               f"probe gather every step; aggregate over instances",
                   "instances": n_inst, "cells_total": int(cells), "cells_min": min(sizes), "cells_max": max(sizes),
                   "seed": int(__import__("efs_amd.workloads", fromlist=["x"]).BATCH_SEED),
                   "tuned_shape": shape, "graph_steps": info["graph_steps"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "batched " + dom_name, "achieved": round(dom_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic) if traffic else None, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": by,
                     "algorithmic_model": ("60 B/cell (FP32 fields and coefficients)" if storage == "fp32" else
                                           f"FP16 fields: {sum(indexed)} of {len(indexed)} instances with 16-bit "
                                           "coefficient indices at 30 B/cell, the others 42 B/cell"),
                     "avg_launch_ms": round(dom_ms, 5),
                     "kernels": {"sweep_e": {"avg_ms": round(avg_e, 5), "GB/s": round(by / (avg_e * 1e-3) / 1e9, 1)},
                                 "sweep_h": {"avg_ms": round(avg_h, 5), "GB/s": round(by / (avg_h * 1e-3) / 1e9, 1)}}},
        "cpu_baseline": base, **extra,
        "checks": {"host_build_s": round(t_build, 1)},
    }
    print(json.dumps(result))


def patch_outputs(eng, w, slices, steps, want_ff=True):
    """S11 over 1-3 GHz (201 points), resonance and Dmax at the NF2FF
    frequency from an engine after `steps` steps."""
    from efs_amd import farfield, ports
    r = eng.readProbeGatherBuffered()
    tr = ports.probe_traces(r["data"], r["numSamples"], r["stride"], w["probes"], slices, 1, w["dT"])
    freqs = np.linspace(1e9, 3e9, 201)
    s11 = 20 * np.log10(np.abs(ports.port_s11(tr[0], tr[1], freqs)))
    k = int(np.argmin(s11))
    out = {"s11_dB": s11, "f_res_GHz": float(freqs[k] / 1e9), "s11_min_dB": float(s11[k])}
    if want_ff:
        P = w["nf2ff"]["numPoints"]
        acc = eng.readNF2FFAccumulation()
        accE, accH = np.asarray(acc["accumE"][0]), np.asarray(acc["accumH"][0])
        theta, phi = farfield.sphere_grid()
        ff = eng.computeNF2FFfarField({"pointMeta": w["pointMeta"], "theta": theta, "phi": phi,
                                       "center": [0, 0, 0], "frequency": float(w["freqs"][0]), "radius": 1.0})
        p = farfield.rad_power(accE, accH, w["pointMeta"], P)
        d = farfield.directivity(ff, len(theta), len(phi), p)
        out.update({"Dmax_dBi": float(d["DmaxdBi"]), "rad_power_W": p})
    return out


def run_patch(args):
    """The reference's patch-antenna example (86 x 87 x 71 graded mesh, Mur,
    lumped port, NF2FF at 2.4 GHz, 30,000 steps) on the plain engine."""
    from efs_amd import workloads
    from efs_amd.engine import PROF_SWEEP_E, PROF_SWEEP_H, HIPEngine
    w = workloads.patch_antenna()
    n = int(np.prod(w["gridSize"]))
    K = args.steps if args.steps != 2000 else w["nTS"]
    g = HIPEngine()
    if not g.initGPU():
        raise RuntimeError("no HIP device")
    slices = workloads.configure_patch(g, w, nf2ff=True, max_samples=K, storage=args.storage)
    g.sync()
    t0 = time.perf_counter()
    g.iterate(K)
    g.sync()
    elapsed = time.perf_counter() - t0
    value = n * K / elapsed / 1e6
    res = patch_outputs(g, w, slices, K)
    g.profileEnable(True)
    g.iterate(args.profile_steps)
    g.sync()
    ms_e, ce = g.profileRead(PROF_SWEEP_E)
    ms_h, ch = g.profileRead(PROF_SWEEP_H)
    flags = g.pathFlags()
    g.destroy()
    avg_e, avg_h = ms_e / ce, ms_h / ch
    b = sweep_bytes(n, 0, args.storage, flags)
    dom_name, dom_ms = ("sweep_h", avg_h) if avg_h >= avg_e else ("sweep_e", avg_e)
    extra = {"results": {"f_res_GHz": res["f_res_GHz"], "s11_min_dB": res["s11_min_dB"],
                         "Dmax_dBi_2p4GHz": res["Dmax_dBi"]}}
    if not args.no_compare:
        print("# FP64 run of the same engine", file=sys.stderr, flush=True)
        g64 = HIPEngine()
        g64.initGPU()
        so = workloads.configure_patch(g64, w, nf2ff=True, max_samples=K, storage="fp64")
        g64.iterate(K)
        ref = patch_outputs(g64, w, so, K)
        g64.destroy()
        extra["precision"] = {"against": FP64_NOTE,
                              "s11_max_abs_err_dB_1_3GHz": float(np.max(np.abs(res["s11_dB"] - ref["s11_dB"]))),
                              "Dmax_err_dB": float(abs(res["Dmax_dBi"] - ref["Dmax_dBi"])),
                              "fp64": {"f_res_GHz": ref["f_res_GHz"], "s11_min_dB": ref["s11_min_dB"],
                                       "Dmax_dBi": ref["Dmax_dBi"]}}
    base = None if args.no_cpu_baseline else cpu_baseline(
        w, args.cpu_seconds, what=f"the same patch case ({'x'.join(map(str, w['gridSize']))}, Mur + port)")
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "MCells/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": round(elapsed / K * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32" else "f16 storage, f32 compute", "data": "synthetic",
        "config": {"workload": "microstrip patch antenna of the reference example: 32x40 mm patch, 1.524 mm "
                               "eps_r 3.38 substrate, 60x60 mm ground, 50-ohm lumped port, Gaussian f0=fc=3 GHz, "
                               "Mur on 6 faces, NF2FF box at 2.4 GHz, whole run timed from zero fields",
                   "grid": w["gridSize"], "dT": w["dT"], "path": sorted(flags), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(b / (dom_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(b / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "traffic_source": None, "algorithmic_bytes_per_launch": b,
                     "algorithmic_model": byte_model(args.storage, flags),
                     "avg_launch_ms": round(dom_ms, 5),
                     "kernels": {"sweep_e": {"avg_ms": round(avg_e, 5)}, "sweep_h": {"avg_ms": round(avg_h, 5)}}},
        "cpu_baseline": base, **extra,
    }
    print(json.dumps(result))


def run_dispersive(args):
    """Dispersive scatterers under a TFSF plane wave (192^3, UPML-8, Drude
    sphere + Lorentz slab with ADE, 15,000 steps, NF2FF at 4 frequencies)
    and the bistatic RCS cut at 64 angles in the xy plane."""
    from efs_amd import farfield, workloads
    from efs_amd.engine import PROF_SWEEP_E, PROF_SWEEP_H, HIPEngine
    t_build = time.perf_counter()
    w = workloads.dispersive_scatter(args.size if args.size != 256 else 192)
    t_build = time.perf_counter() - t_build
    print(f"# built {w['gridSize']} in {t_build:.1f} s: {w['info']}", file=sys.stderr, flush=True)
    n = int(np.prod(w["gridSize"]))
    K = args.steps if args.steps != 2000 else w["nTS"]
    g = HIPEngine()
    if not g.initGPU():
        raise RuntimeError("no HIP device")
    workloads.configure_dispersive(g, w, storage=args.storage)
    blocks_scale = g.tuneSweeps(args.tune_reps)
    g.sync()
    t0 = time.perf_counter()
    g.iterate(K)
    g.sync()
    elapsed = time.perf_counter() - t0
    value = n * K / elapsed / 1e6
    einc = farfield.incident_spectrum(w["signal"], 2 * np.pi * w["freqs"], w["dT"])
    sph = w["info"]["sphere"]
    centre = [c * 1e-3 for c in sph["centre"]]
    cut, rcs_main = {}, []
    acc_main = g.readNF2FFAccumulation()
    for k, fr in enumerate(w["freqs"]):
        ff = g.computeNF2FFfarField({"pointMeta": w["pointMeta"], "theta": w["theta"], "phi": w["phi"],
                                     "center": centre, "frequency": float(fr), "radius": 1.0, "freqIndex": k})
        rcs_main.append(farfield.rcs(ff, einc[k]))
        s = 10 * np.log10(np.maximum(rcs_main[-1], 1e-30))
        cut[f"{fr / 1e9:g}GHz"] = {"forward_dBsm": round(float(s[0]), 2),
                                   "back_dBsm": round(float(s[len(s) // 2]), 2),
                                   "max_dBsm": round(float(s.max()), 2)}
    g.profileEnable(True)
    g.iterate(args.profile_steps)
    g.sync()
    ms_e, ce = g.profileRead(PROF_SWEEP_E)
    ms_h, ch = g.profileRead(PROF_SWEEP_H)
    flags = g.pathFlags()
    g.destroy()
    avg_e, avg_h = ms_e / ce, ms_h / ch
    b = sweep_bytes(n, w["n_pml_cells"], args.storage, flags)
    dom_name, dom_ms = ("sweep_h", avg_h) if avg_h >= avg_e else ("sweep_e", avg_e)
    extra = {"results": {"rcs_xy_cut": cut, "angles": len(w["phi"])}}
    if not args.no_compare:
        print("# FP64 run of the same engine", file=sys.stderr, flush=True)
        g64 = HIPEngine()
        g64.initGPU()
        workloads.configure_dispersive(g64, w, storage="fp64")
        g64.iterate(K)
        a64 = g64.readNF2FFAccumulation(fp64=True)
        cut_err = 0.0
        for k, fr in enumerate(w["freqs"]):
            args_ff = {"pointMeta": w["pointMeta"], "theta": w["theta"], "phi": w["phi"], "center": centre,
                       "frequency": float(fr), "radius": 1.0, "freqIndex": k}
            r64 = farfield.rcs(g64.computeNF2FFfarField(args_ff), einc[k])
            r = rcs_main[k]
            keep = r64 >= r64.max() * 1e-3  # within 30 dB of the cut's maximum
            cut_err = max(cut_err, float(np.max(np.abs(10 * np.log10(r[keep] / r64[keep])))))
        g64.destroy()
        # the fields while the wave crosses the grid (at the end of the run
        # they ring down to FP16 noise level, where relative errors mean
        # nothing): both forms again over the first 400 steps
        snap = {}
        for st in (args.storage, "fp64"):
            gs = HIPEngine()
            gs.initGPU()
            workloads.configure_dispersive(gs, w, storage=st)
            gs.iterate(400)
            snap[st] = gs.getFields(fp64=st == "fp64")
            gs.destroy()
        extra["precision"] = {"against": FP64_NOTE,
                              "E_rel_l2_step400": rel_l2(snap[args.storage]["volt"], snap["fp64"]["volt"]),
                              "H_rel_l2_step400": rel_l2(snap[args.storage]["curr"], snap["fp64"]["curr"]),
                              "nf2ff_E_rel_l2": rel_l2(acc_main["accumE"], a64["accumE"]),
                              "rcs_cut_max_abs_err_dB_within_30dB": cut_err}
    base = None if args.no_cpu_baseline else cpu_baseline(
        w, args.cpu_seconds, what=f"the same {w['gridSize'][0]}^3 UPML + ADE + TFSF workload")
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "MCells/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": round(elapsed / K * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32" else "f16 storage, f32 compute", "data": "synthetic",
        "config": {"workload": "dispersive scatterers: 192^3 x 1 mm, UPML-8, Drude sphere (12 GHz plasma "
                               "frequency) beside a Lorentz slab (eps_inf 2, delta eps 3, 6 GHz pole), ADE with "
                               "write-back, TFSF plane wave along +x (Gaussian f0 6 GHz, fc 5 GHz), NF2FF at 3/5/7/9 "
                               "GHz, bistatic RCS cut at 64 angles; whole run timed from zero fields",
                   "grid": w["gridSize"], "dT": w["dT"], "path": sorted(flags), "parallelism": "single GPU",
                   "sweep_blocks_scale": blocks_scale, "ade_cells": [o["numCells"] for o in w["ade"]["orders"]],
                   "tfsf_points": [len(w["tfsf"]["voltagePoints"]["field_idx"]),
                                   len(w["tfsf"]["currentPoints"]["field_idx"])]},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(b / (dom_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(b / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "traffic_source": None, "algorithmic_bytes_per_launch": b,
                     "algorithmic_model": byte_model(args.storage, flags),
                     "avg_launch_ms": round(dom_ms, 5),
                     "kernels": {"sweep_e": {"avg_ms": round(avg_e, 5)}, "sweep_h": {"avg_ms": round(avg_h, 5)}}},
        "cpu_baseline": base, **extra,
    }
    print(json.dumps(result))


def large_outputs(g, w, slices, K, f_idx):
    """Per-port S11 (dB) over 2-18 GHz and Dmax at freqs[f_idx] of an engine
    after K steps."""
    from efs_amd import farfield, ports
    r = g.readProbeGatherBuffered()
    tr = ports.probe_traces(r["data"], r["numSamples"], r["stride"], w["probes"], slices, 1, w["dT"])
    f = np.linspace(2e9, 18e9, 81)
    s11 = np.stack([20 * np.log10(np.abs(ports.port_s11(tr[2 * p], tr[2 * p + 1], f)))
                    for p in range(len(w["probes"]) // 2)])
    theta, phi = farfield.sphere_grid(5, 5)
    ff = g.computeNF2FFfarField({"pointMeta": w["pointMeta"], "theta": theta, "phi": phi, "center": [0.2, 0.2, 0.15],
                                 "frequency": float(w["freqs"][f_idx]), "radius": 1.0, "freqIndex": f_idx})
    acc = g.readNF2FFAccumulation()
    P = w["nf2ff"]["numPoints"]
    p_rad = farfield.rad_power(acc["accumE"][f_idx], acc["accumH"][f_idx], w["pointMeta"], P)
    d = farfield.directivity(ff, len(theta), len(phi), p_rad)
    return s11, float(d["DmaxdBi"])


def run_large(args):
    """The large mixed-precision workload (800 x 800 x 600 = 384 M cells,
    UPML-8, 200 seeded boxes, a 4 x 4 lumped-port array, NF2FF at 32
    frequencies with K = 64) on one GPU; FP16 storage by default here."""
    from efs_amd import ports, workloads
    from efs_amd.engine import PROF_SWEEP_E, PROF_SWEEP_H, HIPEngine

    def say(*a):
        print("#", *a, file=sys.stderr, flush=True)
    threads = host_threads()
    t_b = time.perf_counter()
    w = workloads.large_array(threads=threads, progress=lambda m: say(m, f"{time.perf_counter() - t_b:.1f} s"))
    n = int(np.prod(w["gridSize"]))
    K = args.steps if args.steps != 2000 else w["nTS"]
    f_idx = int(np.argmin(np.abs(w["freqs"] - w["f0"])))

    def run(storage, profile):
        g = HIPEngine()
        if not g.initGPU():
            raise RuntimeError("no HIP device")
        slices = workloads.configure_large(g, w, storage=storage, max_samples=K)
        scale = g.tuneSweeps(args.tune_reps)
        g.sync()
        say(storage, "configured", f"{time.perf_counter() - t_b:.1f} s")
        t0 = time.perf_counter()
        done = 0
        while done < K:  # a progress line per chunk; the stream stays busy in between
            step = min(2000, K - done)
            g.iterate(step)
            g.sync()
            done += step
            say(storage, "steps", done)
        elapsed = time.perf_counter() - t0
        s11, dmax = large_outputs(g, w, slices, K, f_idx)
        kern = None
        if profile:
            g.profileEnable(True)
            g.iterate(args.profile_steps)
            g.sync()
            ms_e, ce = g.profileRead(PROF_SWEEP_E)
            ms_h, ch = g.profileRead(PROF_SWEEP_H)
            kern = (ms_e / ce, ms_h / ch)
        flags = g.pathFlags()
        g.destroy()
        return {"value": n * K / elapsed / 1e6, "elapsed": elapsed, "s11": s11, "dmax": dmax, "kern": kern,
                "flags": flags, "blocks_scale": scale}

    main = run(args.storage, True)
    extra = {"results": {"Dmax_dBi_at_f0": round(main["dmax"], 3),
                         "s11_min_dB_per_port": [round(float(v), 2) for v in main["s11"].min(axis=1)]}}
    if args.storage == "fp16" and not args.no_compare:
        ref = run("fp32", False)
        extra["fp32_value"] = round(ref["value"], 1)
        extra["speedup_vs_fp32"] = round(main["value"] / ref["value"], 3)
        extra["precision_vs_fp32_same_grid"] = {
            "s11_max_abs_err_dB_2_18GHz": float(np.max(np.abs(main["s11"] - ref["s11"]))),
            "Dmax_err_dB": abs(main["dmax"] - ref["dmax"])}
    K64 = K if args.fp64_steps is None else args.fp64_steps
    if not args.no_compare and K64 > 0:
        say("FP64 run of the same engine,", K64, "steps")
        if K64 == K:
            ref64 = run("fp64", False)
        else:  # a shorter run: the main form again over the same steps
            K_full, K = K, K64
            ref64, short = run("fp64", False), run(args.storage, False)
            K = K_full
        cmp = main if K64 == K else short
        extra["precision_vs_fp64"] = {"against": FP64_NOTE, "steps": K64, "fp64_value": round(ref64["value"], 1),
                                      "s11_max_abs_err_dB_2_18GHz": float(np.max(np.abs(cmp["s11"] - ref64["s11"]))),
                                      "Dmax_err_dB": abs(cmp["dmax"] - ref64["dmax"])}
    kern = main["kern"]
    b = sweep_bytes(n, w["n_pml_cells"], args.storage, main["flags"])
    dom_name, dom_ms = ("sweep_h", kern[1]) if kern[1] >= kern[0] else ("sweep_e", kern[0])
    result = {
        "metric": METRIC, "value": round(main["value"], 1), "unit": "MCells/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": round(main["elapsed"] / K * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.storage == "fp32" else "f16 storage, f32 compute",
        "data": "synthetic",
        "config": {"workload": "large mixed-precision grid: 800 x 800 x 600 x 0.5 mm (384 M cells), UPML-8, 200 seeded "
                               "boxes (150 dielectric, 50 PEC plates), 4 x 4 array of 50-ohm lumped ports (Gaussian f0 "
                               "10 GHz, fc 8 GHz), NF2FF six-face box at 32 frequencies (K = 64); whole run timed from "
                               "zero fields", "grid": w["gridSize"], "dT": w["dT"], "path": sorted(main["flags"]),
                   "nf2ff_points": w["nf2ff"]["numPoints"], "sweep_blocks_scale": main["blocks_scale"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(b / (dom_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(b / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "traffic_source": None, "algorithmic_bytes_per_launch": b,
                     "algorithmic_model": byte_model(args.storage, main["flags"]),
                     "avg_launch_ms": round(dom_ms, 5),
                     "kernels": {"sweep_e": {"avg_ms": round(kern[0], 5)}, "sweep_h": {"avg_ms": round(kern[1], 5)}}},
        "cpu_baseline": None, **extra,
    }
    print(json.dumps(result))


def time_stress_window(eng, K, E):
    """The timed steps of the stress workload: K steps with the field
    energy read every E steps, and once at the end when K is not a multiple
    of E (so a window shorter than E still holds one energy call).  Returns
    the energies read."""
    energies, done = [], 0
    while done < K:
        k = min(E, K - done)
        eng.iterate(k)
        done += k
        if done % E == 0 or done == K:
            energies.append(eng.energy())
    return energies


def calibrate_copy():
    import torch
    a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    del a, b


def stream_reference(reps=20):
    """Achievable HBM streaming rates on this GPU, for reading the sweeps'
    real traffic against (efs_amd/csrc/stream_ref.hip, HIP events, arrays
    larger than the last-level cache, best over 1024-16384 workgroups):
    - add: out_c = a_c + 0.1 * b_c over FP32 arrays of 64 M floats (k = 1:
      2 reads + 1 write; k = 3: 6 reads + 3 writes like a sweep's field
      streams);
    - copy: o = a over 256 M floats (1 GiB each way), 4 independent 16-byte
      loads in flight per thread, plain or non-temporal accesses;
    - memcpy: hipMemcpyAsync device to device of the same arrays;
    - read: a sum over the 1 GiB array (read only).
    "GB/s" is the best read+write rate (add, copy, memcpy); the read-only
    rate is reported beside it."""
    import ctypes
    path = os.path.join(ROOT, "efs_amd", "libefs_stream_ref.so")
    if not os.path.exists(path):  # an auxiliary figure: the bench line goes on without it
        print(f"stream reference library not built: {path}", file=sys.stderr)
        return None
    lib = ctypes.CDLL(path)
    lib.efs_stream_ref.restype = ctypes.c_int
    lib.efs_stream_ref.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_double)]
    rates = {}
    n = 64 * 1024 * 1024
    for k in (1, 3):
        for grid in (1024, 2048, 4096, 8192):
            ms = ctypes.c_double()
            if lib.efs_stream_ref(0, n, k, grid, reps, ctypes.byref(ms)) != 0:
                return None
            gbs = 3 * k * 4 * n / (ms.value * 1e-3) / 1e9
            key = f"add_k{k}"
            if key not in rates or gbs > rates[key]["GB/s"]:
                rates[key] = {"GB/s": round(gbs, 1), "workgroups": grid, "avg_ms": round(ms.value, 4)}
    if hasattr(lib, "efs_stream_kind"):
        lib.efs_stream_kind.restype = ctypes.c_int
        lib.efs_stream_kind.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_double)]
        n = 256 * 1024 * 1024
        for kind, key, nbytes in ((0, "copy", 8 * n), (1, "copy_nontemporal", 8 * n), (2, "read", 4 * n),
                                  (3, "memcpy", 8 * n)):
            for grid in ((1024, 2048, 4096, 8192, 16384) if kind != 3 else (1,)):
                ms = ctypes.c_double()
                if lib.efs_stream_kind(0, kind, n, grid, reps, ctypes.byref(ms)) != 0:
                    return None
                gbs = nbytes / (ms.value * 1e-3) / 1e9
                if key not in rates or gbs > rates[key]["GB/s"]:
                    rates[key] = {"GB/s": round(gbs, 1), "avg_ms": round(ms.value, 4)}
                    if kind != 3:
                        rates[key]["workgroups"] = grid
    rw = {k: v for k, v in rates.items() if k != "read"}
    best_key = max(rw, key=lambda k: rw[k]["GB/s"])
    best = {"GB/s": rw[best_key]["GB/s"], "best_of": best_key, "rates": rates,
            "kernel": "best read+write rate of: out_c = a_c + 0.1 * b_c (k = 1, 3), copy (plain, non-temporal), "
                      "hipMemcpy device to device; arrays larger than the last-level cache"}
    if "read" in rates:
        best["read_only_GB/s"] = rates["read"]["GB/s"]
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--energy-interval", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--storage", choices=["fp32", "fp16"], default="fp32",
                    help="field storage: fp32 (plain engine) or fp16 (per-component power-of-two scales, "
                         "FP32 updates)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one elementwise copy of 1 GiB -> 1 GiB first (reference kernel for HBM counters)")
    ap.add_argument("--workload", choices=["stress", "batch", "patch", "dispersive", "large"], default="stress",
                    help="stress: the 256^3 UPML stress grid; batch: the batched patch-antenna design sweep; "
                         "patch: the reference's patch-antenna example on the plain engine; dispersive: "
                         "Drude/Lorentz scatterers under a TFSF plane wave (192^3) with the RCS cut; large: the "
                         "384 M-cell mixed-precision grid with a 4 x 4 port array")
    ap.add_argument("--instances", type=int, default=64, help="batch: number of patch instances")
    ap.add_argument("--plain-steps", type=int, default=200, help="batch: steps per instance of the plain baseline")
    ap.add_argument("--batch-shape", default=None,
                    help="batch: fixed launch shape mode,tile_threads,z_cells,waves,blocks instead of the autotuner")
    ap.add_argument("--s11-steps", type=int, default=20000,
                    help="batch: steps of the S11 comparison against the FP64 storage form")
    ap.add_argument("--precision-steps", type=int, default=512,
                    help="stress: steps of the comparison against the FP64 storage form (full grid)")
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="large: steps of the FP64-storage run at full size (default: the run's own steps; "
                         "0 skips it)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the comparison runs (FP32 speed and FP64 error) of --storage fp16 / batch")
    ap.add_argument("--tune-reps", type=int, default=5,
                    help="sweep pairs the sweep-shape tuner times per candidate before the run (0: planner's shape)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; multi-GPU runs are not part "
                                   "of this build (use --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload in ("batch", "patch", "dispersive", "large"):
        if args.calibrate_copy:
            calibrate_copy()
        return {"batch": run_batch, "patch": run_patch, "dispersive": run_dispersive,
                "large": run_large}[args.workload](args)

    if args.calibrate_copy:
        calibrate_copy()

    from efs_amd import workloads
    from efs_amd.engine import PROF_STEP, PROF_SWEEP_E, PROF_SWEEP_H, HIPEngine

    t_build = time.perf_counter()
    w = workloads.stress_grid(args.size)
    t_build = time.perf_counter() - t_build
    n = int(np.prod(w["gridSize"]))

    eng = HIPEngine()
    if not eng.initGPU():
        raise RuntimeError("no HIP device")
    workloads.configure_engine(eng, w, storage=args.storage)
    blocks_scale = eng.tuneSweeps(args.tune_reps)

    K, E = args.steps, args.energy_interval
    kb = int(w["nf2ff"]["kBlock"])
    # The NF2FF surface is sampled every step and folded into the
    # frequency bins every kb samples.  The untimed lead-in (warmup plus
    # `align` steps) ends the timed window on a fold, so the window holds
    # at least one fold at any step count, and every window of kb steps
    # or more holds floor-or-ceil(K / kb) of them.
    align = (-(args.warmup + K)) % kb
    lead = args.warmup + align
    eng.iterate(lead)
    eng.energy()
    eng.sync()
    folds = (lead + K) // kb - lead // kb
    t0 = time.perf_counter()
    energies = time_stress_window(eng, K, E)
    eng.sync()
    elapsed = time.perf_counter() - t0
    mcells = n * K / elapsed / 1e6

    # kernel durations from HIP events on the engine's stream
    eng.profileEnable(True)
    eng.iterate(args.profile_steps)
    eng.sync()
    ms_e, ce = eng.profileRead(PROF_SWEEP_E)
    ms_h, ch = eng.profileRead(PROF_SWEEP_H)
    ms_s, cs = eng.profileRead(PROF_STEP)
    eng.profileEnable(False)
    avg_e, avg_h, avg_s = ms_e / ce, ms_h / ch, ms_s / cs
    flags_main = eng.pathFlags()
    b = sweep_bytes(n, w["n_pml_cells"], args.storage, flags_main)
    gbs_e = b / (avg_e * 1e-3) / 1e9
    gbs_h = b / (avg_h * 1e-3) / 1e9
    dom_name, dom_ms = ("sweep_h", avg_h) if avg_h >= avg_e else ("sweep_e", avg_e)
    dom_sym = sweep_kernel_symbol(flags_main, dom_name[-1])
    traffic = measured_traffic(dom_sym, args.size, args.storage)
    dom_gbs = b / (dom_ms * 1e-3) / 1e9
    acc = eng.readNF2FFAccumulation()
    finite = bool(np.all(np.isfinite(acc["accumE"])) and np.all(np.isfinite(energies)))

    extra = {}
    eng.destroy()
    if args.storage == "fp16" and not args.no_compare:
        # the same steps on the plain FP32 engine
        e32 = HIPEngine()
        e32.initGPU()
        workloads.configure_engine(e32, w)
        e32.iterate(lead)
        e32.sync()
        t1 = time.perf_counter()
        time_stress_window(e32, K, E)
        e32.sync()
        v32 = n * K / (time.perf_counter() - t1) / 1e6
        e32.destroy()
        extra = {"fp32_value": round(v32, 1), "speedup_vs_fp32": round(mcells / v32, 3)}
    if not args.no_compare:
        print("# FP64 run of the same engine", file=sys.stderr, flush=True)
        extra["precision"] = precision_stress(w, args.storage, args.precision_steps)
    base = None if args.no_cpu_baseline else cpu_baseline(w, args.cpu_seconds)
    step_bytes = 2 * b
    sref = stream_reference()
    if sref and traffic:
        # the dominant sweep's real (counter-measured) bytes per launch over
        # its launch time, against the plain streaming rate
        sref["sweep_real_GB/s"] = round(traffic / (dom_ms * 1e-3) / 1e9, 1)
        sref["sweep_real_frac"] = round(sref["sweep_real_GB/s"] / sref["GB/s"], 4)
    result = {
        "metric": METRIC,
        "value": round(mcells, 1),
        "unit": "MCells/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32" else "f16 storage, f32 compute",
        "data": "synthetic",
        "config": {
            "workload": (f"{'plain FP32' if args.storage == 'fp32' else 'FP16-storage (FP32 update)'} stress grid "
                         f"{args.size}^3, 1 mm cells, UPML-8 on 6 faces, 40 dielectric boxes, "
                         f"10 PEC plates, 50-ohm lumped port, Gaussian f0=5 GHz fc=4 GHz, energy every {E} steps, "
                         f"NF2FF 6-face box {w['nf2ff']['numPoints']} points x 16 frequencies (1-9 GHz)"),
            "grid": w["gridSize"], "pml_cells": w["n_pml_cells"], "dT": w["dT"], "seed": w["seed"],
            "nf2ff_k_block": w["nf2ff"]["kBlock"], "sweep_blocks_scale": blocks_scale, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic else None,
            "traffic_source": (os.path.relpath(TRAFFIC_FILES[args.storage], ROOT) + f" [{dom_sym}]") if traffic else None,
            "algorithmic_bytes_per_launch": b,
            "algorithmic_model": byte_model(args.storage, flags_main),
            "avg_launch_ms": round(dom_ms, 5),
            "kernels": {"sweep_e": {"avg_ms": round(avg_e, 5), "GB/s": round(gbs_e, 1)},
                        "sweep_h": {"avg_ms": round(avg_h, 5), "GB/s": round(gbs_h, 1)},
                        "step": {"avg_ms": round(avg_s, 5),
                                 "sweeps_GB/s": round(step_bytes / (avg_s * 1e-3) / 1e9, 1)}},
            "stream_reference": sref,
        },
        "cpu_baseline": base,
        **extra,
        "checks": {"finite": finite, "energy_last": energies[-1] if energies else None,
                   "host_build_s": round(t_build, 1),
                   "timed_window": {"steps": K, "nf2ff_folds": folds, "energy_calls": len(energies),
                                    "untimed_lead_in_steps": lead, "nf2ff_k_block": kb}},
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
