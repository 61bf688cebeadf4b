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


def sweep_bytes(n_cells, n_pml, storage="fp32"):
    """Algorithmic HBM bytes of one E (or H) sweep launch.
    Per cell: read V, vv, vi, H (3 components each) and write V = 15 values:
    60 B with FP32 fields, 42 B with FP16 fields (9 field values x 2 B + 6
    FP32 coefficients).  Per UPML cell in addition: flux read + write and
    three UPML coefficients for 3 components = 60 B."""
    per_cell = 42.0 if storage == "fp16" else 60.0
    return per_cell * n_cells + 60.0 * n_pml


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_sweeps_256cube_stress_traffic.json")


def sweep_kernel_symbol(flags, side):
    """Name of the kernel that runs the E or H sweep for the path flags."""
    if "vec4" not in flags:
        return f"efs::sweep_{side}1<true>"
    suffix = ""
    if "fused" in flags:
        suffix = "f" + ("c" if "coef_idx" in flags else "") + ("p" if "upml_idx" in flags else "")
    elif "upml4" in flags:
        suffix = "mc" if "coef_idx" in flags else "m"
    else:
        suffix = "c" if "coef_idx" in flags else ""
    return f"efs::sweep_{side}4{suffix}" + ("_h" if "f16" in flags else "")


def measured_traffic(kernel, size):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement of this workload (None if absent)."""
    if size != 256 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    return t.get(kernel, {}).get("hbm_bytes_per_launch")


def precision_sample(storage, n=48, steps=150):
    """Error of the GPU engine against the FP64 oracle on a bounded sample of
    the same workload (n^3 stress grid, same builder and extensions)."""
    from efs_amd import workloads
    from efs_amd.engine import HIPEngine
    from oracle import oracle as orc
    w = workloads.stress_grid(n, pml=8, nfreq=3, n_boxes=6, n_plates=2)
    o = orc.Oracle(w["gridSize"], w["coefficients"], fp64=True)
    o.configurePML(w["pmlRegions"])
    o.configureExcitation(w["excitation"])
    nf = w["nf2ff"]
    o.configureNF2FF(nf["surfaceIndices"], nf["omega"], nf["dT"], nf["primalEdgeLens"], nf["dualEdgeLens"])
    o.iterate(steps)
    ref = o.getFields()
    refE, refH = o.readNF2FF()
    g = HIPEngine()
    g.initGPU()
    workloads.configure_engine(g, w, storage=storage)
    g.iterate(steps)
    f = g.getFields()
    acc = g.readNF2FFAccumulation()

    def rl2(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out = {"against": "FP64 oracle (C restatement, double storage)", "grid": [n, n, n], "steps": steps,
           "volt_rel_l2": rl2(f["volt"], ref["volt"]), "curr_rel_l2": rl2(f["curr"], ref["curr"]),
           "energy_rel": abs(g.energy() - o.energy()) / o.energy(),
           "nf2ff_rel_l2": max(rl2(np.concatenate([acc["accumE"][k], acc["accumH"][k]]),
                                   np.concatenate([refE[k], refH[k]])) for k in range(len(refE)))}
    g.destroy()
    return out


def cpu_baseline(w, seconds_target=15.0):
    """Time the CPU oracle (all host threads, f32 variant) on a bounded
    sample of the same workload: the same grid, coefficients, UPML and
    excitation; NF2FF and energy are not part of the sample."""
    from oracle import oracle as orc
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    threads = min(threads, 16)
    orc.Oracle.set_threads(threads)
    o = orc.Oracle(w["gridSize"], w["coefficients"])
    o.configurePML(w["pmlRegions"])
    o.configureExcitation(w["excitation"])
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
            "sample": f"{steps} steps of the same {w['gridSize'][0]}^3 UPML+excitation workload "
                      f"({elapsed:.1f} s) with the C oracle (float storage, double arithmetic, OpenMP); "
                      f"NF2FF and energy excluded"}


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
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; multi-GPU runs are not part "
                                   "of this build (use --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b

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

    eng.iterate(args.warmup)
    eng.energy()
    eng.sync()

    K, E = args.steps, args.energy_interval
    energies = []
    t0 = time.perf_counter()
    done = 0
    while done < K:
        k = min(E, K - done)
        eng.iterate(k)
        done += k
        if done % E == 0:
            energies.append(eng.energy())
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
    b = sweep_bytes(n, w["n_pml_cells"], args.storage)
    gbs_e = b / (avg_e * 1e-3) / 1e9
    gbs_h = b / (avg_h * 1e-3) / 1e9
    dom_name, dom_ms = ("sweep_h", avg_h) if avg_h >= avg_e else ("sweep_e", avg_e)
    dom_sym = sweep_kernel_symbol(eng.pathFlags(), dom_name[-1])
    traffic = measured_traffic(dom_sym, args.size)
    dom_gbs = b / (dom_ms * 1e-3) / 1e9
    acc = eng.readNF2FFAccumulation()
    finite = bool(np.all(np.isfinite(acc["accumE"])) and np.all(np.isfinite(energies)))

    extra = {}
    if args.storage == "fp16":
        # the same steps on the plain FP32 engine, and the error against FP64
        eng.destroy()
        e32 = HIPEngine()
        e32.initGPU()
        workloads.configure_engine(e32, w)
        e32.iterate(args.warmup)
        e32.sync()
        t1 = time.perf_counter()
        done = 0
        while done < K:
            k = min(E, K - done)
            e32.iterate(k)
            done += k
            if done % E == 0:
                e32.energy()
        e32.sync()
        v32 = n * K / (time.perf_counter() - t1) / 1e6
        e32.destroy()
        extra = {"fp32_value": round(v32, 1), "speedup_vs_fp32": round(mcells / v32, 3),
                 "precision": precision_sample("fp16")}
    base = None if args.no_cpu_baseline else cpu_baseline(w, args.cpu_seconds)
    step_bytes = 2 * b
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
            "workload": f"plain FP32 stress grid {args.size}^3, 1 mm cells, UPML-8 on 6 faces, 40 dielectric boxes, "
                        f"10 PEC plates, 50-ohm lumped port, Gaussian f0=5 GHz fc=4 GHz, energy every {E} steps, "
                        f"NF2FF 6-face box {w['nf2ff']['numPoints']} points x 16 frequencies (1-9 GHz)",
            "grid": w["gridSize"], "pml_cells": w["n_pml_cells"], "dT": w["dT"], "seed": w["seed"],
            "nf2ff_k_block": w["nf2ff"]["kBlock"], "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic else None,
            "traffic_source": (os.path.relpath(TRAFFIC_FILE, ROOT) + f" [{dom_sym}]") if traffic else None,
            "algorithmic_bytes_per_launch": b,
            "avg_launch_ms": round(dom_ms, 5),
            "kernels": {"sweep_e": {"avg_ms": round(avg_e, 5), "GB/s": round(gbs_e, 1)},
                        "sweep_h": {"avg_ms": round(avg_h, 5), "GB/s": round(gbs_h, 1)},
                        "step": {"avg_ms": round(avg_s, 5),
                                 "sweeps_GB/s": round(step_bytes / (avg_s * 1e-3) / 1e9, 1)}},
        },
        "cpu_baseline": base,
        **extra,
        "checks": {"finite": finite, "energy_last": energies[-1] if energies else None,
                   "host_build_s": round(t_build, 1)},
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
