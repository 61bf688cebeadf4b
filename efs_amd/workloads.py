"""Synthetic workloads built with the FP64 operator builder.

``stress_grid`` is the plain-FP32 stress case: an n^3 grid of 1 mm cells
with 8-cell UPML on all six faces, 40 seeded dielectric boxes (eps_r in
U(1,10), kappa in U(0,0.05) S/m), 10 PEC plates, a 50-ohm lumped port fed
by a Gaussian pulse (f0 = 5 GHz, fc = 4 GHz), field energy every 100 steps
and a six-face NF2FF box at 16 frequencies from 1 to 9 GHz.  n = 256 is the
full-size workload; smaller n keep the same structure for tests.
"""
import numpy as np

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
    # lumped port: 2 z-edges in series at the centre, 50 ohm
    mat.lumped_resistor(2, (c, c, c - 1), 2, 50.0)
    coeffs = op.build_coefficients(mesh, mat, dT)
    regions = op.build_upml(mesh, coeffs, [pml] * 6, dT)
    signal = op.gaussian_pulse(5e9, 4e9, dT)
    lin = lambda x, y, z: x * n * n + y * n + z  # noqa: E731
    excitation = {"signal": signal, "delay": np.zeros(2, np.uint32), "amp": np.full(2, 0.5, np.float32),
                  "dir": np.full(2, 2, np.uint32), "pos": np.array([lin(c, c, c - 1), lin(c, c, c)], np.uint32),
                  "period": 0}
    pts, meta, primal, dual = op.nf2ff_box(mesh, [pml + 1] * 3, [n - pml - 2] * 3)
    freqs = np.linspace(1e9, 9e9, nfreq)
    nf2ff = {"surfaceIndices": pts, "numPoints": len(pts) // 3, "omega": list(2 * np.pi * freqs), "dT": dT,
             "primalEdgeLens": primal, "dualEdgeLens": dual, "kBlock": k_block}
    n_pml = sum(int(np.prod(r["numLines"])) for r in regions)
    return {"gridSize": [n, n, n], "coefficients": op.to_f32(coeffs), "pmlRegions": regions,
            "excitation": excitation, "nf2ff": nf2ff, "pointMeta": meta, "freqs": freqs, "dT": dT,
            "seed": seed, "n_pml_cells": n_pml, "pml": pml}


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
