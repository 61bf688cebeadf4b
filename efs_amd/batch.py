"""Python binding of the batched form (include/efs_batch.h): many
independent grids advanced together on one GPU.

Each instance is configured like one WebGPUEngine (init coefficients,
configureMur, configureExcitation, probe indices of a buffered gather) and
then every instance steps in lockstep: one sweep launch per half-step covers
all of them, and blocks of steps replay a captured HIP graph.  Results per
instance are those of a single engine on the same arrays.

There is no fallback: without the native library this module raises.
"""
import ctypes as C

import numpy as np

from . import engine as _eng

STORAGE = {"fp32": 0, "fp16": 1}


class BatchEngine:
    def __init__(self, device=0, storage="fp32"):
        self.lib = _eng.load_library()
        h = C.c_void_p()
        rc = self.lib.efs_batch_create(C.byref(h), int(device))
        if rc != 0:
            raise RuntimeError(self.lib.efs_batch_last_error(None).decode())
        self.h = h
        self._grids = []
        self._ngather = []
        self._final = False
        self.storage = storage
        self._check(self.lib.efs_batch_set_storage(self.h, STORAGE[storage]))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.efs_batch_last_error(self.h).decode())

    def destroy(self):
        if self.h:
            self.lib.efs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -- configuration (before finalize) -------------------------------------
    def add(self, gridSize, coefficients):
        """init(gridSize, {vv, vi, ii, iv}) of a new instance; returns its id."""
        nx, ny, nz = (int(v) for v in gridSize)
        c = {k: _eng.f32(coefficients[k]) for k in ("vv", "vi", "ii", "iv")}
        for k, v in c.items():
            if v.size != 3 * nx * ny * nz:
                raise RuntimeError(f"coefficient array {k} has {v.size} values, expected {3 * nx * ny * nz}")
        out = C.c_uint32()
        self._check(self.lib.efs_batch_add(self.h, nx, ny, nz, c["vv"], c["vi"], c["ii"], c["iv"], C.byref(out)))
        self._grids.append((nx, ny, nz))
        self._ngather.append(0)
        return out.value

    def configureExcitation(self, i, exc):
        sig = _eng.f32(exc["signal"])
        n = len(exc["pos"])
        self._check(self.lib.efs_batch_cfg_excitation(self.h, i, sig, len(sig), _eng.u32(exc["delay"]),
                                                      _eng.f32(exc["amp"]), _eng.u32(exc["dir"]),
                                                      _eng.u32(exc["pos"]), n, int(exc.get("period", 0))))

    def configureMur(self, i, cfg):
        if "coeff_nyP" in cfg:
            coeff = np.concatenate([_eng.f32(cfg["coeff_nyP"]), _eng.f32(cfg["coeff_nyPP"])])
            nidx = np.concatenate([_eng.u32(cfg["normal_idx_nyP"]), _eng.u32(cfg["normal_idx_nyPP"])])
            sidx = np.concatenate([_eng.u32(cfg["shifted_idx_nyP"]), _eng.u32(cfg["shifted_idx_nyPP"])])
        else:
            nidx, sidx = _eng.u32(cfg["normal_idx"]), _eng.u32(cfg["shifted_idx"])
            c = cfg["coeff"]
            coeff = np.full(len(nidx), c, dtype=np.float32) if np.isscalar(c) else _eng.f32(c)
        self._check(self.lib.efs_batch_cfg_mur(self.h, i, len(nidx), _eng.f32(coeff), _eng.u32(nidx),
                                               _eng.u32(sidx)))

    def configureProbeGather(self, i, indices):
        idx = _eng.u32(indices)
        self._check(self.lib.efs_batch_cfg_gather(self.h, i, idx, len(idx)))
        self._ngather[i] = len(idx)

    def finalize(self, sampleInterval=1, maxSamples=0):
        self._check(self.lib.efs_batch_finalize(self.h, int(sampleInterval), int(maxSamples)))
        self._final = True

    def tune(self, steps=32):
        """Autotune the launch shape (form, column tile width, occupancy,
        block count) on whole steps; leaves the batch reset.  Returns
        shape()."""
        b, w = C.c_uint32(), C.c_uint32()
        self._check(self.lib.efs_batch_tune(self.h, int(steps), C.byref(b), C.byref(w)))
        return self.shape()

    MODES = {"launch": 0, "persistent": 1}

    def setShape(self, mode="launch", tile_threads=256, z_cells=4, waves=4, blocks=2048):
        """Force a launch shape: mode "launch" (one launch per phase, graph
        replays; lanes of 4 z cells, or 8 in FP16 storage) or "persistent"
        (one cooperative launch per up to 64 steps, grid barriers between
        phases; 256-lane tiles of 4-cell lanes, 3 waves per SIMD).  Fields
        and the step count are kept."""
        self._check(self.lib.efs_batch_set_shape(self.h, self.MODES[mode], int(tile_threads), int(z_cells),
                                                 int(waves), int(blocks)))

    def shape(self):
        m, t, z, w, b, g = C.c_int(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self.lib.efs_batch_get_shape(self.h, C.byref(m), C.byref(t), C.byref(z), C.byref(w), C.byref(b),
                                                 C.byref(g)))
        mode = {v: k for k, v in self.MODES.items()}[m.value]
        return {"mode": mode, "tile_threads": t.value, "z_cells": z.value, "waves": w.value, "blocks": b.value,
                "grid": g.value}

    # -- running -------------------------------------------------------------
    def reset(self):
        self._check(self.lib.efs_batch_reset(self.h))

    def iterate(self, n):
        self._check(self.lib.efs_batch_iterate(self.h, int(n)))

    def sync(self):
        self._check(self.lib.efs_batch_sync(self.h))

    @property
    def numTS(self):
        v = C.c_uint32()
        self._check(self.lib.efs_batch_num_ts(self.h, C.byref(v)))
        return v.value

    def info(self):
        b, w, g, c = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        self._check(self.lib.efs_batch_info(self.h, C.byref(b), C.byref(w), C.byref(g), C.byref(c)))
        return {"blocks": b.value, "waves": w.value, "graph_steps": g.value, "cells": c.value,
                "instances": len(self._grids)}

    def coefIndexed(self, i):
        """True when instance i keeps its coefficients as 16-bit indices into
        tables (both sides), False for FP32 coefficient arrays."""
        v = C.c_uint32()
        self._check(self.lib.efs_batch_coef_form(self.h, i, C.byref(v)))
        return bool(v.value)

    def murSplit(self):
        """Mur entries applied by the point kernels ("listed") and by the E
        sweep itself ("in_sweep", z-face entries)."""
        a, s = C.c_uint32(), C.c_uint32()
        self._check(self.lib.efs_batch_mur_split(self.h, C.byref(a), C.byref(s)))
        return {"listed": a.value, "in_sweep": s.value}

    def readProbeGatherBuffered(self, i):
        """{data, numSamples, stride} of instance i (readProbeGatherBuffered
        form, src/webgpu-engine.mjs:2689-2715)."""
        ns = C.c_uint32()
        self._check(self.lib.efs_batch_read_gather(self.h, i, None, C.byref(ns)))
        stride = 2 * self._ngather[i]
        out = np.zeros(max(ns.value * stride, 1), dtype=np.float32)
        if ns.value and stride:
            self._check(self.lib.efs_batch_read_gather(self.h, i, out.ctypes.data, C.byref(ns)))
        return {"data": out[:ns.value * stride], "numSamples": ns.value, "stride": stride}

    def getFields(self, i):
        nx, ny, nz = self._grids[i]
        v = np.zeros(3 * nx * ny * nz, dtype=np.float32)
        c = np.zeros_like(v)
        self._check(self.lib.efs_batch_get_fields(self.h, i, v, c))
        return {"volt": v, "curr": c}

    def profile(self, enable):
        self._check(self.lib.efs_batch_profile(self.h, 1 if enable else 0))

    def profileRead(self):
        me, ne, mh, nh = C.c_double(), C.c_uint32(), C.c_double(), C.c_uint32()
        self._check(self.lib.efs_batch_profile_read(self.h, C.byref(me), C.byref(ne), C.byref(mh), C.byref(nh)))
        return {"e_ms": me.value, "e_n": ne.value, "h_ms": mh.value, "h_n": nh.value}
