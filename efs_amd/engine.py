"""Python host binding of the MI355X engine (libefs_engine.so, C ABI in
include/efs_engine.h).

``HIPEngine`` keeps the method names, argument objects and return shapes of
the reference's ``WebGPUEngine`` (src/webgpu-engine.mjs:994-3498) so code
written against that class ports line by line; synchronous Python returns
stand in for the JavaScript Promises.  Misuse raises ``RuntimeError`` with
the engine's message, as the reference throws ``Error``.

There is no fallback: if the native library is missing, importing the
engine raises.  Build it with ``python -m efs_amd.build`` (or
``__graft_entry__.build()``).
"""
import ctypes as C
import os

import numpy as np

from . import build as _build

LIB_PATH = _build.ENGINE_LIB

EFS_OK, EFS_ERR_NO_DEVICE, EFS_ERR_STATE, EFS_ERR_ARG, EFS_ERR_HIP, EFS_ERR_NOMEM = 0, 1, 2, 3, 4, 5
PROF_SWEEP_E, PROF_SWEEP_H, PROF_STEP = 0, 1, 2

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None

class EfsSizes(C.Structure):
    """efs_sizes of include/efs_engine.h."""
    _fields_ = [(n, C.c_uint32) for n in ("nx", "ny", "nz", "gather_n", "gather_buffered_n", "gather_buffered_max",
                                          "nf2ff_points", "nf2ff_freqs", "steady_state_n")]


# Every entry point of include/efs_engine.h with its ctypes signature.
SIGNATURES = {
    "efs_version": (C.c_char_p, []),
    "efs_device_count": (C.c_int, []),
    "efs_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "efs_destroy": (None, [C.c_void_p]),
    "efs_last_error": (C.c_char_p, [C.c_void_p]),
    "efs_init": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, _f32p, _f32p, _f32p, _f32p]),
    "efs_cfg_excitation": (C.c_int, [C.c_void_p, _f32p, C.c_uint32, _u32p, _f32p, _u32p, _u32p, C.c_uint32,
                                     C.c_uint32]),
    "efs_cfg_pml": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _u32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
    "efs_cfg_mur": (C.c_int, [C.c_void_p, C.c_uint32, _f32p, _u32p, _u32p]),
    "efs_cfg_ade_clear": (C.c_int, [C.c_void_p]),
    "efs_cfg_ade_add": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _f32p, _f32p, _f32p,
                                  _f32p, _f32p, _f32p]),
    "efs_set_ade_apply": (C.c_int, [C.c_void_p, C.c_int]),
    "efs_cfg_tfsf": (C.c_int, [C.c_void_p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _u32p, _f32p, _f32p,
                               C.c_uint32, _u32p, _u32p, _f32p, _f32p]),
    "efs_cfg_rlc": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _u32p, _u32p, _f32p]),
    "efs_iterate": (C.c_int, [C.c_void_p, C.c_uint32]),
    "efs_sync": (C.c_int, [C.c_void_p]),
    "efs_num_ts": (C.c_uint32, [C.c_void_p]),
    "efs_set_num_ts": (C.c_int, [C.c_void_p, C.c_uint32]),
    "efs_get_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "efs_upload_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "efs_cfg_gather": (C.c_int, [C.c_void_p, _u32p, C.c_uint32]),
    "efs_read_gather": (C.c_int, [C.c_void_p, _f32p]),
    "efs_cfg_gather_buffered": (C.c_int, [C.c_void_p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "efs_read_gather_buffered": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
    "efs_cfg_nf2ff": (C.c_int, [C.c_void_p, _u32p, C.c_uint32, _f64p, C.c_uint32, C.c_double, _f32p, _u32p,
                                C.c_uint32, C.c_uint32, C.c_uint32]),
    "efs_read_nf2ff": (C.c_int, [C.c_void_p, _f32p, _f32p]),
    "efs_read_nf2ff_f64": (C.c_int, [C.c_void_p, _f64p, _f64p]),
    "efs_get_fields_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "efs_farfield": (C.c_int, [C.c_void_p, C.c_uint32, _f32p, _f32p, C.c_uint32, _f32p, C.c_uint32, _f64p,
                               C.c_double, C.c_double, _f32p]),
    "efs_farfield_mirror": (C.c_int, [C.c_void_p, C.c_uint32, _f32p, _f32p, C.c_uint32, _f32p, C.c_uint32, _f64p,
                                      C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, _f32p]),
    "efs_farfield_surface": (C.c_int, [C.c_void_p, C.c_uint32, _f32p, _f64p, _f64p, _f32p, C.c_uint32, _f32p,
                                       C.c_uint32, _f64p, C.c_double, C.c_double, _f64p]),
    "efs_energy": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "efs_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "efs_profile_read": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "efs_path_flags": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "efs_tune_sweeps": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)]),
    "efs_get_sizes": (C.c_int, [C.c_void_p, C.POINTER(EfsSizes)]),
    "efs_set_nf2ff_compensated": (C.c_int, [C.c_void_p, C.c_int]),
    "efs_farfield_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                       C.c_uint32, _f64p, C.c_double, C.c_double, C.c_void_p]),
    "efs_cfg_pbc": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _f64p]),
    "efs_set_storage": (C.c_int, [C.c_void_p, C.c_int]),
    "efs_field_scales": (C.c_int, [C.c_void_p, _f32p]),
    "efs_cfg_steady_state": (C.c_int, [C.c_void_p, _u32p, C.c_uint32, C.c_uint32]),
    "efs_ss_set": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32]),
    "efs_ss_read": (C.c_int, [C.c_void_p, _f32p, _f32p, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]),
    # batched form (include/efs_batch.h)
    "efs_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "efs_batch_destroy": (None, [C.c_void_p]),
    "efs_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "efs_batch_set_storage": (C.c_int, [C.c_void_p, C.c_int]),
    "efs_batch_add": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, _f32p, _f32p, _f32p, _f32p,
                                C.POINTER(C.c_uint32)]),
    "efs_batch_cfg_excitation": (C.c_int, [C.c_void_p, C.c_uint32, _f32p, C.c_uint32, _u32p, _f32p, _u32p, _u32p,
                                           C.c_uint32, C.c_uint32]),
    "efs_batch_cfg_mur": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _f32p, _u32p, _u32p]),
    "efs_batch_cfg_gather": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, C.c_uint32]),
    "efs_batch_finalize": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "efs_batch_tune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "efs_batch_reset": (C.c_int, [C.c_void_p]),
    "efs_batch_iterate": (C.c_int, [C.c_void_p, C.c_uint32]),
    "efs_batch_sync": (C.c_int, [C.c_void_p]),
    "efs_batch_num_ts": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "efs_batch_read_gather": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)]),
    "efs_batch_get_fields": (C.c_int, [C.c_void_p, C.c_uint32, _f32p, _f32p]),
    "efs_batch_sizes": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, C.POINTER(C.c_uint32)]),
    "efs_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint64)]),
    "efs_batch_mur_split": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "efs_batch_set_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "efs_batch_get_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "efs_batch_coef_form": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "efs_batch_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "efs_batch_profile_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
}


def load_library(path=LIB_PATH):
    """Load libefs_engine.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP engine library not built: {path} (run `python -m efs_amd.build`)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def device_count():
    return load_library().efs_device_count()


def f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())


def u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).ravel())


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


class HIPEngine:
    """WebGPUEngine-compatible engine on one MI355X (HIP device)."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.device_index = device
        self.ctx = None
        self.numLines = None
        self.totalCells = 0
        self._energy_last = 0.0
        self._energy_pending = None
        self._nf2ff = None
        self._gather_n = 0
        self._gather_b = None

    # -- lifetime --------------------------------------------------------
    def initGPU(self):
        """initGPU() (src/webgpu-engine.mjs:1107-1135): False without a GPU."""
        h = C.c_void_p()
        rc = self.lib.efs_create(C.byref(h), int(self.device_index))
        if rc == EFS_ERR_NO_DEVICE:
            return False
        if rc != EFS_OK:
            raise RuntimeError(self.lib.efs_last_error(None).decode())
        self.ctx = h
        return True

    def destroy(self):
        if self.ctx:
            self.lib.efs_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _check(self, rc):
        if rc != EFS_OK:
            msg = self.lib.efs_last_error(self.ctx).decode() if self.ctx else "engine not created"
            raise RuntimeError(msg)

    def _need_device(self):
        if not self.ctx:
            raise RuntimeError("HIP device not initialized. Call initGPU() first.")

    # -- init / configure -----------------------------------------------
    def init(self, gridSize, coefficients, storage="fp32"):
        """init(gridSize, {vv, vi, ii, iv}) (src/webgpu-engine.mjs:1143-1185).
        storage="fp16" keeps the fields as FP16 with per-component
        power-of-two scales (updates stay FP32); storage="fp64" is the FP64
        run of the same engine (fields, UPML flux, Mur state and NF2FF in
        double; no ADE, TFSF, RLC, PBC or steady state); reads return FP32."""
        self._need_device()
        codes = {"fp32": 0, "fp16": 1, "fp64": 2}
        if storage not in codes:
            raise ValueError("storage must be 'fp32', 'fp16' or 'fp64'")
        self._check(self.lib.efs_set_storage(self.ctx, codes[storage]))
        nx, ny, nz = (int(v) for v in gridSize)
        c = coefficients
        self._check(self.lib.efs_init(self.ctx, nx, ny, nz, f32(c["vv"]), f32(c["vi"]), f32(c["ii"]),
                                      f32(c["iv"])))
        self.numLines = [nx, ny, nz]
        self.totalCells = nx * ny * nz
        self._energy_last = 0.0
        self._energy_pending = None
        self._nf2ff = None
        self._gather_n = 0
        self._gather_b = None

    def configureExcitation(self, excitation):
        """configureExcitation (src/webgpu-engine.mjs:1199-1227)."""
        self._need_device()
        e = excitation
        sig, amp = f32(e["signal"]), f32(e["amp"])
        self._check(self.lib.efs_cfg_excitation(self.ctx, sig, len(sig), u32(e["delay"]), amp, u32(e["dir"]),
                                                u32(e["pos"]), len(amp), int(e.get("period", 0) or 0)))

    def configurePML(self, pmlRegions):
        """configurePML (src/webgpu-engine.mjs:1235-1323)."""
        self._need_device()
        regs = list(pmlRegions)
        start = u32([r["startPos"] for r in regs]) if regs else u32([0])
        num = u32([r["numLines"] for r in regs]) if regs else u32([0])
        cat = {}
        for k in ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn"):
            cat[k] = f32(np.concatenate([f32(r[k]) for r in regs])) if regs else f32([0])
        self._check(self.lib.efs_cfg_pml(self.ctx, len(regs), start, num, cat["vv"], cat["vvfo"], cat["vvfn"],
                                         cat["ii"], cat["iifo"], cat["iifn"]))

    def configureMur(self, config):
        """configureMur (src/webgpu-engine.mjs:1822-1918): single or dual form."""
        self._need_device()
        if "coeff_nyP" in config:
            coeff = np.concatenate([f32(config["coeff_nyP"]), f32(config["coeff_nyPP"])])
            nidx = np.concatenate([u32(config["normal_idx_nyP"]), u32(config["normal_idx_nyPP"])])
            sidx = np.concatenate([u32(config["shifted_idx_nyP"]), u32(config["shifted_idx_nyPP"])])
        else:
            nidx, sidx = u32(config["normal_idx"]), u32(config["shifted_idx"])
            c = config["coeff"]
            coeff = np.full(len(nidx), c, dtype=np.float32) if np.isscalar(c) else f32(c)
        self._check(self.lib.efs_cfg_mur(self.ctx, len(nidx), f32(coeff), u32(nidx), u32(sidx)))

    def configurePBC(self, config):
        """configurePBC({axes: [{axis, phase}]}) (src/webgpu-engine.mjs:1976-2155)."""
        self._need_device()
        axes = config.get("axes") or []
        if not axes:
            return
        self._check(self.lib.efs_cfg_pbc(self.ctx, len(axes), u32([a["axis"] for a in axes]),
                                         np.ascontiguousarray([float(a.get("phase") or 0) for a in axes])))

    def configureSteadyState(self, config):
        """configureSteadyState({probe_idx, periodSamples, threshold})
        (src/webgpu-engine.mjs:2161-2217)."""
        self._need_device()
        idx = u32(config["probe_idx"])
        self._ss_n = len(idx)
        self._ss_threshold = float(config.get("threshold") or 1e-6)
        self._ss_period = int(config["periodSamples"])
        self._check(self.lib.efs_cfg_steady_state(self.ctx, idx, len(idx), self._ss_period))

    def _ss_state(self):
        cur, rec = C.c_uint32(0), C.c_int(0)
        n = max(getattr(self, "_ss_n", 0), 1)
        e1, e2 = np.zeros(n, np.float32), np.zeros(n, np.float32)
        self._check(self.lib.efs_ss_read(self.ctx, e1, e2, C.byref(cur), C.byref(rec)))
        k = getattr(self, "_ss_n", 0)
        return e1[:k], e2[:k], cur.value, bool(rec.value)

    @property
    def ssRecording(self):
        return self._ss_state()[3]

    @ssRecording.setter
    def ssRecording(self, on):
        self._check(self.lib.efs_ss_set(self.ctx, 1 if on else 0, self._ss_state()[2]))

    @property
    def ssCurrentSample(self):
        return self._ss_state()[2]

    @ssCurrentSample.setter
    def ssCurrentSample(self, v):
        self._check(self.lib.efs_ss_set(self.ctx, 1 if self._ss_state()[3] else 0, int(v)))

    def readSteadyState(self):
        """The two period energies and the sample counter."""
        e1, e2, cur, _ = self._ss_state()
        return {"energy_period1": e1, "energy_period2": e2, "currentSample": cur}

    def checkConvergence(self):
        """Both periods recorded and max |e2-e1|/|e1| < threshold (the CPU
        engine's rule, src/webgpu-fdtd.mjs:1077-1095)."""
        e1, e2, cur, _ = self._ss_state()
        if not hasattr(self, "_ss_period") or cur < 2 * self._ss_period:
            return False
        e1 = e1.astype(np.float64)
        e2 = e2.astype(np.float64)
        ok = np.abs(e1) >= 1e-30
        ratio = np.abs(e2[ok] - e1[ok]) / np.abs(e1[ok])
        return bool((ratio.max() if ratio.size else 0.0) < self._ss_threshold)

    def configureLorentzADE(self, config, apply=False):
        """configureLorentzADE (src/webgpu-engine.mjs:1358-1475).  apply=True
        selects the openEMS write-back (field -= ADE state)."""
        self._need_device()
        self._check(self.lib.efs_cfg_ade_clear(self.ctx))
        for order in config["orders"]:
            n = int(order["numCells"])
            for d in order["directions"]:
                def co(key):
                    v = d.get(key)
                    return f32(np.zeros(n) if v is None else v)
                self._check(self.lib.efs_cfg_ade_add(self.ctx, 1 if order["hasLorentz"] else 0, int(d["dir"]), n,
                                                     u32(d["pos_idx"]), co("v_int_ADE"), co("v_ext_ADE"),
                                                     co("v_Lor_ADE"), co("i_int_ADE"), co("i_ext_ADE"),
                                                     co("i_Lor_ADE")))
        self._check(self.lib.efs_set_ade_apply(self.ctx, 1 if apply else 0))

    configureLorentz = configureLorentzADE

    def configureTFSF(self, config):
        """configureTFSF (src/webgpu-engine.mjs:1527-1645)."""
        self._need_device()
        sig = f32(config["signal"])
        # a list of point records, or a dict of columns of the same keys
        vp = config.get("voltagePoints")
        cp = config.get("currentPoints")

        def cols(pts):
            if pts is None or len(pts) == 0:
                return 0, (u32([0]), u32([0]), f32([0]), f32([0]))
            if isinstance(pts, dict):
                return len(pts["field_idx"]), (u32(pts["field_idx"]), u32(pts["delay_int"]),
                                               f32(pts["delay_frac"]), f32(pts["amp"]))
            return len(pts), (u32([p["field_idx"] for p in pts]), u32([p["delay_int"] for p in pts]),
                              f32([p["delay_frac"] for p in pts]), f32([p["amp"] for p in pts]))
        nv, v = cols(vp)
        ni, i = cols(cp)
        self._check(self.lib.efs_cfg_tfsf(self.ctx, sig, len(sig), int(config.get("period", 0) or 0), nv, *v,
                                          ni, *i))

    RLC_KEYS = ("i2v", "ilv", "vvd", "vv2", "vj1", "vj2", "ib0", "b1", "b2")

    def configureRLC(self, config):
        """configureRLC (src/webgpu-engine.mjs:1722-1794)."""
        self._need_device()
        el = config["elements"]
        if not el:
            return
        coeffs = f32([[e.get(k, 0.0) for k in self.RLC_KEYS] for e in el])
        self._check(self.lib.efs_cfg_rlc(self.ctx, len(el), u32([e["field_idx"] for e in el]),
                                         u32([e["direction"] for e in el]), u32([e["type_flag"] for e in el]),
                                         coeffs))

    # -- stepping ----------------------------------------------------------
    def iterate(self, numSteps):
        """iterate(n) (src/webgpu-engine.mjs:2407-2504): enqueue, no wait."""
        self._need_device()
        self._check(self.lib.efs_iterate(self.ctx, int(numSteps)))

    def sync(self):
        self._check(self.lib.efs_sync(self.ctx))

    @property
    def numTS(self):
        return self.lib.efs_num_ts(self.ctx) if self.ctx else 0

    @numTS.setter
    def numTS(self, v):
        self._check(self.lib.efs_set_num_ts(self.ctx, int(v)))

    # -- fields ------------------------------------------------------------
    def getFields(self, fp64=False):
        """getFields() (src/webgpu-engine.mjs:2533-2557); fp64=True returns
        float64 arrays (the FP64 storage form's values; efs_get_fields_f64)."""
        self._need_device()
        dt = np.float64 if fp64 else np.float32
        v = np.empty(3 * self.totalCells, dtype=dt)
        c = np.empty(3 * self.totalCells, dtype=dt)
        fn = self.lib.efs_get_fields_f64 if fp64 else self.lib.efs_get_fields
        self._check(fn(self.ctx, v.ctypes.data, c.ctypes.data))
        return {"volt": v, "curr": c}

    def uploadFields(self, volt, curr):
        """uploadFields(v, c) (src/webgpu-engine.mjs:2564-2571)."""
        self._need_device()
        v = None if volt is None else f32(volt)
        c = None if curr is None else f32(curr)
        self._check(self.lib.efs_upload_fields(self.ctx, None if v is None else v.ctypes.data,
                                               None if c is None else c.ctypes.data))

    def readProbe(self, positions):
        """readProbe (src/webgpu-engine.mjs:2512-2531): [volt, curr] per position."""
        f = self.getFields()
        nx, ny, nz = self.numLines
        out = np.zeros(2 * len(positions), dtype=np.float32)
        for i, p in enumerate(positions):
            j = p["component"] * self.totalCells + p["x"] * ny * nz + p["y"] * nz + p["z"]
            out[2 * i], out[2 * i + 1] = f["volt"][j], f["curr"][j]
        return out

    # -- probe gather --------------------------------------------------------
    def configureProbeGather(self, indices):
        idx = u32(indices)
        if len(idx) == 0:
            return
        self._check(self.lib.efs_cfg_gather(self.ctx, idx, len(idx)))
        self._gather_n = len(idx)

    def readProbeGather(self):
        if not self._gather_n:
            return None
        out = np.empty(2 * self._gather_n, dtype=np.float32)
        self._check(self.lib.efs_read_gather(self.ctx, out))
        return out

    def configureProbeGatherBuffered(self, indices, maxSamples, sampleInterval):
        idx = u32(indices)
        if len(idx) == 0:
            return
        self._check(self.lib.efs_cfg_gather_buffered(self.ctx, idx, len(idx), int(maxSamples), int(sampleInterval)))
        self._gather_b = (len(idx), int(maxSamples))

    def readProbeGatherBuffered(self):
        """{data, numSamples, stride} or None (src/webgpu-engine.mjs:2689-2694)."""
        if not self._gather_b:
            return None
        n, mx = self._gather_b
        ns = C.c_uint32(0)
        self._check(self.lib.efs_read_gather_buffered(self.ctx, None, C.byref(ns)))
        if ns.value == 0:
            return None
        out = np.empty(ns.value * 2 * n, dtype=np.float32)
        self._check(self.lib.efs_read_gather_buffered(self.ctx, out.ctypes.data, C.byref(ns)))
        return {"data": out, "numSamples": ns.value, "stride": 2 * n}

    # -- NF2FF -----------------------------------------------------------------
    def configureNF2FFAccumulation(self, config):
        """configureNF2FFAccumulation (src/webgpu-engine.mjs:2729-2827).
        ``omega`` may be a scalar (reference form) or a list of F values;
        ``kBlock`` sets the buffered steps per DFT fold (default 64);
        ``compensated`` (default True) keeps the accumulators' rounding
        errors (TwoSum per fold), False adds with a plain FP32 +=."""
        pts = u32(config["surfaceIndices"])
        P = len(pts) // 3
        if "numPoints" in config and int(config["numPoints"]) != P:
            raise ValueError(f"configureNF2FFAccumulation: numPoints {config['numPoints']} does not match "
                             f"surfaceIndices ({P} points)")
        if P == 0:
            return
        om = f64(np.atleast_1d(config["omega"]))
        els = [f32(a) for a in list(config["primalEdgeLens"]) + list(config["dualEdgeLens"])]
        self._check(self.lib.efs_set_nf2ff_compensated(self.ctx, 0 if config.get("compensated", True) is False else 1))
        self._check(self.lib.efs_cfg_nf2ff(self.ctx, pts, P, om, len(om), float(config["dT"]),
                                           f32(np.concatenate(els)), u32([len(a) for a in els]),
                                           int(config.get("maxTS", 0) or 0), int(config.get("windowType", 0) or 0),
                                           int(config.get("kBlock", 0) or 0)))
        self._nf2ff = (P, len(om), np.isscalar(config["omega"]))

    def readNF2FFAccumulation(self, fp64=False):
        """{accumE, accumH}: 6P floats each for a scalar omega, else [F, 6P];
        fp64=True returns float64 (efs_read_nf2ff_f64)."""
        if not self._nf2ff:
            return None
        P, F, scalar = self._nf2ff
        dt = np.float64 if fp64 else np.float32
        e = np.empty(F * 6 * P, dtype=dt)
        h = np.empty(F * 6 * P, dtype=dt)
        self._check((self.lib.efs_read_nf2ff_f64 if fp64 else self.lib.efs_read_nf2ff)(self.ctx, e, h))
        if scalar:
            return {"accumE": e, "accumH": h}
        return {"accumE": e.reshape(F, 6 * P), "accumH": h.reshape(F, 6 * P)}

    def computeNF2FFfarField(self, config):
        """computeNF2FFfarField (src/webgpu-engine.mjs:2871-2967): 5 floats
        per angle (P_rad, Et_re, Et_im, Ep_re, Ep_im).  Optional
        config["mirror"] = {"type": "PEC"|"PMC", "direction": 0..2,
        "position": p, "imageNormals": bool} adds the surface's image in
        that plane (src/nf2ff.mjs:377-398; efs_farfield_mirror).  By default
        the reflected faces keep their normals, as the reference does;
        imageNormals=True reverses the normals of the reflected faces
        perpendicular to the plane (image theory)."""
        th, ph = f32(config["theta"]), f32(config["phi"])
        out = np.empty(5 * len(th) * len(ph), dtype=np.float32)
        m = config.get("mirror")
        if m is None:
            self._check(self.lib.efs_farfield(self.ctx, int(config.get("freqIndex", 0)), f32(config["pointMeta"]),
                                              th, len(th), ph, len(ph), f64(config["center"]),
                                              float(config["frequency"]), float(config["radius"]), out))
            return out
        mtype = {"PEC": 0, "PMC": 1}[str(m.get("type", "PEC")).upper()]
        self._check(self.lib.efs_farfield_mirror(self.ctx, int(config.get("freqIndex", 0)), f32(config["pointMeta"]),
                                                 th, len(th), ph, len(ph), f64(config["center"]),
                                                 float(config["frequency"]), float(config["radius"]),
                                                 int(m["direction"]), float(m.get("position", 0.0)), mtype,
                                                 1 if m.get("imageNormals") else 0, out))
        return out

    def farfieldSurface(self, point_meta, acc_e, acc_h, theta, phi, center, frequency, radius=1.0):
        """Far field of caller-supplied surface points (efs_farfield_surface):
        point_meta 8 floats per point, acc_e / acc_h 6 doubles per point;
        returns 5 doubles per angle (P_rad, Et re/im, Ep re/im).  Needs
        initGPU only."""
        meta = f32(point_meta)
        P = meta.size // 8
        ae, ah = f64(acc_e), f64(acc_h)
        if ae.size != 6 * P or ah.size != 6 * P:
            raise RuntimeError(f"surface fields need 6 values per point ({6 * P}), got {ae.size} and {ah.size}")
        th, ph = f32(theta), f32(phi)
        out = np.empty(5 * len(th) * len(ph), dtype=np.float64)
        self._check(self.lib.efs_farfield_surface(self.ctx, P, meta, ae, ah, th, len(th), ph, len(ph), f64(center),
                                                  float(frequency), float(radius), out))
        return out

    # -- energy ----------------------------------------------------------------
    def energy(self):
        """Current field energy (synchronous, FP64 reduction)."""
        out = C.c_double(0.0)
        self._check(self.lib.efs_energy(self.ctx, C.byref(out)))
        return out.value

    def computeEnergy(self):
        """computeEnergy() (src/webgpu-engine.mjs:2974-3021): returns the
        previous call's value (0 on the first call), as the reference's
        double-buffered readback does."""
        prev = self._energy_pending
        self._energy_pending = self.energy()
        if prev is not None:
            self._energy_last = prev
        return self._energy_last

    def flushEnergy(self):
        """flushEnergy() (src/webgpu-engine.mjs:3027-3037)."""
        if self._energy_pending is not None:
            self._energy_last = self._energy_pending
            self._energy_pending = None
        return self._energy_last

    # -- profiling ---------------------------------------------------------------
    def profileEnable(self, on=True):
        self._check(self.lib.efs_profile_enable(self.ctx, 1 if on else 0))

    def profileRead(self, which):
        ms, n = C.c_double(0), C.c_uint64(0)
        self._check(self.lib.efs_profile_read(self.ctx, int(which), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def fieldScales(self):
        """Per-component storage scales: volt 0..2, curr 0..2."""
        out = np.zeros(6, np.float32)
        self._check(self.lib.efs_field_scales(self.ctx, out))
        return out

    def tuneSweeps(self, reps=5):
        """Pick the sweep workgroup target by timing (efs_tune_sweeps); the
        fields are left as they were.  Returns the chosen factor on the
        planner's target (1.0 where the grid's path is not tuned)."""
        s = C.c_double(0.0)
        self._check(self.lib.efs_tune_sweeps(self.ctx, int(reps), C.byref(s)))
        return s.value

    def pathFlags(self):
        """Forms in use: set of 'vec4', 'coef_idx', 'upml4', 'fused', 'upml_idx',
        'f16', 'nf2ff_capture', 'coef_idx16', 'f64', 'side_work', 'sweep_epilogue', 'fuse_tfsf_ade',
        'fuse_tfsf_hooks'."""
        f = C.c_uint32(0)
        self._check(self.lib.efs_path_flags(self.ctx, C.byref(f)))
        names = ("vec4", "coef_idx", "upml4", "fused", "upml_idx", "f16", "nf2ff_capture", "coef_idx16", "f64",
                 "side_work", "sweep_epilogue", "fuse_tfsf_ade", "fuse_tfsf_hooks")
        return {n for b, n in enumerate(names) if f.value & (1 << b)}


WebGPUEngine = HIPEngine  # drop-in name for code written against the reference class
