"""ctypes binding of the CPU oracle (oracle/fdtd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product engine
(efs_amd) never imports this module.

``Oracle`` mirrors the method names of the reference's ``CPUFDTDEngine``
(src/webgpu-fdtd.mjs:13-1202) so parity tests read like the reference's own
tests: ``configureExcitation``, ``configurePML``, ``configureLorentz``,
``configureTFSF``, ``configureRLC``, ``configureMur``, ``step``, ``iterate``,
``getFields``, ``uploadFields``; plus the WebGPUEngine hooks it checks
(NF2FF accumulation, far field, energy, buffered probe gather).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdtd_oracle.so")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


def build():
    """Compile the oracle library with its Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    for pfx in ("orc32_", "orc64_"):
        def f(name, res, *args):
            fn = getattr(lib, pfx + name)
            fn.restype = res
            fn.argtypes = list(args)
        f("create", vp, C.c_uint32, C.c_uint32, C.c_uint32, _f64p, _f64p, _f64p, _f64p)
        f("destroy", None, vp)
        f("set_threads", None, C.c_int)
        f("max_threads", C.c_int)
        f("num_ts", C.c_uint32, vp)
        f("set_num_ts", None, vp, C.c_uint32)
        f("get_fields", None, vp, _f64p, _f64p)
        f("upload_fields", None, vp, _f64p, _f64p)
        for n in ("update_voltages", "update_currents", "apply_excitation", "update_volt_ade",
                  "update_curr_ade", "apply_tfsf_voltage", "apply_tfsf_current", "pre_voltage_rlc",
                  "apply_rlc", "mur_pre", "mur_post", "mur_apply", "step", "clear_pml", "clear_ade"):
            f(n, None, vp)
        f("cfg_excitation", None, vp, _f64p, C.c_uint32, _u32p, _f64p, _u32p, _u32p, C.c_uint32, C.c_uint32)
        f("add_pml_region", None, vp, _u32p, _u32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p)
        f("pml_step", None, vp, C.c_int)
        f("get_pml_flux", None, vp, C.c_uint32, _f64p, _f64p)
        f("set_pml_flux", None, vp, C.c_uint32, _f64p, _f64p)
        f("set_ade_apply", None, vp, C.c_int)
        f("add_ade_dir", None, vp, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
          _f64p, _f64p, _f64p, _f64p, _f64p, _f64p)
        f("get_ade_state", None, vp, C.c_uint32, _f64p, _f64p, _f64p, _f64p)
        f("cfg_tfsf", None, vp, _f64p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _u32p, _f64p, _f64p,
          C.c_uint32, _u32p, _u32p, _f64p, _f64p)
        f("cfg_rlc", None, vp, C.c_uint32, _u32p, _u32p, _u32p, _f64p)
        f("get_rlc_state", None, vp, _f64p, _f64p, _f64p)
        f("cfg_mur", None, vp, C.c_uint32, _f64p, _u32p, _u32p)
        f("cfg_nf2ff", None, vp, _u32p, C.c_uint32, _f64p, C.c_uint32, C.c_double,
          C.POINTER(C.c_void_p), _u32p, C.c_uint32, C.c_uint32)
        f("nf2ff_accumulate", None, vp, C.c_uint32)
        f("read_nf2ff", None, vp, _f64p, _f64p)
        f("farfield", None, _f64p, _f64p, _f64p, C.c_uint32, _f64p, C.c_uint32, _f64p, C.c_uint32,
          _f64p, C.c_double, C.c_double, _f64p)
        f("farfield_mirror", None, _f64p, _f64p, _f64p, C.c_uint32, _f64p, C.c_uint32, _f64p, C.c_uint32,
          _f64p, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, _f64p)
        f("energy", C.c_double, vp)
        f("cfg_gather_buffered", None, vp, _u32p, C.c_uint32, C.c_uint32, C.c_uint32)
        f("read_gather_buffered", C.c_uint32, vp, _f64p)
        f("iterate", None, vp, C.c_uint32)
        f("cfg_steady_state", None, vp, _u32p, C.c_uint32, C.c_uint32, C.c_double)
        f("ss_set", None, vp, C.c_int, C.c_uint32)
        f("ss_read", C.c_uint32, vp, _f64p, _f64p)
        f("ss_accumulate", None, vp)
        f("ss_converged", C.c_int, vp)
        f("cfg_pbc", None, vp, C.c_uint32, _u32p, _f64p)
        f("pbc_voltage", None, vp)
        f("pbc_current", None, vp)
    _lib = lib
    return lib


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).ravel())


def _zeros_or(a, n):
    return _f64(np.zeros(n) if a is None else a)


class Oracle:
    """CPU FDTD engine: f32 variant (JS semantics) or f64 variant."""

    def __init__(self, numLines, coefficients, fp64=False):
        self.lib = _load()
        self.p = "orc64_" if fp64 else "orc32_"
        self.numLines = [int(v) for v in numLines]
        nx, ny, nz = self.numLines
        self.totalCells = nx * ny * nz
        c = coefficients
        self.h = self._f("create")(nx, ny, nz, _f64(c["vv"]), _f64(c["vi"]), _f64(c["ii"]), _f64(c["iv"]))
        self._pml = []
        self._ade = []
        self._rlc_n = 0

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self._f("destroy")(h)
            self.h = None

    @staticmethod
    def set_threads(n):
        lib = _load()
        lib.orc32_set_threads(int(n))
        lib.orc64_set_threads(int(n))

    @staticmethod
    def max_threads():
        return _load().orc32_max_threads()

    # -- state ---------------------------------------------------------
    @property
    def numTS(self):
        return self._f("num_ts")(self.h)

    @numTS.setter
    def numTS(self, v):
        self._f("set_num_ts")(self.h, int(v))

    def idx(self, n, x, y, z):
        nx, ny, nz = self.numLines
        return n * self.totalCells + x * ny * nz + y * nz + z

    def getFields(self):
        v = np.zeros(3 * self.totalCells)
        c = np.zeros(3 * self.totalCells)
        self._f("get_fields")(self.h, v, c)
        return {"volt": v, "curr": c}

    def uploadFields(self, volt, curr):
        self._f("upload_fields")(self.h, _f64(volt), _f64(curr))

    # -- configuration ---------------------------------------------------
    def configureExcitation(self, exc):
        sig = _f64(exc["signal"])
        amp = _f64(exc["amp"])
        self._f("cfg_excitation")(self.h, sig, len(sig), _u32(exc["delay"]), amp, _u32(exc["dir"]),
                                  _u32(exc["pos"]), len(amp), int(exc.get("period", 0) or 0))

    def configurePML(self, regions):
        self._f("clear_pml")(self.h)
        self._pml = []
        for r in regions:
            self._f("add_pml_region")(self.h, _u32(r["startPos"]), _u32(r["numLines"]), _f64(r["vv"]),
                                      _f64(r["vvfo"]), _f64(r["vvfn"]), _f64(r["ii"]), _f64(r["iifo"]),
                                      _f64(r["iifn"]))
            self._pml.append(3 * int(np.prod(r["numLines"])))

    def pmlFlux(self, r):
        n = self._pml[r]
        v, c = np.zeros(n), np.zeros(n)
        self._f("get_pml_flux")(self.h, r, v, c)
        return v, c

    def setPmlFlux(self, r, vflux, cflux):
        self._f("set_pml_flux")(self.h, r, _f64(vflux), _f64(cflux))

    def configureLorentz(self, config, apply=False):
        self._f("clear_ade")(self.h)
        self._f("set_ade_apply")(self.h, 1 if apply else 0)
        self._ade = []
        for order in config["orders"]:
            n = int(order["numCells"])
            for d in order["directions"]:
                self._f("add_ade_dir")(self.h, 1 if order["hasLorentz"] else 0, int(d["dir"]), n,
                                       _u32(d["pos_idx"]), _zeros_or(d.get("v_int_ADE"), n),
                                       _zeros_or(d.get("v_ext_ADE"), n), _zeros_or(d.get("v_Lor_ADE"), n),
                                       _zeros_or(d.get("i_int_ADE"), n), _zeros_or(d.get("i_ext_ADE"), n),
                                       _zeros_or(d.get("i_Lor_ADE"), n))
                self._ade.append(n)

    def adeState(self, k):
        n = self._ade[k]
        out = [np.zeros(n) for _ in range(4)]
        self._f("get_ade_state")(self.h, k, *out)
        return dict(zip(("volt_ADE", "volt_Lor_ADE", "curr_ADE", "curr_Lor_ADE"), out))

    def configureTFSF(self, cfg):
        sig = _f64(cfg["signal"])
        # a list of point records, or a dict of columns of the same keys
        vp = cfg.get("voltagePoints")
        cp = cfg.get("currentPoints")

        def cols(pts):
            if pts is None or len(pts) == 0:
                return 0, (_u32([0]), _u32([0]), _f64([0]), _f64([0]))
            if isinstance(pts, dict):
                return len(pts["field_idx"]), (_u32(pts["field_idx"]), _u32(pts["delay_int"]),
                                               _f64(pts["delay_frac"]), _f64(pts["amp"]))
            return len(pts), (_u32([p["field_idx"] for p in pts]), _u32([p["delay_int"] for p in pts]),
                              _f64([p["delay_frac"] for p in pts]), _f64([p["amp"] for p in pts]))
        nv, v = cols(vp)
        ni, i = cols(cp)
        self._f("cfg_tfsf")(self.h, sig, len(sig), int(cfg.get("period", 0) or 0), nv, *v, ni, *i)

    RLC_KEYS = ("i2v", "ilv", "vvd", "vv2", "vj1", "vj2", "ib0", "b1", "b2")

    def configureRLC(self, cfg):
        el = cfg["elements"]
        self._rlc_n = len(el)
        coeffs = _f64([[e.get(k, 0.0) for k in self.RLC_KEYS] for e in el])
        self._f("cfg_rlc")(self.h, len(el), _u32([e["field_idx"] for e in el]),
                           _u32([e["direction"] for e in el]), _u32([e["type_flag"] for e in el]), coeffs)

    def rlcState(self):
        n = self._rlc_n
        vdn, jn, vil = np.zeros(3 * n), np.zeros(3 * n), np.zeros(max(n, 1))
        self._f("get_rlc_state")(self.h, vdn, jn, vil)
        return {"Vdn": vdn, "Jn": jn, "VIl": vil[:n]}

    def configureMur(self, cfg):
        """Single-component ({coeff, normal_idx, shifted_idx}) or
        dual-component ({coeff_nyP, ...}) form, as configureMur accepts."""
        if "coeff_nyP" in cfg:
            coeff = np.concatenate([_f64(cfg["coeff_nyP"]), _f64(cfg["coeff_nyPP"])])
            nidx = np.concatenate([_u32(cfg["normal_idx_nyP"]), _u32(cfg["normal_idx_nyPP"])])
            sidx = np.concatenate([_u32(cfg["shifted_idx_nyP"]), _u32(cfg["shifted_idx_nyPP"])])
        else:
            nidx = _u32(cfg["normal_idx"])
            sidx = _u32(cfg["shifted_idx"])
            c = cfg["coeff"]
            coeff = np.full(len(nidx), float(c)) if np.isscalar(c) else _f64(c)
        self._f("cfg_mur")(self.h, len(nidx), _f64(coeff), _u32(nidx), _u32(sidx))

    def configureNF2FF(self, points, omegas, dT, primalEdgeLens, dualEdgeLens, maxTS=0, windowType=0):
        pts = _u32(points)
        om = _f64(np.atleast_1d(omegas))
        els = [_f64(a) for a in list(primalEdgeLens) + list(dualEdgeLens)]
        self._nf_keep = els
        ptrs = (C.c_void_p * 6)(*[a.ctypes.data for a in els])
        lens = _u32([len(a) for a in els])
        self._nf_P = len(pts) // 3
        self._nf_F = len(om)
        self._f("cfg_nf2ff")(self.h, pts, self._nf_P, om, self._nf_F, float(dT), ptrs, lens, int(maxTS),
                             int(windowType))

    def configureNF2FFAccumulation(self, cfg):
        """The engine's configuration dict form (surfaceIndices, omega or
        frequency, dT, primal/dualEdgeLens, maxTS, windowType)."""
        om = cfg["omega"] if "omega" in cfg else 2 * np.pi * np.atleast_1d(cfg["frequency"])
        self.configureNF2FF(cfg["surfaceIndices"], om, cfg["dT"], cfg["primalEdgeLens"], cfg["dualEdgeLens"],
                            cfg.get("maxTS", 0), cfg.get("windowType", 0))

    def nf2ffAccumulate(self, ts):
        self._f("nf2ff_accumulate")(self.h, int(ts))

    def readNF2FF(self):
        n = self._nf_F * 6 * self._nf_P
        e, h = np.zeros(n), np.zeros(n)
        self._f("read_nf2ff")(self.h, e, h)
        return e.reshape(self._nf_F, 6 * self._nf_P), h.reshape(self._nf_F, 6 * self._nf_P)

    def configureProbeGatherBuffered(self, indices, maxSamples, sampleInterval):
        idx = _u32(indices)
        self._g_n = len(idx)
        self._g_max = int(maxSamples)
        self._f("cfg_gather_buffered")(self.h, idx, len(idx), int(maxSamples), int(sampleInterval))

    def readProbeGatherBuffered(self):
        out = np.zeros(max(self._g_max * 2 * self._g_n, 1))
        ns = self._f("read_gather_buffered")(self.h, out)
        return out[: ns * 2 * self._g_n].reshape(ns, 2 * self._g_n)

    def energy(self):
        return self._f("energy")(self.h)

    # -- steady state (CPUFDTDEngine.configureSteadyState / accumulateEnergy /
    #    checkConvergence, src/webgpu-fdtd.mjs:1037-1095) ---------------------
    def configureSteadyState(self, cfg):
        idx = _u32(cfg["probe_idx"])
        self._ss_n = len(idx)
        self._f("cfg_steady_state")(self.h, idx, len(idx), int(cfg["periodSamples"]),
                                    float(cfg.get("threshold") or 1e-6))

    def ssSet(self, recording, currentSample=None):
        if currentSample is None:
            currentSample = self.ssRead()["currentSample"]
        self._f("ss_set")(self.h, 1 if recording else 0, int(currentSample))

    def ssRead(self):
        e1, e2 = np.zeros(max(self._ss_n, 1)), np.zeros(max(self._ss_n, 1))
        cur = self._f("ss_read")(self.h, e1, e2)
        return {"energy_period1": e1[: self._ss_n], "energy_period2": e2[: self._ss_n], "currentSample": cur}

    def accumulateEnergy(self):
        self._f("ss_accumulate")(self.h)

    def checkConvergence(self):
        return bool(self._f("ss_converged")(self.h))

    # -- periodic boundaries (WebGPUEngine.configurePBC, src/webgpu-engine.mjs:1976-2155)
    def configurePBC(self, cfg):
        axes = cfg.get("axes") or []
        self._f("cfg_pbc")(self.h, len(axes), _u32([a["axis"] for a in axes] or [0]),
                           _f64([a.get("phase", 0) or 0 for a in axes] or [0]))

    def pbcVoltage(self):
        self._f("pbc_voltage")(self.h)

    def pbcCurrent(self):
        self._f("pbc_current")(self.h)

    # -- stepping ----------------------------------------------------------
    def __getattr__(self, name):
        # individual phases: updateVoltages, applyExcitation, murPreVoltage, ...
        table = {"updateVoltages": "update_voltages", "updateCurrents": "update_currents",
                 "applyExcitation": "apply_excitation", "updateVoltADE": "update_volt_ade",
                 "updateCurrADE": "update_curr_ade", "applyTFSFVoltage": "apply_tfsf_voltage",
                 "applyTFSFCurrent": "apply_tfsf_current", "preVoltageRLC": "pre_voltage_rlc",
                 "applyRLC": "apply_rlc", "murPreVoltage": "mur_pre", "murPostVoltage": "mur_post",
                 "murApply": "mur_apply", "step": "step"}
        if name in table:
            fn = self._f(table[name])
            return lambda: fn(self.h)
        if name in ("preVoltageUpdatePML", "postVoltageUpdatePML", "preCurrentUpdatePML",
                    "postCurrentUpdatePML"):
            mode = ("preVoltageUpdatePML", "postVoltageUpdatePML", "preCurrentUpdatePML",
                    "postCurrentUpdatePML").index(name)
            return lambda: self._f("pml_step")(self.h, mode)
        raise AttributeError(name)

    def updateRLC(self):
        self.preVoltageRLC()
        self.applyRLC()

    def iterate(self, n):
        self._f("iterate")(self.h, int(n))


def farfield(accE, accH, pointMeta, theta, phi, center, frequency, radius, fp64=True, mirror=None):
    """Far-field per (theta, phi): 5 values [P_rad, Et_re, Et_im, Ep_re, Ep_im].
    mirror: {"type": "PEC"|"PMC", "direction": 0..2, "position": p,
    "imageNormals": bool} adds the image of the surface in that plane; the
    reflected faces keep their normals (the reference's rule,
    src/nf2ff.mjs:377-398) unless imageNormals is set."""
    lib = _load()
    P = len(pointMeta) // 8
    th, ph = _f64(theta), _f64(phi)
    out = np.zeros(5 * len(th) * len(ph))
    if mirror is None:
        fn = lib.orc64_farfield if fp64 else lib.orc32_farfield
        fn(_f64(accE), _f64(accH), _f64(pointMeta), P, th, len(th), ph, len(ph), _f64(center),
           float(frequency), float(radius), out)
        return out
    fn = lib.orc64_farfield_mirror if fp64 else lib.orc32_farfield_mirror
    fn(_f64(accE), _f64(accH), _f64(pointMeta), P, th, len(th), ph, len(ph), _f64(center), float(frequency),
       float(radius), int(mirror["direction"]), float(mirror.get("position", 0.0)),
       0 if str(mirror.get("type", "PEC")).upper() == "PEC" else 1, 0 if mirror.get("imageNormals") else 1, out)
    return out
