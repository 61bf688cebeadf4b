"""efs_farfield_noalloc: the far field with nothing allocated (device
inputs and outputs from the caller, partial sums in the library's static
per-GPU scratch).  Called through the C ABI, its output must equal
efs_farfield's on the same accumulators to the FP32 tolerance (relative L2
<= 1e-5, the engine's bar for floating-point results), on two calls in a
row -- the second with fewer angles, so the scratch holds the first call's
partials when it starts -- and a request beyond the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class Dev:
    """Device buffers through the HIP runtime the engine library uses."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host -> device
        self.bufs.append(p)
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n, dtype=np.float32):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device -> host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def test_farfield_noalloc_matches_farfield():
    from efs_amd import engine, operator as op
    lib = engine.load_library()
    n = 24
    mesh = op.uniform_mesh((n, n, n), 1e-3)
    dT = mesh.timestep()
    mat = op.Materials(mesh)
    op.pec_boundary(mat)
    port, _ = op.lumped_port(mesh, mat, 2, (12, 12, 11), (12, 12, 13), 50.0, excite=100.0)
    co = op.to_f32(op.build_coefficients(mesh, mat, dT))
    exc = {"signal": op.gaussian_pulse(20e9, 15e9, dT), "delay": [0] * len(port["pos"]), "amp": port["amp"],
           "dir": port["dir"], "pos": port["pos"], "period": 0}
    pts, meta, primal, dual = op.nf2ff_box(mesh, [4] * 3, [19] * 3)
    freqs = np.array([12e9, 18e9])
    g = engine.HIPEngine()
    assert g.initGPU()
    g.init([n] * 3, co)
    g.configureExcitation(exc)
    g.configureNF2FFAccumulation({"surfaceIndices": pts, "omega": list(2 * np.pi * freqs), "dT": dT,
                                  "primalEdgeLens": primal, "dualEdgeLens": dual})
    g.iterate(150)  # leaves a partly filled K-step block: both entry points fold it first
    d = Dev()
    try:
        meta_d = d.put(np.asarray(meta, np.float32).ravel())
        center = np.array([12e-3, 12e-3, 12e-3])
        for call, (nt, nph, fi) in enumerate(((37, 73, 1), (9, 12, 0))):  # second call: fewer angles
            theta = np.linspace(0, np.pi, nt).astype(np.float32)
            phi = np.linspace(0, 2 * np.pi, nph, endpoint=False).astype(np.float32)
            say("call", call, "angles", nt * nph, "freq", fi)
            want = g.computeNF2FFfarField({"pointMeta": meta, "theta": theta, "phi": phi, "center": center,
                                           "frequency": float(freqs[fi]), "radius": 1.0, "freqIndex": fi})
            th_d, ph_d, out_d = d.put(theta), d.put(phi), d.empty(5 * nt * nph * 4)
            rc = lib.efs_farfield_noalloc(g.ctx, fi, meta_d, th_d, nt, ph_d, nph, center, float(freqs[fi]), 1.0,
                                          out_d)
            assert rc == 0, lib.efs_last_error(g.ctx)
            assert lib.efs_sync(g.ctx) == 0
            got = d.get(out_d, 5 * nt * nph)
            err = rel_l2(got, want)
            say("relative L2 vs efs_farfield", err)
            assert np.all(np.isfinite(got)) and err <= 1e-5
        # beyond the scratch: refused with a message, nothing launched
        big = 65537
        th_d, ph_d, out_d = d.put(np.zeros(big, np.float32)), d.put(np.zeros(1, np.float32)), d.empty(5 * big * 4)
        rc = lib.efs_farfield_noalloc(g.ctx, 0, meta_d, th_d, big, ph_d, 1, center, 12e9, 1.0, out_d)
        assert rc == engine.EFS_ERR_ARG
        assert b"scratch" in lib.efs_last_error(g.ctx)
    finally:
        g.sync()
        d.free()
        g.destroy()
