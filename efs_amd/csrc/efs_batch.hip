// efs_batch.hip -- batched form: many independent grids (instances) advanced
// together, one launch per half-step for all of them.
//
// Implements include/efs_batch.h.  Per instance the arithmetic is the
// single-grid engine's (the same e4_march / h4_march device code, the same
// point-kernel formulas) in the same phase order (src/webgpu-fdtd.mjs:
// 1094-1172): Mur pre, E sweep, Mur post, excitation, Mur apply, H sweep,
// [FP16 range control], gather, numTS++.
//
// Layout.  Every instance owns six device arrays (volt, curr, vv, vi, ii,
// iv), NIJK with Nz padded to a multiple of 4 so that every instance runs
// the 4-cells-per-lane sweeps.  Padding cells have zero coefficients (they
// stay zero), and the last real z plane of the H coefficients is set to
// ii = 1, iv = 0, which leaves those cells unchanged exactly as the
// reference's last-plane rule does (src/webgpu-engine.mjs:88); the real
// cells never read a padding cell otherwise (E reads z-1, H reads z+1 only
// below the last real plane).
//
// Launch shape.  A packed table of blocks (instance, column tile, x range)
// covers all instances; each block has about the same cost, so one launch
// balances uneven grids.  The table is ordered by instance, then x, and
// workgroup ids are remapped per XCD, so the blocks an XCD runs are
// neighbours in one instance (shared halo planes in its L2).  The step
// counter lives on the device (excitation and gather read it, the gather
// kernel advances it), so a block of steps is captured once into a HIP
// graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/efs_batch.h"
#include "grid_barrier.hpp"
#include "sweep.hpp"
#include "sweep_z8.hpp"

namespace efs {
namespace batch {

struct Inst {
    void* volt;         // T[3*S]
    void* curr;
    const float* vv;    // float[3*S] each (instances without 16-bit indices)
    const float* vi;
    const float* ii;
    const float* iv;
    const uint16_t* ie; // 16-bit coefficient indices, E and H sides, [3*S] each
    const uint16_t* ih;
    const float2* te;   // their tables, [3][tabn] (vv, vi) and (ii, iv)
    const float2* th;
    uint64_t S;         // component stride
    uint32_t nx, ny, nz;  // nz padded to a multiple of 4
    uint32_t cols;      // ny * nz
    uint32_t tabn_e, tabn_h;  // table rows per component; 0 = FP32 arrays
    ZMur zm;                  // z-face Mur done in the E sweep
};

struct Block {
    uint32_t inst, tile, x0, x1;
};

constexpr int kPt = 256;
constexpr uint32_t kGraphSteps = 32;  // steps per captured graph; also the FP16 range interval
constexpr int kRangeLo = 6, kRangeHi = 12, kRangeTarget = 10, kRangeSibling = 14;

// ------------------------------------------------------------------ sweeps

// One block of the table: the instance's column tile (TPB lanes of Z = 4
// or 8 z cells) marched over [x0, x1).  COH: the scales are read with
// coherent loads (the persistent kernel rewrites them between phases).
template <bool HF, typename T, int TPB, bool COH = false, int Z = 4>
__device__ __forceinline__ void bsweep_block(const Inst* __restrict__ insts, const Block d, const float* fscale) {
    const Inst I = insts[d.inst];
    const uint32_t c0 = (d.tile * TPB + threadIdx.x) * (uint32_t)Z;
    if (c0 >= I.cols || d.x0 >= d.x1) return;
    SweepGeom g{};
    g.nx = I.nx;
    g.ny = I.ny;
    g.nz = I.nz;
    g.cols = I.cols;
    g.S = I.S;
    Lane4 w;
    w.y = c0 / I.nz;
    w.z0 = c0 - w.y * I.nz;
    w.x0 = d.x0;
    w.x1 = d.x1;
    w.region = -1;
    const PmlTable pml{};  // no UPML in the batched form: column mask 0, never read
    const Scales sc = COH ? load_scales_coherent<T>(fscale + 6 * d.inst, HF) : load_scales<T>(fscale + 6 * d.inst, HF);
    if (Z == 8) {
        // rows padded to a multiple of 8 (checked by the host)
        if (HF) {
            if (I.tabn_h) {
                const Coef cf{nullptr, nullptr, nullptr, I.th, nullptr, nullptr, I.ih, I.tabn_h};
                h8_march<kCoefIdx16, T>(static_cast<T*>(I.curr), static_cast<const T*>(I.volt), cf, g, w, sc);
            } else {
                const Coef cf{I.ii, I.iv, nullptr, nullptr, nullptr, nullptr, nullptr, 0u};
                h8_march<kCoefF32, T>(static_cast<T*>(I.curr), static_cast<const T*>(I.volt), cf, g, w, sc);
            }
        } else {
            if (I.tabn_e) {
                const Coef cf{nullptr, nullptr, nullptr, I.te, nullptr, nullptr, I.ie, I.tabn_e};
                e8_march<kCoefIdx16, T>(static_cast<T*>(I.volt), static_cast<const T*>(I.curr), cf, g, w, sc, I.zm);
            } else {
                const Coef cf{I.vv, I.vi, nullptr, nullptr, nullptr, nullptr, nullptr, 0u};
                e8_march<kCoefF32, T>(static_cast<T*>(I.volt), static_cast<const T*>(I.curr), cf, g, w, sc, I.zm);
            }
        }
        return;
    }
    // the coefficient form is per instance (uniform over the workgroup)
    if (HF) {
        if (I.tabn_h) {
            const Coef cf{nullptr, nullptr, nullptr, I.th, nullptr, nullptr, I.ih, I.tabn_h};
            h4_march<kMarchInterior, kCoefIdx16, false, T>(static_cast<T*>(I.curr), static_cast<const T*>(I.volt),
                                                           cf, nullptr, nullptr, g, pml, nullptr, nullptr, w, 0u, sc,
                                                           NfCapture{});
        } else {
            const Coef cf{I.ii, I.iv, nullptr, nullptr, nullptr, nullptr, nullptr, 0u};
            h4_march<kMarchInterior, kCoefF32, false, T>(static_cast<T*>(I.curr), static_cast<const T*>(I.volt), cf,
                                                         nullptr, nullptr, g, pml, nullptr, nullptr, w, 0u, sc,
                                                           NfCapture{});
        }
    } else {
        const ZMur& zm = I.zm;
        if (I.tabn_e) {
            const Coef cf{nullptr, nullptr, nullptr, I.te, nullptr, nullptr, I.ie, I.tabn_e};
            e4_march<kMarchInterior, kCoefIdx16, false, T>(static_cast<T*>(I.volt), static_cast<const T*>(I.curr),
                                                           cf, nullptr, nullptr, g, pml, nullptr, nullptr, w, 0u, sc,
                                                           NfCapture{}, zm);
        } else {
            const Coef cf{I.vv, I.vi, nullptr, nullptr, nullptr, nullptr, nullptr, 0u};
            e4_march<kMarchInterior, kCoefF32, false, T>(static_cast<T*>(I.volt), static_cast<const T*>(I.curr), cf,
                                                         nullptr, nullptr, g, pml, nullptr, nullptr, w, 0u, sc,
                                                         NfCapture{}, zm);
        }
    }
}

// One kernel per (side, storage, occupancy W, tile width TPB, z cells per
// lane Z); the autotuner picks W, TPB and Z.
template <bool HF, typename T, int W, int TPB, int Z>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(W, W))) void bsweep(
    const Inst* __restrict__ insts, const Block* __restrict__ tab, uint32_t nblocks,
    const float* __restrict__ fscale) {
    bsweep_block<HF, T, TPB, false, Z>(insts, tab[xcd_remap(blockIdx.x, nblocks)], fscale);
}

// ------------------------------------------------------------ point kernels

// A point of an instance: its field array through a FieldView of that
// instance (3 components S apart, its 3 scales).
__device__ __forceinline__ FieldView bview(const Inst& I, bool curr, const float* fscale, uint32_t inst,
                                           uint32_t half) {
    return FieldView{curr ? I.curr : I.volt, fscale + 6 * inst + (curr ? 3 : 0), I.S, half};
}

// Mur ABC (src/webgpu-fdtd.mjs:992-1023), the single-grid engine's scheme
// (k_mur_pre / k_mur_post_some / k_mur_finish): one entry per normal cell
// (its last), ordered by (instance, cell); entries whose shifted cell is
// written between post and apply post early (flag kMurEarly).
constexpr uint32_t kMurEarly = 0x80000000u;
struct MurPt {
    uint32_t nidx, sidx;  // normal and shifted cell (device index in the instance)
    float coeff;
    uint32_t inst;        // instance | kMurEarly
};

// Point work, one item per call; the point kernels and the persistent
// kernel share these bodies.
__device__ __forceinline__ void bmur_pre_one(const Inst* __restrict__ insts, const float* fscale, uint32_t half,
                                             const MurPt* __restrict__ pts, float* saved, uint32_t k) {
    const MurPt p = pts[k];
    const uint32_t q = p.inst & ~kMurEarly;
    const FieldView v = bview(insts[q], false, fscale, q, half);
    saved[k] = __builtin_fmaf(-p.coeff, v.ld(p.nidx), v.ld(p.sidx));
}

__device__ __forceinline__ void bmur_post_one(const Inst* __restrict__ insts, const float* fscale, uint32_t half,
                                              const MurPt* __restrict__ pts, const uint32_t* __restrict__ list,
                                              float* saved, uint32_t j) {
    const uint32_t k = list[j];
    const MurPt p = pts[k];
    const uint32_t q = p.inst & ~kMurEarly;
    const FieldView v = bview(insts[q], false, fscale, q, half);
    saved[k] = __builtin_fmaf(p.coeff, v.ld(p.sidx), saved[k]);
}

__device__ __forceinline__ void bmur_finish_one(const Inst* __restrict__ insts, const float* fscale, uint32_t half,
                                                const MurPt* __restrict__ pts, const float* saved, uint32_t k) {
    const MurPt p = pts[k];
    const uint32_t q = p.inst & ~kMurEarly;
    const FieldView v = bview(insts[q], false, fscale, q, half);
    const float s = (p.inst & kMurEarly) ? saved[k] : __builtin_fmaf(p.coeff, v.ld(p.sidx), saved[k]);
    v.st(p.nidx, s);
}

// Excitation (src/webgpu-fdtd.mjs:161-183 semantics, as k_excitation) with
// a signal per instance: point k reads sig[off[k] + e], e < len[k].
struct ExcPoint {
    uint32_t inst, gidx, delay, period;
    uint32_t off, len;
    float amp;
    uint32_t pad;
};

__device__ __forceinline__ void bexc_one(const Inst* __restrict__ insts, const float* fscale, uint32_t half,
                                         const ExcPoint* __restrict__ pts, const float* __restrict__ sig, uint32_t k,
                                         uint32_t t) {
    const ExcPoint p = pts[k];
    int64_t e = (int64_t)t - (int64_t)p.delay;
    if (e < 0) e = 0;
    if (p.period > 0) e %= (int64_t)p.period;
    if (e >= (int64_t)p.len) e = 0;
    bview(insts[p.inst], false, fscale, p.inst, half).add(p.gidx, p.amp * sig[p.off + e]);
}

// Gather entry k at step t (the step's fields; sample slot from the
// interval, as the single-grid buffered gather).
__device__ __forceinline__ void bgather_one(const Inst* __restrict__ insts, const float* fscale, uint32_t half,
                                            const uint32_t* __restrict__ inst, const uint32_t* __restrict__ idx,
                                            uint32_t n, float* ring, uint32_t s, uint32_t k) {
    float* out = ring + (size_t)s * 2 * n;
    const Inst& I = insts[inst[k]];
    out[2 * k] = bview(I, false, fscale, inst[k], half).ld(idx[k]);
    out[2 * k + 1] = bview(I, true, fscale, inst[k], half).ld(idx[k]);
}

__device__ __forceinline__ bool gather_slot(uint32_t t, uint32_t n, uint32_t interval, uint32_t max_samples,
                                            uint32_t& s) {
    if (!n || !interval || (t + 1) % interval != 0) return false;
    s = (t + 1) / interval - 1;
    return s < max_samples;
}

__global__ void k_bmur_pre(const Inst* __restrict__ insts, const float* __restrict__ fscale, uint32_t half,
                           const MurPt* __restrict__ pts, float* __restrict__ saved, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) bmur_pre_one(insts, fscale, half, pts, saved, k);
}

__global__ void k_bmur_post_some(const Inst* __restrict__ insts, const float* __restrict__ fscale, uint32_t half,
                                 const MurPt* __restrict__ pts, const uint32_t* __restrict__ list,
                                 float* __restrict__ saved, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) bmur_post_one(insts, fscale, half, pts, list, saved, j);
}

__global__ void k_bmur_finish(const Inst* __restrict__ insts, const float* __restrict__ fscale, uint32_t half,
                              const MurPt* __restrict__ pts, const float* __restrict__ saved, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) bmur_finish_one(insts, fscale, half, pts, saved, k);
}

__global__ void k_bexcitation(const Inst* __restrict__ insts, const float* __restrict__ fscale, uint32_t half,
                              const ExcPoint* __restrict__ pts, const float* __restrict__ sig, uint32_t n,
                              const uint32_t* __restrict__ ts) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) bexc_one(insts, fscale, half, pts, sig, k, *ts);
}

// Buffered gather of every instance's probes (one workgroup), then the step
// counter advances: every thread reads the counter before the barrier, one
// thread writes it after the second.
__global__ __launch_bounds__(kPt) void k_bgather_tick(const Inst* __restrict__ insts, const float* __restrict__ fscale,
                                                     uint32_t half, const uint32_t* __restrict__ inst,
                                                     const uint32_t* __restrict__ idx, uint32_t n,
                                                     float* __restrict__ ring, uint32_t interval,
                                                     uint32_t max_samples, uint32_t* __restrict__ ts) {
    const uint32_t t = *ts;
    __syncthreads();
    uint32_t s;
    if (gather_slot(t, n, interval, max_samples, s))
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) bgather_one(insts, fscale, half, inst, idx, n, ring, s, k);
    __syncthreads();
    if (threadIdx.x == 0) ts[0] = t + 1;
}

// FP16 range control per instance (the single-grid k_absmax16 /
// k_range_plan / k_range_apply, one segment per instance and component).
struct Seg {
    f16* p;     // first stored value of the component
    uint32_t n; // values (multiple of 4)
    uint32_t slot;  // inst * 6 + k
};

// Largest |stored| of chunk c (of nchunks) of segment s, one atomic per
// workgroup (every thread of the workgroup calls it).
__device__ __forceinline__ void babsmax_item(const Seg* __restrict__ segs, unsigned int* amax, uint32_t seg,
                                             uint32_t c, uint32_t nchunks) {
    const Seg s = segs[seg];
    float m = 0.f;
    for (uint32_t i = 4 * (c * blockDim.x + threadIdx.x); i < s.n; i += 4 * nchunks * blockDim.x) {
        const v4 v = __builtin_convertvector(*reinterpret_cast<const h4*>(s.p + i), v4);
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        if (v.x != v.x || v.y != v.y || v.z != v.z || v.w != v.w) m = __builtin_nanf("");
    }
    // block maximum (waves, then LDS), one atomic per workgroup
    __shared__ float s_m[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_down(m, off, 64);
        m = (o != o || m != m) ? __builtin_nanf("") : fmaxf(m, o);
    }
    if ((threadIdx.x & 63u) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / 64u; ++w) {
            const float o = s_m[w];
            m = (o != o || m != m) ? __builtin_nanf("") : fmaxf(m, o);
        }
        atomicMax(amax + s.slot, __float_as_uint(m));
    }
    __syncthreads();
}

__device__ __forceinline__ void bplan_one(unsigned int* amax, float* fscale, const float* __restrict__ smin,
                                          float* factor, const uint32_t* __restrict__ src_end, uint32_t q,
                                          uint32_t t) {
    const uint32_t active = (uint64_t)t + 1 < (uint64_t)src_end[q];
    unsigned int* am = amax + 6 * q;
    float* sc = fscale + 6 * q;
    const float* sm = smin + 6 * q;
    float* fa = factor + 6 * q;
    float m[6], arr[2] = {0.f, 0.f};
    for (int k = 0; k < 6; ++k) {
        m[k] = __uint_as_float(am[k]);
        am[k] = 0u;
        if (m[k] < __builtin_huge_valf()) arr[k / 3] = fmaxf(arr[k / 3], m[k] * sc[k]);
    }
    for (int k = 0; k < 6; ++k) {
        float f = 1.0f;
        if (m[k] > 0.0f && m[k] < __builtin_huge_valf()) {
            const int e = ilogbf(m[k]);
            if (e < kRangeLo || e >= kRangeHi) {
                int shift = kRangeTarget - e;
                if (shift > 0) {
                    const int floor_e = active ? ilogbf(sm[k]) : ilogbf(arr[k / 3]) - kRangeSibling;
                    shift = min(shift, max(0, ilogbf(sc[k]) - floor_e));
                }
                f = ldexpf(1.0f, shift);
                sc[k] = ldexpf(sc[k], -shift);
            }
        }
        fa[k] = f;
    }
}

__device__ __forceinline__ void bapply_item(const Seg* __restrict__ segs, const float* factor, uint32_t seg,
                                            uint32_t c, uint32_t nchunks) {
    const Seg s = segs[seg];
    const float f = factor[s.slot];
    if (f == 1.0f) return;
    for (uint32_t i = 4 * (c * blockDim.x + threadIdx.x); i < s.n; i += 4 * nchunks * blockDim.x) {
        h4* q = reinterpret_cast<h4*>(s.p + i);
        *q = __builtin_convertvector(clamp16(__builtin_convertvector(*q, v4) * f), h4);
    }
}

__global__ __launch_bounds__(256) void k_babsmax(const Seg* __restrict__ segs, unsigned int* __restrict__ amax) {
    babsmax_item(segs, amax, blockIdx.y, blockIdx.x, gridDim.x);
}

__global__ void k_bplan(unsigned int* __restrict__ amax, float* __restrict__ fscale,
                        const float* __restrict__ smin, float* __restrict__ factor,
                        const uint32_t* __restrict__ src_end, uint32_t ninst, const uint32_t* __restrict__ ts) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < ninst) bplan_one(amax, fscale, smin, factor, src_end, q, *ts);
}

__global__ __launch_bounds__(256) void k_bapply(const Seg* __restrict__ segs, const float* __restrict__ factor) {
    bapply_item(segs, factor, blockIdx.y, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------- persistent kernel

// Everything one step reads and writes, for the persistent kernel.
struct PArgs {
    const Inst* insts;
    const Block* tab;
    uint32_t nblocks;
    uint32_t half;
    float* fscale;
    // Mur
    const MurPt* mur;
    const uint32_t* mur_early;
    float* mur_saved;
    uint32_t mur_n, mur_ne;
    // excitation
    const ExcPoint* exc;
    const float* sig;
    uint32_t exc_n;
    // gather
    uint32_t g_n;
    const uint32_t* g_inst;
    const uint32_t* g_idx;
    float* ring;
    uint32_t g_interval, g_max;
    // FP16 range control
    const Seg* segs;
    uint32_t nsegs, seg_chunks;
    unsigned int* amax;
    const float* smin;
    float* factor;
    const uint32_t* src_end;
    uint32_t ninst;
    // step counter, barrier, steps of this launch
    uint32_t* ts;
    GridBarState* bar;
    uint32_t flat;
    uint32_t nsteps;
};

// nsteps steps of every instance in one launch: the phases of a step
// (Mur pre, E sweep, early Mur post, excitation, Mur post + apply, H sweep,
// [FP16 range control], gather) run in the order of the launch form, each
// over the whole grid of resident workgroups, separated by grid barriers.
// The next step's Mur pre shares the gather's phase (both only read the
// fields).  Sweep blocks are dealt to workgroups round-robin through the
// same XCD-contiguous table order as the launch form.  Results are the
// launch form's bits.  On a barrier timeout every workgroup leaves; the
// error word tells the host.
template <typename T, int W>
__global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(W, W))) void bpersist(PArgs a) {
    GridBar gb;
    if (!gb.init(a.bar, a.flat != 0)) return;
    const uint32_t G = gridDim.x, wg = blockIdx.x;
    const uint32_t gtid = wg * blockDim.x + threadIdx.x, gthreads = G * blockDim.x;
    const uint32_t t0 = __hip_atomic_load(a.ts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t half = a.half;
    for (uint32_t k = gtid; k < a.mur_n; k += gthreads) bmur_pre_one(a.insts, a.fscale, half, a.mur, a.mur_saved, k);
    if (!gb.sync()) return;
    for (uint32_t s = 0; s < a.nsteps; ++s) {
        const uint32_t t = t0 + s;
        for (uint32_t j = wg; j < a.nblocks; j += G)
            bsweep_block<false, T, kSweepThreads, true>(a.insts, a.tab[xcd_remap(j, a.nblocks)], a.fscale);
        if (!gb.sync()) return;
        if (a.mur_ne) {
            for (uint32_t j = gtid; j < a.mur_ne; j += gthreads)
                bmur_post_one(a.insts, a.fscale, half, a.mur, a.mur_early, a.mur_saved, j);
            if (!gb.sync()) return;
        }
        if (a.exc_n) {
            for (uint32_t k = gtid; k < a.exc_n; k += gthreads) bexc_one(a.insts, a.fscale, half, a.exc, a.sig, k, t);
            if (!gb.sync()) return;
        }
        if (a.mur_n) {
            for (uint32_t k = gtid; k < a.mur_n; k += gthreads)
                bmur_finish_one(a.insts, a.fscale, half, a.mur, a.mur_saved, k);
            if (!gb.sync()) return;
        }
        for (uint32_t j = wg; j < a.nblocks; j += G)
            bsweep_block<true, T, kSweepThreads, true>(a.insts, a.tab[xcd_remap(j, a.nblocks)], a.fscale);
        if (!gb.sync()) return;
        if (half && (t + 1) % kGraphSteps == 0) {
            const uint32_t items = a.nsegs * a.seg_chunks;
            for (uint32_t it = wg; it < items; it += G) babsmax_item(a.segs, a.amax, it / a.seg_chunks, it % a.seg_chunks, a.seg_chunks);
            if (!gb.sync()) return;
            for (uint32_t q = gtid; q < a.ninst; q += gthreads) bplan_one(a.amax, a.fscale, a.smin, a.factor, a.src_end, q, t);
            if (!gb.sync()) return;
            for (uint32_t it = wg; it < items; it += G) bapply_item(a.segs, a.factor, it / a.seg_chunks, it % a.seg_chunks, a.seg_chunks);
            if (!gb.sync()) return;
        }
        uint32_t slot;
        if (gather_slot(t, a.g_n, a.g_interval, a.g_max, slot))
            for (uint32_t k = gtid; k < a.g_n; k += gthreads)
                bgather_one(a.insts, a.fscale, half, a.g_inst, a.g_idx, a.g_n, a.ring, slot, k);
        if (s + 1 < a.nsteps) {
            for (uint32_t k = gtid; k < a.mur_n; k += gthreads)
                bmur_pre_one(a.insts, a.fscale, half, a.mur, a.mur_saved, k);
            if (!gb.sync()) return;
        }
    }
    if (gtid == 0) __hip_atomic_store(a.ts, t0 + a.nsteps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Occupancy of the persistent kernel: both marches and the point work live
// in one kernel, which needs up to 168 registers without spilling (at 4
// waves per SIMD, 128, it spills).
constexpr uint32_t kPersistWaves = 3;
template __global__ void bpersist<float, kPersistWaves>(PArgs);
template __global__ void bpersist<f16, kPersistWaves>(PArgs);

}  // namespace batch
}  // namespace efs

using namespace efs;
using namespace efs::batch;

namespace {

std::string g_batch_create_error;

struct HostInst {
    uint32_t nx, ny, nz, nzp;
    uint64_t N, S;
    void *volt = nullptr, *curr = nullptr;
    float *vv = nullptr, *vi = nullptr, *ii = nullptr, *iv = nullptr;
    uint16_t *ie = nullptr, *ih = nullptr;
    float2 *te = nullptr, *th = nullptr;
    uint32_t tabn_e = 0, tabn_h = 0;
    ZMur zm;  // z-face Mur of the E sweep (set at finalize)
    // configured point lists (instance-local device indices)
    std::vector<float> sig;
    std::vector<ExcPoint> exc;  // off relative to this instance's signal
    uint32_t exc_period = 0;
    std::vector<float> mur_c;
    std::vector<uint32_t> mur_n, mur_s;
    std::vector<uint32_t> gat;
    float scale0[6] = {1, 1, 1, 1, 1, 1};  // FP16 initial scales (also the range floor while the source runs)
    uint32_t src_end = 0;
};

}  // namespace

struct efs_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool final = false;
    bool half = false;
    std::vector<HostInst> inst;
    uint64_t cells = 0;
    uint32_t numTS = 0;

    Inst* d_inst = nullptr;
    Block* d_tab = nullptr;
    uint32_t nblocks = 0;
    uint32_t want = 2048;  // blocks asked of the planner
    uint32_t waves = 4;
    uint32_t tpb = kSweepThreads;  // lanes per column tile (launch form)
    uint32_t zc = 4;               // z cells per lane (launch form: 4, or 8 when every row is padded to 8)
    // Rows are padded to a multiple of zpad cells: 8 (default) lets the
    // sweeps run 8-cell lanes (16-byte FP16 accesses) and was faster for
    // 4-cell lanes too (64-instance design sweep: FP32 47.3k -> 48.5k,
    // FP16 68.4k -> 71.8k MCells/s); EFS_BATCH_ZPAD=4 pads to 4.
    uint32_t zpad = 8;
    int mode = EFS_BATCH_LAUNCH;
    uint32_t pgrid = 0;            // workgroups of the persistent kernel
    GridBarState* d_bar = nullptr;
    bool bar_flat = false;
    bool persist_ran = false;
    uint32_t* d_ts = nullptr;
    float* d_fscale = nullptr;
    float* d_smin = nullptr;
    float* d_factor = nullptr;
    unsigned int* d_amax = nullptr;
    uint32_t* d_src_end = nullptr;
    Seg* d_segs = nullptr;
    uint32_t nsegs = 0;
    uint32_t seg_chunks = 1;

    // Mur
    uint32_t mur_n = 0, mur_ne = 0;
    uint32_t mur_swept = 0;  // z-face entries the E sweep applies
    MurPt* d_mur = nullptr;
    uint32_t* d_mur_early = nullptr;
    float* d_mur_saved = nullptr;
    // excitation
    uint32_t exc_n = 0;
    ExcPoint* d_exc = nullptr;
    float* d_sig = nullptr;
    // gather
    uint32_t g_n = 0, g_interval = 0, g_max = 0;
    std::vector<uint32_t> g_off;  // per instance: first gather entry
    uint32_t *d_g_inst = nullptr, *d_g_idx = nullptr;
    float* d_ring = nullptr;

    // graph of kGraphSteps steps
    hipGraphExec_t graph = nullptr;
    bool graph_ok = false;

    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    struct Pend {
        int side;
        hipEvent_t a, b;
    };
    std::vector<Pend> pend;
    double prof_ms[2] = {0, 0};
    uint32_t prof_n[2] = {0, 0};
};

namespace {

int bfail(efs_batch* b, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    else g_batch_create_error = buf;
    return code;
}

#define BHIP(b, expr)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return bfail(b, EFS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
#define BTRY(expr)               \
    do {                         \
        int rc_ = (expr);        \
        if (rc_) return rc_;     \
    } while (0)

template <class T>
int balloc(efs_batch* b, T*& p, size_t n) {
    if (p) hipFree(p);
    p = nullptr;
    hipError_t e = hipMalloc((void**)&p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        return bfail(b, e == hipErrorOutOfMemory ? EFS_ERR_NOMEM : EFS_ERR_HIP, "hipMalloc(%zu bytes): %s",
                     (n ? n : 1) * sizeof(T), hipGetErrorString(e));
    }
    return EFS_OK;
}

template <class T>
int bupload(efs_batch* b, T*& p, const T* src, size_t n) {
    BTRY(balloc(b, p, n));
    if (n) BHIP(b, hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    return EFS_OK;
}

template <class T>
void bfree(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

uint32_t nblk(size_t n) { return (uint32_t)((n + kPt - 1) / kPt); }

// Instance-local interface index (component*N + x*Ny*Nz + y*Nz + z) ->
// device index (component*S + x*Ny*Nzp + y*Nzp + z).
uint32_t btr(const HostInst& h, uint64_t g) {
    const uint64_t n = g / h.N, r = g - n * h.N;
    const uint64_t z = r % h.nz, xy = r / h.nz;
    return (uint32_t)(n * h.S + xy * h.nzp + z);
}

// Equal-cost block table: every (instance, column tile of tpb lanes) is cut
// along x into chunks of about total/want cells, ordered by instance, then x.
std::vector<Block> plan_blocks(const efs_batch* b, uint32_t want, uint32_t tpb, uint32_t zc) {
    double total = 0.0;
    for (const HostInst& h : b->inst) total += (double)h.nx * h.ny * h.nzp;
    const double target = total / std::max(1u, want);
    std::vector<Block> tab;
    for (uint32_t q = 0; q < (uint32_t)b->inst.size(); ++q) {
        const HostInst& h = b->inst[q];
        const uint64_t cols = (uint64_t)h.ny * h.nzp;
        const uint32_t tcells = zc * tpb;
        const uint32_t T = (uint32_t)((cols + tcells - 1) / tcells);
        std::vector<Block> mine;
        for (uint32_t t = 0; t < T; ++t) {
            const double cells = (double)(std::min<uint64_t>(cols, (uint64_t)(t + 1) * tcells) - (uint64_t)t * tcells);
            const double cost = cells * h.nx;
            const uint32_t n = (uint32_t)std::max(1.0, std::min((double)std::max(1u, h.nx / 2),
                                                                std::floor(cost / target + 0.5)));
            for (uint32_t j = 0; j < n; ++j) {
                const uint32_t x0 = (uint32_t)((uint64_t)h.nx * j / n), x1 = (uint32_t)((uint64_t)h.nx * (j + 1) / n);
                if (x1 > x0) mine.push_back(Block{q, t, x0, x1});
            }
        }
        std::stable_sort(mine.begin(), mine.end(), [](const Block& a, const Block& c) {
            return a.x0 + a.x1 != c.x0 + c.x1 ? a.x0 + a.x1 < c.x0 + c.x1 : a.tile < c.tile;
        });
        tab.insert(tab.end(), mine.begin(), mine.end());
    }
    return tab;
}

void drop_graph(efs_batch* b);

// (Re)build the block table; a captured graph holds the old table, so it is
// dropped (build_graph captures a new one).
int set_blocks(efs_batch* b, uint32_t want) {
    drop_graph(b);
    b->want = want;
    // the persistent kernel runs 256-lane tiles
    const bool pers = b->mode == EFS_BATCH_PERSISTENT;
    std::vector<Block> tab = plan_blocks(b, want, pers ? (uint32_t)kSweepThreads : b->tpb, pers ? 4u : b->zc);
    BHIP(b, hipStreamSynchronize(b->stream));
    BTRY(bupload(b, b->d_tab, tab.data(), tab.size()));
    b->nblocks = (uint32_t)tab.size();
    return EFS_OK;
}

hipEvent_t take_event(efs_batch* b) {
    if (!b->ev_pool.empty()) {
        hipEvent_t e = b->ev_pool.back();
        b->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

typedef void (*SweepFn)(const Inst*, const Block*, uint32_t, const float*);

template <bool HF, typename T>
SweepFn sweep_kernel_t(uint32_t waves, uint32_t tpb, uint32_t z) {
    // 8-cell lanes (rows padded to 8): 3 waves per SIMD, 147-165 registers
    // (at 4 waves they spill and run 30 % slower)
    if (z == 8) return tpb == 512 ? bsweep<HF, T, 3, 512, 8> : bsweep<HF, T, 3, 256, 8>;
    if (tpb == 128) return waves == 5 ? bsweep<HF, T, 5, 128, 4> : bsweep<HF, T, 4, 128, 4>;
    if (tpb == 512) return waves == 5 ? bsweep<HF, T, 5, 512, 4> : bsweep<HF, T, 4, 512, 4>;
    return waves == 5 ? bsweep<HF, T, 5, 256, 4> : bsweep<HF, T, 4, 256, 4>;
}

SweepFn sweep_kernel(bool hf, bool half, uint32_t waves, uint32_t tpb, uint32_t z) {
    if (hf) return half ? sweep_kernel_t<true, f16>(waves, tpb, z) : sweep_kernel_t<true, float>(waves, tpb, z);
    return half ? sweep_kernel_t<false, f16>(waves, tpb, z) : sweep_kernel_t<false, float>(waves, tpb, z);
}

bool valid_tpb(uint32_t t) { return t == 128 || t == 256 || t == 512; }

void launch_sweep(efs_batch* b, bool hf) {
    hipEvent_t ea = nullptr;
    if (b->prof) {
        ea = take_event(b);
        if (ea) hipEventRecord(ea, b->stream);
    }
    hipLaunchKernelGGL(sweep_kernel(hf, b->half, b->waves, b->tpb, b->zc), dim3(b->nblocks), dim3(b->tpb), 0, b->stream,
                       (const Inst*)b->d_inst, (const Block*)b->d_tab, b->nblocks, (const float*)b->d_fscale);
    if (b->prof && ea) {
        hipEvent_t eb = take_event(b);
        if (eb) {
            hipEventRecord(eb, b->stream);
            b->pend.push_back({hf ? 1 : 0, ea, eb});
        } else {
            b->ev_pool.push_back(ea);
        }
    }
}

// One step; `range` adds the FP16 range control before the gather (steps
// with (t + 1) % kGraphSteps == 0).
void launch_step(efs_batch* b, bool range) {
    const uint32_t half = b->half ? 1u : 0u;
    if (b->mur_n)
        hipLaunchKernelGGL(k_bmur_pre, dim3(nblk(b->mur_n)), dim3(kPt), 0, b->stream, (const Inst*)b->d_inst,
                           (const float*)b->d_fscale, half, (const MurPt*)b->d_mur, b->d_mur_saved, b->mur_n);
    launch_sweep(b, false);
    if (b->mur_ne)
        hipLaunchKernelGGL(k_bmur_post_some, dim3(nblk(b->mur_ne)), dim3(kPt), 0, b->stream, (const Inst*)b->d_inst,
                           (const float*)b->d_fscale, half, (const MurPt*)b->d_mur, (const uint32_t*)b->d_mur_early,
                           b->d_mur_saved, b->mur_ne);
    if (b->exc_n)
        hipLaunchKernelGGL(k_bexcitation, dim3(nblk(b->exc_n)), dim3(kPt), 0, b->stream, (const Inst*)b->d_inst,
                           (const float*)b->d_fscale, half, (const ExcPoint*)b->d_exc, (const float*)b->d_sig,
                           b->exc_n, (const uint32_t*)b->d_ts);
    if (b->mur_n)
        hipLaunchKernelGGL(k_bmur_finish, dim3(nblk(b->mur_n)), dim3(kPt), 0, b->stream, (const Inst*)b->d_inst,
                           (const float*)b->d_fscale, half, (const MurPt*)b->d_mur, (const float*)b->d_mur_saved,
                           b->mur_n);
    launch_sweep(b, true);
    if (b->half && range) {
        const dim3 grid(b->seg_chunks, b->nsegs);
        hipLaunchKernelGGL(k_babsmax, grid, dim3(256), 0, b->stream, (const Seg*)b->d_segs, b->d_amax);
        hipLaunchKernelGGL(k_bplan, dim3(nblk(b->inst.size())), dim3(kPt), 0, b->stream, b->d_amax, b->d_fscale,
                           (const float*)b->d_smin, b->d_factor, (const uint32_t*)b->d_src_end,
                           (uint32_t)b->inst.size(), (const uint32_t*)b->d_ts);
        hipLaunchKernelGGL(k_bapply, grid, dim3(256), 0, b->stream, (const Seg*)b->d_segs,
                           (const float*)b->d_factor);
    }
    hipLaunchKernelGGL(k_bgather_tick, dim3(1), dim3(kPt), 0, b->stream, (const Inst*)b->d_inst,
                       (const float*)b->d_fscale, half, (const uint32_t*)b->d_g_inst, (const uint32_t*)b->d_g_idx,
                       b->g_n, b->d_ring, b->g_interval, b->g_max, b->d_ts);
}

void drop_graph(efs_batch* b) {
    if (b->graph) hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_ok = false;
}

// Capture kGraphSteps steps (the last one with the FP16 range control): a
// replay starting at a multiple of kGraphSteps is exactly those steps.
int build_graph(efs_batch* b) {
    drop_graph(b);
    if (b->mode == EFS_BATCH_PERSISTENT) return EFS_OK;  // one launch already covers many steps
    // EFS_BATCH_GRAPH=0: direct launches only (tools that cannot follow
    // graph replays)
    if (const char* e = std::getenv("EFS_BATCH_GRAPH"))
        if (e[0] == '0') return EFS_OK;
    BHIP(b, hipStreamSynchronize(b->stream));
    hipGraph_t g = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const bool prof = b->prof;
    b->prof = false;
    for (uint32_t s = 0; s < kGraphSteps; ++s) launch_step(b, s + 1 == kGraphSteps);
    b->prof = prof;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) return bfail(b, EFS_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return bfail(b, EFS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    b->graph_ok = true;
    return EFS_OK;
}

// Steps of one persistent launch at most (the host regains control between
// launches; the FP16 range control runs inside at every 32nd step).
constexpr uint32_t kPersistSteps = 64;

// Workgroups of the persistent kernel: every one must be resident at once
// (occupancy of the kernel x compute units).
int persist_grid(efs_batch* b) {
    const void* f = b->half ? reinterpret_cast<const void*>(bpersist<f16, kPersistWaves>)
                            : reinterpret_cast<const void*>(bpersist<float, kPersistWaves>);
    int per_cu = 0, cus = 0, coop = 0;
    BHIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kSweepThreads, 0));
    BHIP(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    BHIP(b, hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, b->device));
    if (!coop) return bfail(b, EFS_ERR_HIP, "device %d has no cooperative launch", b->device);
    if (per_cu <= 0) return bfail(b, EFS_ERR_HIP, "persistent kernel does not fit a compute unit");
    b->pgrid = (uint32_t)(per_cu * cus) & ~7u;  // a multiple of the 8 XCDs
    if (b->pgrid == 0) return bfail(b, EFS_ERR_HIP, "persistent grid is empty");
    if (!b->d_bar) BTRY(balloc(b, b->d_bar, 1));
    return EFS_OK;
}

int launch_persistent(efs_batch* b, uint32_t nsteps) {
    PArgs a{};
    a.insts = b->d_inst;
    a.tab = b->d_tab;
    a.nblocks = b->nblocks;
    a.half = b->half ? 1u : 0u;
    a.fscale = b->d_fscale;
    a.mur = b->d_mur;
    a.mur_early = b->d_mur_early;
    a.mur_saved = b->d_mur_saved;
    a.mur_n = b->mur_n;
    a.mur_ne = b->mur_ne;
    a.exc = b->d_exc;
    a.sig = b->d_sig;
    a.exc_n = b->exc_n;
    a.g_n = b->g_n;
    a.g_inst = b->d_g_inst;
    a.g_idx = b->d_g_idx;
    a.ring = b->d_ring;
    a.g_interval = b->g_interval;
    a.g_max = b->g_max;
    a.segs = b->d_segs;
    a.nsegs = b->half ? b->nsegs : 0u;
    a.seg_chunks = b->seg_chunks;
    a.amax = b->d_amax;
    a.smin = b->d_smin;
    a.factor = b->d_factor;
    a.src_end = b->d_src_end;
    a.ninst = (uint32_t)b->inst.size();
    a.ts = b->d_ts;
    a.bar = b->d_bar;
    a.flat = b->bar_flat ? 1u : 0u;
    a.nsteps = nsteps;
    hipLaunchKernelGGL(k_gridbar_clear, dim3(1), dim3(256), 0, b->stream, reinterpret_cast<uint32_t*>(b->d_bar));
    BHIP(b, hipGetLastError());
    void* args[] = {&a};
    const void* f = b->half ? reinterpret_cast<const void*>(bpersist<f16, kPersistWaves>)
                            : reinterpret_cast<const void*>(bpersist<float, kPersistWaves>);
    BHIP(b, hipLaunchCooperativeKernel(f, dim3(b->pgrid), dim3(kSweepThreads), args, 0, b->stream));
    b->persist_ran = true;
    return EFS_OK;
}

// After persistent launches: a barrier that timed out (grid not resident)
// is reported here (the stream is synchronised).
int check_persistent(efs_batch* b) {
    if (!b->persist_ran) return EFS_OK;
    BHIP(b, hipStreamSynchronize(b->stream));
    uint32_t err = 0;
    BHIP(b, hipMemcpy(&err, reinterpret_cast<const uint32_t*>(b->d_bar) + offsetof(GridBarState, err) / 4,
                      sizeof err, hipMemcpyDeviceToHost));
    b->persist_ran = false;
    if (err) return bfail(b, EFS_ERR_HIP, "persistent kernel: a grid barrier timed out (workgroups not co-resident)");
    return EFS_OK;
}

int run_steps(efs_batch* b, uint32_t n) {
    uint32_t done = 0;
    if (b->mode == EFS_BATCH_PERSISTENT) {
        while (done < n) {
            const uint32_t k = std::min(kPersistSteps, n - done);
            BTRY(launch_persistent(b, k));
            b->numTS += k;
            done += k;
        }
        return EFS_OK;
    }
    while (done < n) {
        const uint32_t t = b->numTS;
        if (b->graph_ok && !b->prof && t % kGraphSteps == 0 && n - done >= kGraphSteps) {
            BHIP(b, hipGraphLaunch(b->graph, b->stream));
            b->numTS += kGraphSteps;
            done += kGraphSteps;
        } else {
            launch_step(b, (t + 1) % kGraphSteps == 0);
            BHIP(b, hipGetLastError());
            b->numTS++;
            done++;
        }
    }
    return EFS_OK;
}

// Distinct (a, b) pairs of each of the 3 components (S values apart):
// 16-bit index per value and a [3][tabn] table (tabn = the largest count,
// rounded up to 64).  False when a component has more than 65536 pairs.
bool pair_index16(const float* a, const float* b, uint64_t S, std::vector<uint16_t>& idx, std::vector<float2>& tab,
                  uint32_t& tabn) {
    constexpr uint32_t H = 1u << 17;  // open-addressing slots (load <= 1/2)
    std::vector<std::vector<float2>> per(3);
    idx.assign(3 * S, 0);
    std::vector<uint64_t> key(H);
    std::vector<int32_t> val(H);
    for (uint32_t n = 0; n < 3; ++n) {
        std::fill(val.begin(), val.end(), -1);
        uint64_t last_k = ~0ull;
        uint16_t last_v = 0;
        for (uint64_t i = n * S; i < (n + 1) * S; ++i) {
            uint32_t ua, ub;
            std::memcpy(&ua, a + i, 4);
            std::memcpy(&ub, b + i, 4);
            const uint64_t k = ((uint64_t)ua << 32) | ub;
            if (k == last_k) {
                idx[i] = last_v;
                continue;
            }
            uint32_t hsh = (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> 47) & (H - 1);
            while (val[hsh] >= 0 && key[hsh] != k) hsh = (hsh + 1) & (H - 1);
            if (val[hsh] < 0) {
                if (per[n].size() >= 65536) return false;
                key[hsh] = k;
                val[hsh] = (int32_t)per[n].size();
                per[n].push_back(float2{a[i], b[i]});
            }
            last_k = k;
            last_v = (uint16_t)val[hsh];
            idx[i] = last_v;
        }
    }
    size_t m = 1;
    for (auto& v : per) m = std::max(m, v.size());
    tabn = (uint32_t)((m + 63) & ~(size_t)63);
    tab.assign(3 * (size_t)tabn, float2{0.f, 0.f});
    for (uint32_t n = 0; n < 3; ++n) std::copy(per[n].begin(), per[n].end(), tab.begin() + n * tabn);
    return true;
}

int need_open(efs_batch* b) {
    if (!b) return EFS_ERR_STATE;
    if (b->final) return bfail(b, EFS_ERR_STATE, "batch already finalized");
    return EFS_OK;
}

int need_final(efs_batch* b) {
    if (!b) return EFS_ERR_STATE;
    if (!b->final) return bfail(b, EFS_ERR_STATE, "batch not finalized: call efs_batch_finalize() first");
    return EFS_OK;
}

int need_inst(efs_batch* b, uint32_t id) {
    if (id >= b->inst.size()) return bfail(b, EFS_ERR_ARG, "instance %u does not exist (%zu added)", id, b->inst.size());
    return EFS_OK;
}

// Initial FP16 scales of one instance from its source (the single-grid
// engine's rule: the largest injected value lands near 2^kRangeTarget).
void source_scales(HostInst& h) {
    float smax = 0.f, va = 0.f;
    for (float s : h.sig) smax = std::max(smax, std::fabs(s));
    for (const ExcPoint& p : h.exc) va = std::max(va, std::fabs(p.amp));
    const float mv = smax * va;
    if (!(mv > 0.f && std::isfinite(mv))) return;
    const float vs = std::ldexp(1.0f, std::ilogb(mv) - kRangeTarget);
    for (int k = 0; k < 3; ++k) {
        h.scale0[k] = vs;
        h.scale0[3 + k] = std::ldexp(vs, -8);
    }
}

int zero_state(efs_batch* b) {
    BHIP(b, hipStreamSynchronize(b->stream));
    for (HostInst& h : b->inst) {
        const size_t fb = (b->half ? 2 : 4) * 3 * h.S;
        BHIP(b, hipMemset(h.volt, 0, fb));
        BHIP(b, hipMemset(h.curr, 0, fb));
    }
    if (b->mur_n) BHIP(b, hipMemset(b->d_mur_saved, 0, b->mur_n * sizeof(float)));
    if (b->g_n && b->g_max) BHIP(b, hipMemset(b->d_ring, 0, (size_t)b->g_max * 2 * b->g_n * sizeof(float)));
    BHIP(b, hipMemset(b->d_ts, 0, sizeof(uint32_t)));
    std::vector<float> sc(6 * b->inst.size());
    for (size_t q = 0; q < b->inst.size(); ++q)
        for (int k = 0; k < 6; ++k) sc[6 * q + k] = b->inst[q].scale0[k];
    BHIP(b, hipMemcpy(b->d_fscale, sc.data(), sc.size() * sizeof(float), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_smin, sc.data(), sc.size() * sizeof(float), hipMemcpyHostToDevice));
    BHIP(b, hipMemset(b->d_amax, 0, sc.size() * sizeof(unsigned int)));
    b->numTS = 0;
    return EFS_OK;
}

int drain(efs_batch* b) {
    if (b->pend.empty()) return EFS_OK;
    BHIP(b, hipStreamSynchronize(b->stream));
    for (auto& p : b->pend) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            b->prof_ms[p.side] += ms;
            b->prof_n[p.side]++;
        }
        b->ev_pool.push_back(p.a);
        b->ev_pool.push_back(p.b);
    }
    b->pend.clear();
    return EFS_OK;
}

}  // namespace

// ===================================================================== C ABI

extern "C" {

int efs_batch_create(efs_batch** out, int device) {
    if (!out) return bfail(nullptr, EFS_ERR_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return bfail(nullptr, EFS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return bfail(nullptr, EFS_ERR_ARG, "device %d out of range (%d devices)", device, n);
    efs_batch* b = new efs_batch();
    b->device = device;
    if (const char* e = std::getenv("EFS_BATCH_ZPAD")) b->zpad = std::atoi(e) == 4 ? 4u : 8u;
    {
        // grid-barrier form of the persistent kernel: "flat" = a write-back
        // per workgroup (the reference form, for measurement)
        const char* e = std::getenv("EFS_BATCH_BARRIER");
        b->bar_flat = e && std::strcmp(e, "flat") == 0;
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return bfail(nullptr, EFS_ERR_HIP, "stream creation failed");
    }
    *out = b;
    return EFS_OK;
}

void efs_batch_destroy(efs_batch* b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (HostInst& h : b->inst) {
        bfree(h.volt);
        bfree(h.curr);
        bfree(h.vv);
        bfree(h.vi);
        bfree(h.ii);
        bfree(h.iv);
        bfree(h.ie);
        bfree(h.ih);
        bfree(h.te);
        bfree(h.th);
    }
    bfree(b->d_inst);
    bfree(b->d_tab);
    bfree(b->d_ts);
    bfree(b->d_fscale);
    bfree(b->d_smin);
    bfree(b->d_factor);
    bfree(b->d_amax);
    bfree(b->d_src_end);
    bfree(b->d_segs);
    bfree(b->d_mur);
    bfree(b->d_mur_early);
    bfree(b->d_mur_saved);
    bfree(b->d_exc);
    bfree(b->d_sig);
    bfree(b->d_g_inst);
    bfree(b->d_g_idx);
    bfree(b->d_ring);
    bfree(b->d_bar);
    for (auto& p : b->pend) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (hipEvent_t e : b->ev_pool) hipEventDestroy(e);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

const char* efs_batch_last_error(const efs_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int efs_batch_set_storage(efs_batch* b, int storage) {
    BTRY(need_open(b));
    if (storage != EFS_STORAGE_F32 && storage != EFS_STORAGE_F16)
        return bfail(b, EFS_ERR_ARG, "unknown storage %d", storage);
    if (!b->inst.empty()) return bfail(b, EFS_ERR_STATE, "set the storage before adding instances");
    b->half = storage == EFS_STORAGE_F16;
    return EFS_OK;
}

int efs_batch_add(efs_batch* b, uint32_t nx, uint32_t ny, uint32_t nz, const float* vv, const float* vi,
                  const float* ii, const float* iv, uint32_t* id) {
    BTRY(need_open(b));
    if (nx == 0 || ny == 0 || nz == 0) return bfail(b, EFS_ERR_ARG, "grid size must be positive");
    if (!vv || !vi || !ii || !iv) return bfail(b, EFS_ERR_ARG, "coefficient arrays are required");
    HostInst h;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.nzp = (nz + b->zpad - 1) / b->zpad * b->zpad;
    h.N = (uint64_t)nx * ny * nz;
    const uint64_t Np = (uint64_t)nx * ny * h.nzp;
    h.S = ((Np + 255) & ~(uint64_t)255) + 1088;
    if (3 * h.S >= (1ull << 32)) return bfail(b, EFS_ERR_ARG, "instance too large for 32-bit indices");
    if ((uint64_t)ny * h.nzp >= (1ull << 32) / 4) return bfail(b, EFS_ERR_ARG, "plane too large");
    BHIP(b, hipSetDevice(b->device));
    // padded copies: pad cells zero; H last real z plane ii = 1, iv = 0
    std::vector<float> pa(3 * h.S), pb(3 * h.S);
    for (int side = 0; side < 2; ++side) {
        const float* sa = side ? ii : vv;
        const float* sb = side ? iv : vi;
        std::fill(pa.begin(), pa.end(), 0.0f);
        std::fill(pb.begin(), pb.end(), 0.0f);
        for (uint32_t n = 0; n < 3; ++n)
            for (uint64_t xy = 0; xy < (uint64_t)nx * ny; ++xy) {
                std::memcpy(pa.data() + n * h.S + xy * h.nzp, sa + n * h.N + xy * nz, nz * sizeof(float));
                std::memcpy(pb.data() + n * h.S + xy * h.nzp, sb + n * h.N + xy * nz, nz * sizeof(float));
                if (side) {
                    pa[n * h.S + xy * h.nzp + nz - 1] = 1.0f;
                    pb[n * h.S + xy * h.nzp + nz - 1] = 0.0f;
                }
            }
        // 16-bit index form when every component has at most 65536 distinct
        // (a, b) pairs, else the FP32 arrays
        std::vector<uint16_t> idx;
        std::vector<float2> tab;
        uint32_t tabn = 0;
        int rc = EFS_OK;
        if (pair_index16(pa.data(), pb.data(), h.S, idx, tab, tabn)) {
            uint16_t*& di = side ? h.ih : h.ie;
            float2*& dt = side ? h.th : h.te;
            rc = bupload(b, di, idx.data(), idx.size());
            if (!rc) rc = bupload(b, dt, tab.data(), tab.size());
            (side ? h.tabn_h : h.tabn_e) = tabn;
        } else {
            float*& da = side ? h.ii : h.vv;
            float*& db = side ? h.iv : h.vi;
            rc = bupload(b, da, pa.data(), pa.size());
            if (!rc) rc = bupload(b, db, pb.data(), pb.size());
        }
        if (rc) {
            for (float** q : {&h.vv, &h.vi, &h.ii, &h.iv}) bfree(*q);
            bfree(h.ie);
            bfree(h.ih);
            bfree(h.te);
            bfree(h.th);
            return rc;
        }
    }
    const size_t fb = (b->half ? 2 : 4) * 3 * h.S;
    hipError_t e1 = hipMalloc(&h.volt, fb), e2 = hipMalloc(&h.curr, fb);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        bfree(h.volt);
        bfree(h.curr);
        for (float** q : {&h.vv, &h.vi, &h.ii, &h.iv}) bfree(*q);
        bfree(h.ie);
        bfree(h.ih);
        bfree(h.te);
        bfree(h.th);
        return bfail(b, EFS_ERR_NOMEM, "field allocation failed");
    }
    b->inst.push_back(h);
    b->cells += h.N;
    if (id) *id = (uint32_t)b->inst.size() - 1;
    return EFS_OK;
}

int efs_batch_cfg_excitation(efs_batch* b, uint32_t id, const float* signal, uint32_t len, const uint32_t* delay,
                             const float* amp, const uint32_t* dir, const uint32_t* pos, uint32_t n,
                             uint32_t period) {
    BTRY(need_open(b));
    BTRY(need_inst(b, id));
    HostInst& h = b->inst[id];
    if (n == 0) {
        h.exc.clear();
        h.sig.clear();
        return EFS_OK;
    }
    if (!signal || len == 0) return bfail(b, EFS_ERR_ARG, "excitation signal must not be empty");
    if (!delay || !amp || !dir || !pos) return bfail(b, EFS_ERR_ARG, "excitation arrays are required");
    std::vector<ExcPoint> pts(n);
    for (uint32_t k = 0; k < n; ++k) {
        if (dir[k] > 2) return bfail(b, EFS_ERR_ARG, "excitation %u: dir %u out of range", k, dir[k]);
        if (pos[k] >= h.N) return bfail(b, EFS_ERR_ARG, "excitation %u: pos %u out of range", k, pos[k]);
        pts[k] = ExcPoint{id, btr(h, (uint64_t)dir[k] * h.N + pos[k]), delay[k], period, 0, len, amp[k], 0};
    }
    h.sig.assign(signal, signal + len);
    h.exc = pts;
    h.exc_period = period;
    // first step at which no sample is injected any more (the FP16 range
    // floor follows the source until then)
    uint64_t end = 0;
    if (period == 0)
        for (uint32_t k = 0; k < n; ++k) end = std::max<uint64_t>(end, (uint64_t)delay[k] + len);
    else
        end = 0xffffffffu;
    h.src_end = (uint32_t)std::min<uint64_t>(end, 0xffffffffu);
    return EFS_OK;
}

int efs_batch_cfg_mur(efs_batch* b, uint32_t id, uint32_t n, const float* coeff, const uint32_t* normal_idx,
                      const uint32_t* shifted_idx) {
    BTRY(need_open(b));
    BTRY(need_inst(b, id));
    HostInst& h = b->inst[id];
    if (n && (!coeff || !normal_idx || !shifted_idx)) return bfail(b, EFS_ERR_ARG, "Mur arrays are required");
    std::vector<uint32_t> ni(n), si(n);
    for (uint32_t k = 0; k < n; ++k) {
        if (normal_idx[k] >= 3 * h.N || shifted_idx[k] >= 3 * h.N)
            return bfail(b, EFS_ERR_ARG, "Mur point %u: index out of range", k);
        if (normal_idx[k] / h.N != shifted_idx[k] / h.N)
            return bfail(b, EFS_ERR_ARG, "Mur point %u: normal and shifted indices of different components", k);
        ni[k] = btr(h, normal_idx[k]);
        si[k] = btr(h, shifted_idx[k]);
    }
    h.mur_c.assign(coeff, coeff + n);
    h.mur_n = ni;
    h.mur_s = si;
    return EFS_OK;
}

int efs_batch_cfg_gather(efs_batch* b, uint32_t id, const uint32_t* idx, uint32_t n) {
    BTRY(need_open(b));
    BTRY(need_inst(b, id));
    HostInst& h = b->inst[id];
    if (n && !idx) return bfail(b, EFS_ERR_ARG, "gather indices are required");
    std::vector<uint32_t> g(n);
    for (uint32_t k = 0; k < n; ++k) {
        if (idx[k] >= 3 * h.N) return bfail(b, EFS_ERR_ARG, "gather index %u out of range", idx[k]);
        g[k] = btr(h, idx[k]);
    }
    h.gat = g;
    return EFS_OK;
}

struct MurKeep {
    uint32_t k;
    bool early;
    bool fused = false;
};

// Which kept Mur entries of an instance the E sweep applies itself (ZMur):
// tangential x / y entries of the z = 0 face (shifted cell z = 1) and of the
// z = Nz-1 face (shifted cell z = Nz-2, in the same 4-cell lane: (Nz-1) % 4
// != 0; Nz >= 6), not posting early, whose normal cell is neither an excitation cell
// nor another kept entry's shifted cell.  Nothing writes those cells between
// the sweep and the apply kernel, so the sweep's fused update gives the
// list's numbers.  A (face, component) slot goes to the sweep only when its
// candidates share one coefficient and fill an x-y rectangle exactly (the
// usual whole-face lists do); otherwise its entries stay in the list.
// EFS_BATCH_ZMUR=0 keeps every entry in the list.
void zmur_split(HostInst& h, std::vector<MurKeep>& keep, bool on) {
    h.zm = ZMur{};
    if (!on || keep.empty() || h.nz < 6) return;
    std::vector<uint32_t> shifted, exc;
    for (const MurKeep& e : keep) shifted.push_back(h.mur_s[e.k]);
    for (const ExcPoint& p : h.exc) exc.push_back(p.gidx);
    std::sort(shifted.begin(), shifted.end());
    std::sort(exc.begin(), exc.end());
    std::vector<uint32_t> cand[4];  // positions in keep
    std::vector<uint64_t> row(keep.size());
    for (uint32_t j = 0; j < (uint32_t)keep.size(); ++j) {
        const MurKeep& e = keep[j];
        if (e.early) continue;
        const uint32_t ni = h.mur_n[e.k], si = h.mur_s[e.k];
        const uint64_t n = ni / h.S, r = ni - n * h.S, z = r % h.nzp, xy = r / h.nzp;
        const uint64_t rs = si - n * h.S;  // same component (checked at configure)
        if (n > 1 || rs / h.nzp != xy) continue;
        const uint64_t zs = rs % h.nzp;
        int f = -1;
        if (z == 0 && zs == 1)
            f = 0;
        else if (z == h.nz - 1 && zs == h.nz - 2 && (h.nz - 1) % 4 != 0)
            f = 1;
        if (f < 0) continue;
        if (std::binary_search(shifted.begin(), shifted.end(), ni) || std::binary_search(exc.begin(), exc.end(), ni))
            continue;
        row[j] = xy;
        cand[2 * f + n].push_back(j);
    }
    for (int s = 0; s < 4; ++s) {
        if (cand[s].empty()) continue;
        const float c = h.mur_c[keep[cand[s][0]].k];
        int64_t x0 = INT64_MAX, x1 = -1, y0 = INT64_MAX, y1 = -1;
        bool same = true;
        for (uint32_t j : cand[s]) {
            const float cj = h.mur_c[keep[j].k];
            same = same && std::memcmp(&cj, &c, sizeof c) == 0;
            const int64_t x = (int64_t)(row[j] / h.ny), y = (int64_t)(row[j] % h.ny);
            x0 = std::min(x0, x);
            x1 = std::max(x1, x);
            y0 = std::min(y0, y);
            y1 = std::max(y1, y);
        }
        // normal cells are distinct, so a count equal to the area fills it
        if (!same || (uint64_t)(x1 - x0 + 1) * (uint64_t)(y1 - y0 + 1) != cand[s].size()) continue;
        h.zm.on = 1;
        h.zm.c[s] = c;
        h.zm.x0[s] = (int32_t)x0;
        h.zm.x1[s] = (int32_t)x1;
        h.zm.y0[s] = (int32_t)y0;
        h.zm.y1[s] = (int32_t)y1;
        for (uint32_t j : cand[s]) keep[j].fused = true;
    }
    h.zm.nzr = h.nz;
}

int efs_batch_finalize(efs_batch* b, uint32_t sample_interval, uint32_t max_samples) {
    BTRY(need_open(b));
    if (b->inst.empty()) return bfail(b, EFS_ERR_STATE, "no instance added");
    BHIP(b, hipSetDevice(b->device));
    const uint32_t ni = (uint32_t)b->inst.size();
    std::vector<Inst> it(ni);
    std::vector<uint32_t> src_end(ni);
    std::vector<Seg> segs;
    std::vector<uint32_t> gi, gx;
    std::vector<MurPt> mur;
    std::vector<uint32_t> mur_early;
    std::vector<float> sig;
    std::vector<ExcPoint> exc;
    const char* zenv = std::getenv("EFS_BATCH_ZMUR");
    const bool zmur = !(zenv && zenv[0] == '0');
    uint32_t swept = 0;
    b->g_off.assign(ni + 1, 0);
    for (uint32_t q = 0; q < ni; ++q) {
        HostInst& h = b->inst[q];

        if (b->half) source_scales(h);
        src_end[q] = h.src_end;
        for (uint32_t k = 0; k < 6; ++k)
            segs.push_back(Seg{static_cast<f16*>(k < 3 ? h.volt : h.curr) + (k % 3) * h.S,
                               (uint32_t)((uint64_t)h.nx * h.ny * h.nzp), 6 * q + k});
        {
            // one entry per normal cell (its last), ordered by the cell; an
            // entry posts early when its shifted cell is another entry's
            // normal cell or an excitation cell
            const uint32_t nm = (uint32_t)h.mur_c.size();
            std::vector<uint32_t> ord(nm);
            for (uint32_t k = 0; k < nm; ++k) ord[k] = k;
            std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) { return h.mur_n[a] < h.mur_n[c]; });
            std::vector<uint32_t> written;
            std::vector<MurKeep> keep;
            for (uint32_t j = 0; j < nm; ++j) written.push_back(h.mur_n[ord[j]]);
            for (const ExcPoint& e : h.exc) written.push_back(e.gidx);
            std::sort(written.begin(), written.end());
            for (uint32_t j = 0; j < nm; ++j) {
                const uint32_t k = ord[j];
                if (j + 1 < nm && h.mur_n[ord[j + 1]] == h.mur_n[k]) continue;
                const bool early = std::binary_search(written.begin(), written.end(), h.mur_s[k]);
                keep.push_back({k, early});
            }
            zmur_split(h, keep, zmur);
            for (const auto& e : keep) swept += e.fused ? 1u : 0u;
            for (const auto& e : keep) {
                if (e.fused) continue;
                if (e.early) mur_early.push_back((uint32_t)mur.size());
                mur.push_back(MurPt{h.mur_n[e.k], h.mur_s[e.k], h.mur_c[e.k], q | (e.early ? kMurEarly : 0u)});
            }
        }
        it[q] = Inst{h.volt, h.curr, h.vv, h.vi, h.ii, h.iv, h.ie, h.ih, h.te, h.th,
                     h.S, h.nx, h.ny, h.nzp, h.ny * h.nzp, h.tabn_e, h.tabn_h,
                     h.zm};
        const uint32_t off = (uint32_t)sig.size();
        sig.insert(sig.end(), h.sig.begin(), h.sig.end());
        for (ExcPoint p : h.exc) {
            p.off = off;
            exc.push_back(p);
        }
        b->g_off[q] = (uint32_t)gx.size();
        for (uint32_t v : h.gat) {
            gi.push_back(q);
            gx.push_back(v);
        }
    }
    b->g_off[ni] = (uint32_t)gx.size();
    if (sig.size() >= (1ull << 32)) return bfail(b, EFS_ERR_ARG, "excitation signals too long");
    BTRY(bupload(b, b->d_inst, it.data(), it.size()));
    BTRY(bupload(b, b->d_src_end, src_end.data(), src_end.size()));
    BTRY(balloc(b, b->d_ts, 1));
    BTRY(balloc(b, b->d_fscale, 6 * ni));
    BTRY(balloc(b, b->d_smin, 6 * ni));
    BTRY(balloc(b, b->d_factor, 6 * ni));
    BTRY(balloc(b, b->d_amax, 6 * ni));
    if (b->half) {
        BTRY(bupload(b, b->d_segs, segs.data(), segs.size()));
        b->nsegs = (uint32_t)segs.size();
        uint64_t mx = 0;
        for (const Seg& s : segs) mx = std::max<uint64_t>(mx, s.n);
        b->seg_chunks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(16, (mx / 4 + 255) / 256));
    }
    b->mur_n = (uint32_t)mur.size();
    b->mur_ne = (uint32_t)mur_early.size();
    b->mur_swept = swept;
    if (b->mur_n) {
        BTRY(bupload(b, b->d_mur, mur.data(), mur.size()));
        BTRY(bupload(b, b->d_mur_early, mur_early.data(), mur_early.size()));
        BTRY(balloc(b, b->d_mur_saved, mur.size()));
    }
    b->exc_n = (uint32_t)exc.size();
    if (b->exc_n) {
        BTRY(bupload(b, b->d_exc, exc.data(), exc.size()));
        BTRY(bupload(b, b->d_sig, sig.data(), sig.size()));
    }
    b->g_n = (uint32_t)gx.size();
    b->g_interval = sample_interval;
    b->g_max = b->g_n && sample_interval ? max_samples : 0;
    if (b->g_n) {
        BTRY(bupload(b, b->d_g_inst, gi.data(), gi.size()));
        BTRY(bupload(b, b->d_g_idx, gx.data(), gx.size()));
        BTRY(balloc(b, b->d_ring, (size_t)std::max(1u, b->g_max) * 2 * b->g_n));
    }
    BTRY(set_blocks(b, 2048));
    b->final = true;
    BTRY(zero_state(b));
    return build_graph(b);
}

int efs_batch_reset(efs_batch* b) {
    BTRY(need_final(b));
    BHIP(b, hipSetDevice(b->device));
    return zero_state(b);
}

int efs_batch_tune(efs_batch* b, uint32_t steps, uint32_t* blocks_out, uint32_t* waves_out) {
    BTRY(need_final(b));
    BHIP(b, hipSetDevice(b->device));
    steps = std::max<uint32_t>(kGraphSteps, (steps + kGraphSteps - 1) / kGraphSteps * kGraphSteps);
    struct Cand {
        int mode;
        uint32_t tpb, zc, waves, want;
    };
    std::vector<Cand> cand;
    for (uint32_t tpb : {128u, 256u, 512u})
        for (uint32_t v : {4u, 5u})
            for (uint32_t want : {1024u, 2048u, 4096u, 8192u}) cand.push_back({EFS_BATCH_LAUNCH, tpb, 4u, v, want});
    if (b->zpad == 8)
        for (uint32_t tpb : {256u, 512u})
            for (uint32_t want : {1024u, 2048u, 4096u, 8192u}) cand.push_back({EFS_BATCH_LAUNCH, tpb, 8u, 3u, want});
    // the persistent form is a candidate unless EFS_BATCH_PERSIST=0 (a
    // profiler whose tool cannot follow cooperative launches)
    const char* pe = std::getenv("EFS_BATCH_PERSIST");
    const int mode0 = b->mode;
    b->mode = EFS_BATCH_PERSISTENT;
    const bool can_persist = !(pe && pe[0] == '0') && persist_grid(b) == EFS_OK;
    b->mode = mode0;
    b->err.clear();
    if (can_persist)
        for (uint32_t m : {1u, 2u, 4u, 8u}) cand.push_back({EFS_BATCH_PERSISTENT, 256u, 4u, kPersistWaves, m * b->pgrid});
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    double best = 1e300;
    Cand pick{EFS_BATCH_LAUNCH, kSweepThreads, 4u, 4u, 2048u};
    const bool prof = b->prof;
    b->prof = false;
    int rc = EFS_OK;
    for (const Cand& c : cand) {
        b->mode = c.mode;
        b->tpb = c.tpb;
        b->zc = c.zc;
        b->waves = c.waves;
        if ((rc = set_blocks(b, c.want)) || (rc = zero_state(b)) || (rc = build_graph(b))) break;
        if ((rc = run_steps(b, kGraphSteps))) break;  // warm
        hipEventRecord(e0, b->stream);
        if ((rc = run_steps(b, steps))) break;
        hipEventRecord(e1, b->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = bfail(b, EFS_ERR_HIP, "tuner: timing failed");
            break;
        }
        if (check_persistent(b) != EFS_OK) {  // a shape that cannot run (not resident): skip it
            b->err.clear();
            continue;
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    b->prof = prof;
    if (rc) return rc;
    b->prof_ms[0] = b->prof_ms[1] = 0.0;
    b->prof_n[0] = b->prof_n[1] = 0;
    b->mode = pick.mode;
    b->tpb = pick.tpb;
    b->zc = pick.zc;
    b->waves = pick.waves;
    BTRY(set_blocks(b, pick.want));
    BTRY(zero_state(b));
    BTRY(build_graph(b));
    if (blocks_out) *blocks_out = b->nblocks;
    if (waves_out) *waves_out = b->waves;
    return EFS_OK;
}

int efs_batch_set_shape(efs_batch* b, int mode, uint32_t tile_threads, uint32_t z_cells, uint32_t waves,
                        uint32_t blocks) {
    BTRY(need_final(b));
    BHIP(b, hipSetDevice(b->device));
    if (mode != EFS_BATCH_LAUNCH && mode != EFS_BATCH_PERSISTENT) return bfail(b, EFS_ERR_ARG, "unknown mode %d", mode);
    if (mode == EFS_BATCH_PERSISTENT) {
        if (tile_threads != kSweepThreads || z_cells != 4 || waves != kPersistWaves)
            return bfail(b, EFS_ERR_ARG,
                         "the persistent kernel runs 256-lane tiles of 4-cell lanes at %u waves per SIMD (got %u, %u, %u)",
                         kPersistWaves, tile_threads, z_cells, waves);
    } else if (z_cells == 4) {
        if (!valid_tpb(tile_threads))
            return bfail(b, EFS_ERR_ARG, "tile_threads must be 128, 256 or 512 (got %u)", tile_threads);
        if (waves != 4 && waves != 5) return bfail(b, EFS_ERR_ARG, "waves must be 4 or 5 (got %u)", waves);
    } else if (z_cells == 8) {
        if (b->zpad != 8)
            return bfail(b, EFS_ERR_ARG, "8-cell lanes need rows padded to 8 cells (EFS_BATCH_ZPAD=4 was set)");
        if (tile_threads != 256 && tile_threads != 512)
            return bfail(b, EFS_ERR_ARG, "8-cell lanes run 256- or 512-lane tiles (got %u)", tile_threads);
        if (waves != 3) return bfail(b, EFS_ERR_ARG, "8-cell lanes run 3 waves per SIMD (got %u)", waves);
    } else {
        return bfail(b, EFS_ERR_ARG, "z_cells must be 4 or 8 (got %u)", z_cells);
    }
    if (blocks == 0) return bfail(b, EFS_ERR_ARG, "blocks must be positive");
    BTRY(check_persistent(b));
    const int mode0 = b->mode;
    b->mode = mode;
    if (mode == EFS_BATCH_PERSISTENT) {
        const int rc = persist_grid(b);
        if (rc) {
            b->mode = mode0;
            return rc;
        }
    }
    b->tpb = tile_threads;
    b->zc = z_cells;
    b->waves = waves;
    BTRY(set_blocks(b, blocks));
    return build_graph(b);
}

int efs_batch_get_shape(efs_batch* b, int* mode, uint32_t* tile_threads, uint32_t* z_cells, uint32_t* waves,
                        uint32_t* blocks, uint32_t* grid) {
    BTRY(need_final(b));
    const bool pers = b->mode == EFS_BATCH_PERSISTENT;
    if (mode) *mode = b->mode;
    if (tile_threads) *tile_threads = pers ? (uint32_t)kSweepThreads : b->tpb;
    if (z_cells) *z_cells = pers ? 4u : b->zc;
    if (waves) *waves = b->waves;
    if (blocks) *blocks = b->nblocks;
    if (grid) *grid = pers ? b->pgrid : b->nblocks;
    return EFS_OK;
}

int efs_batch_iterate(efs_batch* b, uint32_t n) {
    BTRY(need_final(b));
    BHIP(b, hipSetDevice(b->device));
    return run_steps(b, n);
}

int efs_batch_sync(efs_batch* b) {
    if (!b) return EFS_ERR_STATE;
    BHIP(b, hipStreamSynchronize(b->stream));
    return check_persistent(b);
}

int efs_batch_num_ts(efs_batch* b, uint32_t* out) {
    if (!b || !out) return EFS_ERR_ARG;
    *out = b->numTS;
    return EFS_OK;
}

int efs_batch_read_gather(efs_batch* b, uint32_t id, float* out, uint32_t* num_samples) {
    BTRY(need_final(b));
    BTRY(need_inst(b, id));
    const uint32_t ns = b->g_interval ? std::min(b->g_max, b->numTS / b->g_interval) : 0;
    if (num_samples) *num_samples = ns;
    if (!out || ns == 0) return EFS_OK;
    const uint32_t k0 = b->g_off[id], k1 = b->g_off[id + 1];
    const uint32_t w = k1 - k0;
    if (w == 0) return EFS_OK;
    BHIP(b, hipSetDevice(b->device));
    BTRY(check_persistent(b));
    BHIP(b, hipStreamSynchronize(b->stream));
    // one strided copy: sample rows of 2*g_n floats, this instance's columns
    BHIP(b, hipMemcpy2D(out, 2 * w * sizeof(float), b->d_ring + 2 * k0, 2 * (size_t)b->g_n * sizeof(float),
                        2 * w * sizeof(float), ns, hipMemcpyDeviceToHost));
    return EFS_OK;
}

int efs_batch_sizes(efs_batch* b, uint32_t id, uint32_t* dims, uint32_t* gather_n) {
    if (!b) return EFS_ERR_STATE;
    BTRY(need_inst(b, id));
    const HostInst& h = b->inst[id];
    if (dims) {
        dims[0] = h.nx;
        dims[1] = h.ny;
        dims[2] = h.nz;
    }
    if (gather_n) *gather_n = (uint32_t)h.gat.size();
    return EFS_OK;
}

int efs_batch_get_fields(efs_batch* b, uint32_t id, float* volt, float* curr) {
    BTRY(need_final(b));
    BTRY(need_inst(b, id));
    if (!volt || !curr) return bfail(b, EFS_ERR_ARG, "output arrays are required");
    BHIP(b, hipSetDevice(b->device));
    BTRY(check_persistent(b));
    BHIP(b, hipStreamSynchronize(b->stream));
    const HostInst& h = b->inst[id];
    float sc[6] = {1, 1, 1, 1, 1, 1};
    if (b->half) BHIP(b, hipMemcpy(sc, b->d_fscale + 6 * id, sizeof sc, hipMemcpyDeviceToHost));
    for (int a = 0; a < 2; ++a) {
        float* dst = a ? curr : volt;
        const void* src = a ? h.curr : h.volt;
        if (b->half) {
            std::vector<_Float16> tmp(3 * h.S);
            BHIP(b, hipMemcpy(tmp.data(), src, tmp.size() * 2, hipMemcpyDeviceToHost));
            for (uint32_t n = 0; n < 3; ++n)
                for (uint64_t xy = 0; xy < (uint64_t)h.nx * h.ny; ++xy)
                    for (uint32_t z = 0; z < h.nz; ++z)
                        dst[n * h.N + xy * h.nz + z] = (float)tmp[n * h.S + xy * h.nzp + z] * sc[3 * a + n];
        } else {
            std::vector<float> tmp(3 * h.S);
            BHIP(b, hipMemcpy(tmp.data(), src, tmp.size() * 4, hipMemcpyDeviceToHost));
            for (uint32_t n = 0; n < 3; ++n)
                for (uint64_t xy = 0; xy < (uint64_t)h.nx * h.ny; ++xy)
                    std::memcpy(dst + n * h.N + xy * h.nz, tmp.data() + n * h.S + xy * h.nzp, h.nz * sizeof(float));
        }
    }
    return EFS_OK;
}

int efs_batch_info(efs_batch* b, uint32_t* blocks, uint32_t* waves, uint32_t* graph_steps, uint64_t* cells) {
    if (!b) return EFS_ERR_STATE;
    if (blocks) *blocks = b->nblocks;
    if (waves) *waves = b->waves;
    if (graph_steps) *graph_steps = b->graph_ok ? kGraphSteps : 0;
    if (cells) *cells = b->cells;
    return EFS_OK;
}

int efs_batch_coef_form(efs_batch* b, uint32_t id, uint32_t* idx16) {
    if (!b) return EFS_ERR_STATE;
    BTRY(need_inst(b, id));
    if (idx16) *idx16 = (b->inst[id].tabn_e && b->inst[id].tabn_h) ? 1u : 0u;
    return EFS_OK;
}

int efs_batch_mur_split(efs_batch* b, uint32_t* listed, uint32_t* in_sweep) {
    BTRY(need_final(b));
    if (listed) *listed = b->mur_n;
    if (in_sweep) *in_sweep = b->mur_swept;
    return EFS_OK;
}

int efs_batch_profile(efs_batch* b, int enable) {
    if (!b) return EFS_ERR_STATE;
    BTRY(drain(b));
    b->prof = enable != 0;
    b->prof_ms[0] = b->prof_ms[1] = 0.0;
    b->prof_n[0] = b->prof_n[1] = 0;
    return EFS_OK;
}

int efs_batch_profile_read(efs_batch* b, double* ms_e, uint32_t* n_e, double* ms_h, uint32_t* n_h) {
    if (!b) return EFS_ERR_STATE;
    BTRY(drain(b));
    if (ms_e) *ms_e = b->prof_ms[0];
    if (n_e) *n_e = b->prof_n[0];
    if (ms_h) *ms_h = b->prof_ms[1];
    if (n_h) *n_h = b->prof_n[1];
    return EFS_OK;
}

}  // extern "C"
