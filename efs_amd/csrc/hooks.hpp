// hooks.hpp -- the per-step point work of one phase in one launch.
//
// Each point kernel (Mur, excitation, probe gather, NF2FF sampling) moves a
// few kilobytes but costs a full launch on the GPU: on a grid of half a
// million cells the six of them took more time than the two sweeps.  The
// kernels here run the work that sits between two sweeps in one launch,
// in an order that keeps every result bit-identical to the separate
// kernels in the reference's phase order (src/webgpu-engine.mjs:2407-2504).
//
// After the E sweep (k_post_e, then k_post_e_b when its lists are not
// empty): Mur post + apply and the excitation.  The reference order is post
// (all entries), excitation, apply (all entries).  Phase A (k_post_e) runs,
// one thread per Mur entry and per excitation entry, every operation that
// no other operation of the phase reads or writes:
//  - Mur entry k forms s = saved[k] + c*V[shifted] (the post) from its
//    shifted cell, which nothing writes in this phase (an entry whose
//    shifted cell is some normal or excitation cell is "early", and every
//    cell an early entry reads is deferred), and applies V[normal] = s
//    unless its normal cell is deferred; a deferred entry keeps s in
//    saved[k];
//  - an excitation entry adds its sample unless its cell is deferred or a
//    Mur normal cell: there the apply overwrites the sum, so the entry has
//    no effect and is dropped.
// Phase B (k_post_e_b, its own launch, so phase A is complete everywhere)
// adds the deferred excitation entries and applies V[normal] = saved[k]
// for the deferred Mur entries; the two lists touch disjoint cells.  The
// host builds the lists; a box of Mur faces defers its edges.
//
// Group form (default): the first grp workgroups of k_post_e do the early
// and the deferred Mur entries and the deferred excitation entries, and
// there is no phase B launch.  The host links each early entry with the
// entries that write its shifted cell and packs the linked sets whole into
// groups of at most kPostThreads Mur entries (one per thread).  A group
// reads the shifted cells of its Mur entries, waits at a workgroup
// barrier, then writes their normal cells and adds its excitation entries.
// Nothing outside a group reads or writes those cells: an entry reading a
// cell that another entry writes is early, so the writer's cell is
// deferred and both are in the same group.
//
// After the H sweep (k_post_h): NF2FF sampling of the step, the probe
// gather and the next step's Mur pre pass.  All three only read the
// fields (the pre pass reads E, which nothing writes between the H sweep
// and the next E sweep), so they run side by side on disjoint workgroups.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nf2ff.hpp"
#include "sparse.hpp"

namespace efs {

constexpr int kPostThreads = 256;

template <class R>
struct PostE {
    // excitation (k_excitation's arguments)
    const float* signal;
    uint32_t len, period;
    const uint32_t *delay, *gidx;
    const float* amp;
    uint32_t exc_n, numTS;
    const uint8_t* exc_defer;   // 1: not in phase A (deferred or dropped)
    const uint32_t* exc_b;      // deferred excitation entries
    uint32_t exc_nb;
    // Mur (device indices)
    const float* coeff;
    const uint32_t *normal, *shifted;
    R* saved;
    uint32_t mur_n;
    const uint8_t* mur_defer;   // 1: phase B (or the group) applies it
    const uint2* mur_b;         // deferred Mur entries: (normal cell, k)
    uint32_t mur_nb;
    // group form: Mur entry of each thread of group g at grp_mur[g * 256 +
    // thread] (~0u: none), excitation entries grp_exc[grp_eoff[g] ..
    // grp_eoff[g + 1])
    const uint32_t *grp_mur, *grp_exc, *grp_eoff;
    uint32_t grp;               // workgroups of the group form (0: none)
};

template <class FV>
__global__ __launch_bounds__(kPostThreads) void k_post_e(FV volt, PostE<typename FV::R> q) {
    typedef typename FV::R R;
    if (blockIdx.x < q.grp) {
        const uint32_t k = q.grp_mur[blockIdx.x * kPostThreads + threadIdx.x];
        R s = 0;
        if (k != ~0u) s = fma1((R)q.coeff[k], volt.ld(q.shifted[k]), q.saved[k]);
        __syncthreads();  // every read above precedes every write below
        if (k != ~0u) volt.st(q.normal[k], s);
        for (uint32_t j = q.grp_eoff[blockIdx.x] + threadIdx.x; j < q.grp_eoff[blockIdx.x + 1]; j += kPostThreads)
            excitation_one(volt, q.signal, q.len, q.period, q.delay, q.amp, q.gidx, q.grp_exc[j], q.numTS);
        return;
    }
    const uint32_t i = (blockIdx.x - q.grp) * blockDim.x + threadIdx.x;
    if (i < q.mur_n) {
        if (q.grp && q.mur_defer[i]) return;  // the group's entry (it reads saved[i])
        const R s = fma1((R)q.coeff[i], volt.ld(q.shifted[i]), q.saved[i]);
        if (q.mur_defer[i]) q.saved[i] = s;
        else volt.st(q.normal[i], s);
    } else if (i - q.mur_n < q.exc_n) {
        const uint32_t k = i - q.mur_n;
        if (!q.exc_defer[k]) excitation_one(volt, q.signal, q.len, q.period, q.delay, q.amp, q.gidx, k, q.numTS);
    }
}

template <class FV>
__global__ __launch_bounds__(kPostThreads) void k_post_e_b(FV volt, PostE<typename FV::R> q) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < q.exc_nb) {
        excitation_one(volt, q.signal, q.len, q.period, q.delay, q.amp, q.gidx, q.exc_b[j], q.numTS);
    } else if (j - q.exc_nb < q.mur_nb) {
        const uint2 e = q.mur_b[j - q.exc_nb];
        volt.st(e.x, q.saved[e.y]);
    }
}

template <class FV>
struct PostH {
    typedef typename FV::R R;
    // NF2FF sampling (k_nf2ff_sample's arguments), workgroups [0, nb_sample)
    const uint32_t* pts;
    const float* el;
    Nf2ffGeom g;
    R* row_out;
    NfCapture cap;
    uint32_t nb_sample;
    // probe gather, the next nb_gather workgroups
    const uint32_t* gidx;
    float* gout;
    uint32_t gn, nb_gather;
    // the next step's Mur pre pass, the next nb_pre workgroups
    const R* vflux;
    const float* coeff;
    const uint32_t *normal_src, *shifted_src;
    R* saved;
    uint32_t mur_n, nb_pre;
    // TFSF on the H side (k_tfsf's arguments), the next nb_tfsf workgroups:
    // H cells that the sampling and the gather of this launch do not read
    // (the host checks)
    const float* t_signal;
    uint32_t t_len, t_period, t_numTS;
    const uint32_t *t_idx, *t_dint;
    const float *t_dfrac, *t_amp;
    uint32_t t_n, nb_tfsf;
    // the next step's ADE update of the E side (k_ade_update's arguments),
    // the rest: it reads E, which nothing in this launch writes
    const R* a_flux;
    const uint32_t* a_src;
    const uint8_t* a_lor;
    const float *a_int, *a_ext, *a_cl;
    R *a_A, *a_L;
    uint32_t a_n;
};

template <class FV>
__global__ void k_post_h(FV volt, FV curr, PostH<FV> q) {
    uint32_t b = blockIdx.x;
    if (b < q.nb_sample) {  // the first workgroups: the same XCD spread as k_nf2ff_sample
        const uint32_t p = xcd_remap(b, q.nb_sample) * blockDim.x + threadIdx.x;
        nf2ff_sample_point(volt, curr, q.pts, q.el, q.g, q.row_out, q.cap, p);
        return;
    }
    b -= q.nb_sample;
    if (b < q.nb_gather) {
        const uint32_t k = b * blockDim.x + threadIdx.x;
        if (k < q.gn) gather_one(volt, curr, q.gidx, q.gout, k);
        return;
    }
    b -= q.nb_gather;
    if (b < q.nb_pre) {
        const uint32_t k = b * blockDim.x + threadIdx.x;
        if (k < q.mur_n) mur_pre_one(volt, q.vflux, q.coeff, q.normal_src, q.shifted_src, q.saved, k);
        return;
    }
    b -= q.nb_pre;
    if (b < q.nb_tfsf) {
        const uint32_t k = b * blockDim.x + threadIdx.x;
        if (k < q.t_n)
            tfsf_one(curr, q.t_signal, q.t_len, q.t_period, q.t_idx, q.t_dint, q.t_dfrac, q.t_amp, q.t_numTS, k);
        return;
    }
    b -= q.nb_tfsf;
    const uint32_t k = b * blockDim.x + threadIdx.x;
    if (k < q.a_n) ade_update_one(volt, q.a_flux, q.a_src, q.a_lor, q.a_int, q.a_ext, q.a_cl, q.a_A, q.a_L, k);
}

// After the E sweep, when the step's only voltage-side work is TFSF and
// the ADE write-back and their cells are disjoint (the host checks): both
// in one launch, the TFSF entries in the first nb_tfsf workgroups.
template <class FV>
struct PostV {
    typedef typename FV::R R;
    const float* signal;
    uint32_t len, period, numTS;
    const uint32_t *t_idx, *t_dint;
    const float *t_dfrac, *t_amp;
    uint32_t t_n, nb_tfsf;
    const uint32_t* a_gidx;
    const R* a_A;
    uint32_t a_n;
};

template <class FV>
__global__ void k_post_v(FV volt, PostV<FV> q) {
    uint32_t b = blockIdx.x;
    if (b < q.nb_tfsf) {
        const uint32_t k = b * blockDim.x + threadIdx.x;
        if (k < q.t_n) tfsf_one(volt, q.signal, q.len, q.period, q.t_idx, q.t_dint, q.t_dfrac, q.t_amp, q.numTS, k);
        return;
    }
    b -= q.nb_tfsf;
    const uint32_t k = b * blockDim.x + threadIdx.x;
    if (k < q.a_n) volt.add(q.a_gidx[k], -q.a_A[k]);
}

// Point work run by extra workgroups of the sweep kernels (FP32 / FP16
// storage).  Each part reads only arrays that the sweep of its kernel does
// not write, so it runs beside the sweep, and the step loses the launches
// (and the cache write-back and invalidate of each kernel boundary) of
// k_post_h:
//  - kernel of the H sweep of step t (HSIDE = true): the E half of the
//    NF2FF sample of step t for the points that read the fields (E is final
//    after the post-E hooks; the H sweep writes only H), the E values of the
//    probe gather, and the Mur pre pass of step t+1;
//  - kernel of the E sweep of step t+1 (or k_side_flush at the end of the
//    iterate call): the H half of the sample of step t, both halves of
//    captured points (from the slabs of step t's H sweep), and the H values
//    of the gather (H is final after step t's post-H hooks; the E sweep
//    writes only E).
// The sample and gather rows are those of k_post_h, bit for bit.  The
// workgroups come first in the grid and their count is a multiple of 8, so
// the sweep's blocks keep their XCD placement.
struct SideWork {
    FieldView volt, curr;
    // NF2FF sample, workgroups [0, nb_sample)
    const uint32_t* pts;
    const float* el;
    Nf2ffGeom g;
    float* row_out;
    NfCapture cap;
    uint32_t nb_sample;
    // probe gather (E or H values of slot gout), the next nb_gather
    const uint32_t* gidx;
    float* gout;
    uint32_t gn, nb_gather;
    // the next step's Mur pre pass (H side only), the rest
    const float* vflux;
    const float* coeff;
    const uint32_t *normal_src, *shifted_src;
    float* saved;
    uint32_t mur_n;
    uint32_t nb;  // workgroups of side work at the start of the grid (0: none)
    // E-sweep epilogue (ep_on): per block-table entry b, Mur entries
    // ep_mur[ep_moff[b] .. ep_moff[b + 1]) as {shifted cell, normal cell,
    // entry k, coefficient bits} (one load each; a shifted cell with
    // kEpInLds set is an index into the block's LDS copy of its first two
    // planes) and excitation entries
    // ep_exc[ep_eoff[b] .. ep_eoff[b + 1]), with the Mur pre values and
    // the excitation arrays of k_post_e
    uint32_t ep_on;
    const uint4* ep_mur;
    const uint32_t *ep_moff, *ep_eoff, *ep_exc;
    const float* m_saved;
    const float* x_signal;
    uint32_t x_len, x_period, numTS;
    const uint32_t *x_delay, *x_gidx;
    const float* x_amp;
};

// Mur entries one thread of an epilogue holds (the host keeps a block's
// list within kEpiPerThread * kSideThreads, else it uses k_post_e).
constexpr int kEpiPerThread = 16;
constexpr uint32_t kEpInLds = 0x80000000u;

constexpr uint32_t kSideThreads = 256;  // the sweep kernels' workgroup size

template <bool HSIDE>
__device__ __forceinline__ void side_work(const SideWork& q, uint32_t b) {
    if (b < q.nb_sample) {
        const uint32_t p = xcd_remap(b, q.nb_sample) * kSideThreads + threadIdx.x;
        nf2ff_sample_point<FieldView, HSIDE ? kSampleE : kSampleRest>(q.volt, q.curr, q.pts, q.el, q.g, q.row_out,
                                                                       q.cap, p);
        return;
    }
    b -= q.nb_sample;
    if (b < q.nb_gather) {
        const uint32_t k = b * kSideThreads + threadIdx.x;
        if (k < q.gn) {
            const uint32_t j = q.gidx[k];
            if (HSIDE)
                q.gout[2 * k] = q.volt.ld(j);
            else
                q.gout[2 * k + 1] = q.curr.ld(j);
        }
        return;
    }
    b -= q.nb_gather;
    if (HSIDE) {
        const uint32_t k = b * kSideThreads + threadIdx.x;
        if (k < q.mur_n) mur_pre_one(q.volt, q.vflux, q.coeff, q.normal_src, q.shifted_src, q.saved, k);
    }
}

// The end of an E-sweep block when every Mur entry has its normal and its
// shifted cell in one block (the host checks; else k_post_e runs): the
// Mur post of the block's entries from its own new E values, then the
// excitation and the Mur apply, in the reference's order (post of all
// entries, excitation, apply).  Reads and writes stay inside the block's
// cells, which no other workgroup of the launch touches, so two workgroup
// barriers order them: one after the march (its values are written), one
// between the reads and the writes.  b: the block-table entry.
__device__ __forceinline__ void sweep_epilogue(const SideWork& q, uint32_t b, const float* s_ep) {
    if (!q.ep_on) return;
    const uint32_t m0 = q.ep_moff[b], m1 = q.ep_moff[b + 1];
    const uint32_t e0 = q.ep_eoff[b], e1 = q.ep_eoff[b + 1];
    if (m0 == m1 && e0 == e1) return;  // uniform over the workgroup
    // the entries and their pre values do not depend on the march: loaded
    // while the block's other waves finish it
    uint4 e[kEpiPerThread];
    float s[kEpiPerThread];
#pragma unroll
    for (int u = 0; u < kEpiPerThread; ++u) {
        const uint32_t j = m0 + u * kSideThreads + threadIdx.x;
        e[u] = j < m1 ? q.ep_mur[j] : make_uint4(0u, ~0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kEpiPerThread; ++u) s[u] = e[u].y != ~0u ? q.m_saved[e[u].z] : 0.f;
    __syncthreads();  // the block's new E values are written
#pragma unroll
    for (int u = 0; u < kEpiPerThread; ++u)
        if (e[u].y != ~0u) {
            float v;
            if (e[u].x & kEpInLds) {  // the value as stored, times the component's scale (FieldView::ld)
                const uint32_t li = e[u].x & ~kEpInLds;
                v = s_ep[li];
                if (q.volt.half) v = v * q.volt.scale[(li / kEpLdsCells) % 3];
            } else {
                v = q.volt.ld(e[u].x);
            }
            s[u] = fma1(__uint_as_float(e[u].w), v, s[u]);
        }
    __syncthreads();  // every read above precedes every write below
#pragma unroll
    for (int u = 0; u < kEpiPerThread; ++u)
        if (e[u].y != ~0u) q.volt.st(e[u].y, s[u]);
    for (uint32_t j = e0 + threadIdx.x; j < e1; j += kSideThreads)
        excitation_one(q.volt, q.x_signal, q.x_len, q.x_period, q.x_delay, q.x_amp, q.x_gidx, q.ep_exc[j], q.numTS);
}

// The E-sweep side of the last step of an iterate call, on its own.
__global__ __launch_bounds__(kSideThreads) void k_side_flush(SideWork q) { side_work<false>(q, blockIdx.x); }

}  // namespace efs
