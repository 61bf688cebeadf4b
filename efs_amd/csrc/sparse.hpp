// sparse.hpp -- point-list extensions applied around the fused sweeps.
//
// Each kernel is one thread per point.  Reads that the reference performs
// between its UPML pre- and post-passes (ADE update, Mur pre) see the UPML
// flux, not the field, inside PML regions (the pre-pass has swapped them,
// src/webgpu-fdtd.mjs:1094-1172).  Because the sweeps do pre+core+post in
// one pass, those reads go through an alias index resolved on the host:
// bit 31 set = read flux[idx & 0x7fffffff], clear = read field[idx].
// Additions use float atomics so repeated indices accumulate correctly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep.hpp"

namespace efs {

constexpr uint32_t kAliasFlux = 0x80000000u;
constexpr int kPointThreads = 256;

// The kernels that the FP64 storage form also runs (excitation, Mur,
// gather, NF2FF sampling) are templates over the field view FV: FieldView
// (FP32 / FP16 storage, FP32 arithmetic) or FieldView64 (FP64 storage and
// arithmetic); R = FV::R is the arithmetic type, and state arrays the
// kernels keep (Mur's saved values, UPML flux) are R.
template <class FV>
__device__ __forceinline__ typename FV::R read_alias(const FV& field, const typename FV::R* __restrict__ flux,
                                                     uint32_t a) {
    return (a & kAliasFlux) ? flux[a & ~kAliasFlux] : field.ld(a);
}

// apply_excitation (src/webgpu-fdtd.mjs:161-183 semantics: clamp at 0,
// optional period, past-the-end reads signal[0]).
template <class FV>
__device__ __forceinline__ void excitation_one(const FV& volt, const float* __restrict__ signal, uint32_t len,
                                               uint32_t period, const uint32_t* __restrict__ delay,
                                               const float* __restrict__ amp, const uint32_t* __restrict__ gidx,
                                               uint32_t k, uint32_t numTS) {
    typedef typename FV::R R;
    int64_t e = (int64_t)numTS - (int64_t)delay[k];
    if (e < 0) e = 0;
    if (period > 0) e %= (int64_t)period;
    if (e >= (int64_t)len) e = 0;
    volt.add(gidx[k], (R)amp[k] * (R)signal[e]);
}

template <class FV>
__global__ void k_excitation(FV volt, const float* __restrict__ signal, uint32_t len,
                             uint32_t period, const uint32_t* __restrict__ delay,
                             const float* __restrict__ amp, const uint32_t* __restrict__ gidx, uint32_t n,
                             uint32_t numTS) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    excitation_one(volt, signal, len, period, delay, amp, gidx, k, numTS);
}

// Mur ABC (src/webgpu-fdtd.mjs:992-1023): per entry
//   pre    s  = V[sh] - c*V[nor]     (before the E sweep)
//   post   s += c*V[sh]              (after it)
//   apply  V[nor] = s                (after the excitation; the last entry of
//                                     a cell wins, as in the reference's loop)
// The host keeps only each cell's last entry.  `finish` does post + apply in
// one pass; entries whose shifted cell is written between post and apply
// (another entry's normal cell, or an excitation / ADE cell) get their post
// earlier, at its place in the step, from `k_mur_post_some`, and `finish`
// then only applies them (flag bit).  Explicit FMAs fix the rounding.
template <class FV>
__device__ __forceinline__ void mur_pre_one(const FV& volt, const typename FV::R* __restrict__ vflux,
                                            const float* __restrict__ coeff, const uint32_t* __restrict__ normal_src,
                                            const uint32_t* __restrict__ shifted_src,
                                            typename FV::R* __restrict__ saved, uint32_t k) {
    typedef typename FV::R R;
    saved[k] = fma1(-(R)coeff[k], read_alias(volt, vflux, normal_src[k]), read_alias(volt, vflux, shifted_src[k]));
}

template <class FV>
__global__ void k_mur_pre(FV volt, const typename FV::R* __restrict__ vflux,
                          const float* __restrict__ coeff, const uint32_t* __restrict__ normal_src,
                          const uint32_t* __restrict__ shifted_src, typename FV::R* __restrict__ saved, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    mur_pre_one(volt, vflux, coeff, normal_src, shifted_src, saved, k);
}

template <class FV>
__global__ void k_mur_post_some(FV volt, const float* __restrict__ coeff, const uint32_t* __restrict__ shifted,
                                const uint32_t* __restrict__ list, typename FV::R* __restrict__ saved, uint32_t n) {
    typedef typename FV::R R;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t k = list[j];
    saved[k] = fma1((R)coeff[k], volt.ld(shifted[k]), saved[k]);
}

template <class FV>
__global__ void k_mur_finish(FV volt, const float* __restrict__ coeff, const uint32_t* __restrict__ normal,
                             const uint32_t* __restrict__ shifted, const uint8_t* __restrict__ early,
                             const typename FV::R* __restrict__ saved, uint32_t n) {
    typedef typename FV::R R;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const R s = early[k] ? saved[k] : fma1((R)coeff[k], volt.ld(shifted[k]), saved[k]);
    volt.st(normal[k], s);
}

// Lorentz / Drude ADE state update for all (order, direction) entries at
// once (src/webgpu-fdtd.mjs:405-454).  lorentz[k] selects the Lorentz form.
// The state A, Lr and the flux are in the arithmetic type R (double in the
// FP64 storage form); the coefficients stay the FP32 values given.
template <class FV>
__device__ __forceinline__ void ade_update_one(const FV& field, const typename FV::R* __restrict__ flux,
                                               const uint32_t* __restrict__ src, const uint8_t* __restrict__ lorentz,
                                               const float* __restrict__ c_int, const float* __restrict__ c_ext,
                                               const float* __restrict__ c_lor, typename FV::R* __restrict__ A,
                                               typename FV::R* __restrict__ Lr, uint32_t k) {
    typedef typename FV::R R;
    const R V = read_alias(field, flux, src[k]);
    if (lorentz[k]) {
        const R l = Lr[k] + (R)c_lor[k] * A[k];
        Lr[k] = l;
        A[k] = (R)c_int[k] * A[k] + (R)c_ext[k] * (V - l);
    } else {
        A[k] = (R)c_int[k] * A[k] + (R)c_ext[k] * V;
    }
}

template <class FV>
__global__ void k_ade_update(FV field, const typename FV::R* __restrict__ flux,
                             const uint32_t* __restrict__ src, const uint8_t* __restrict__ lorentz,
                             const float* __restrict__ c_int, const float* __restrict__ c_ext,
                             const float* __restrict__ c_lor, typename FV::R* __restrict__ A,
                             typename FV::R* __restrict__ Lr, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    ade_update_one(field, flux, src, lorentz, c_int, c_ext, c_lor, A, Lr, k);
}

// openEMS write-back of the ADE state (optional): field -= A.
template <class FV>
__global__ void k_ade_apply(FV field, const uint32_t* __restrict__ gidx, const typename FV::R* __restrict__ A,
                            uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    field.add(gidx[k], -A[k]);
}

// TFSF injection (src/webgpu-fdtd.mjs:683-739): linear interpolation
// between signal[d] and signal[d+1]; skipped before the delay and at or
// past the last sample.
template <class FV>
__device__ __forceinline__ void tfsf_one(const FV& field, const float* __restrict__ signal, uint32_t len,
                                         uint32_t period, const uint32_t* __restrict__ gidx,
                                         const uint32_t* __restrict__ dint, const float* __restrict__ dfrac,
                                         const float* __restrict__ amp, uint32_t numTS, uint32_t k) {
    typedef typename FV::R R;
    int64_t d = (int64_t)numTS - (int64_t)dint[k];
    if (d < 0) return;
    if (period > 0) d %= (int64_t)period;
    if (d >= (int64_t)len - 1) return;
    const R delta = dfrac[k];
    const R sig = ((R)1 - delta) * (R)signal[d] + delta * (R)signal[d + 1];
    field.add(gidx[k], (R)amp[k] * sig);
}

template <class FV>
__global__ void k_tfsf(FV field, const float* __restrict__ signal, uint32_t len,
                       uint32_t period, const uint32_t* __restrict__ gidx, const uint32_t* __restrict__ dint,
                       const float* __restrict__ dfrac, const float* __restrict__ amp, uint32_t n,
                       uint32_t numTS) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    tfsf_one(field, signal, len, period, gidx, dint, dfrac, amp, numTS, k);
}

// Lumped RLC, CPU-engine order: history shift + parallel inductor current
// before the E sweep, series update + write-back after excitation
// (src/webgpu-fdtd.mjs:775-846).  c: 9 coefficients per element.
__global__ void k_rlc_pre(const uint32_t* __restrict__ type_flag, const float* __restrict__ c,
                          float* __restrict__ vdn, float* __restrict__ vil, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float* V = vdn + 3 * k;
    V[2] = V[1];
    V[1] = V[0];
    if (type_flag[k] == 0) vil[k] += c[9 * k + 0] * c[9 * k + 1] * V[1];
}

__global__ void k_rlc_apply(FieldView volt, const uint32_t* __restrict__ gidx,
                            const uint32_t* __restrict__ type_flag, const float* __restrict__ c,
                            float* __restrict__ vdn, float* __restrict__ jn, const float* __restrict__ vil,
                            uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float* V = vdn + 3 * k;
    float* J = jn + 3 * k;
    const float* q = c + 9 * k;  // i2v ilv vvd vv2 vj1 vj2 ib0 b1 b2
    V[0] = volt.ld(gidx[k]);
    J[2] = J[1];
    J[1] = J[0];
    if (type_flag[k] == 1) {
        const float v0 = q[2] * (V[0] - vil[k] + q[3] * V[2] + q[4] * J[1] + q[5] * J[2]);
        V[0] = v0;
        J[0] = q[6] * (v0 - V[2]) - q[7] * q[6] * J[1] - q[8] * q[6] * J[2];
        volt.st(gidx[k], v0);
    }
}

// Periodic boundary copy for one axis (PBC copy WGSL,
// src/webgpu-engine.mjs:475-584): thread (i0, i1) of the b0 x b1 face copies
// both tangential components from plane src to plane dst of the axis; with
// a Bloch phase the pair (field, imaginary part) is rotated: dst = c*re -
// s*im, im = s*re + c*im, the imaginary part kept per face cell.
struct PbcAxis {
    uint32_t axis, t0, t1;   // periodic axis and its tangential axes
    uint32_t f1;             // face extent along t1 (row length of the imaginary buffers)
    uint32_t b0, b1;         // face cells handled: f0 x f1 (E) or (f0-1) x (f1-1) (H)
    uint32_t src, dst;       // source / destination plane index along the axis
    uint32_t imag_off;       // this face's first imaginary buffer (component t0; t1 at +plane)
    uint32_t plane;          // f0 * f1
    float cs, sn;
    uint32_t has_imag;
};

__global__ void k_pbc(FieldView field, float* __restrict__ imag, PbcAxis a, uint32_t ny, uint32_t nz, uint64_t S) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.b0 * a.b1) return;
    const uint32_t i0 = t / a.b1, i1 = t - i0 * a.b1;
    uint32_t ps[3], pd[3];
    ps[a.axis] = a.src;
    pd[a.axis] = a.dst;
    ps[a.t0] = pd[a.t0] = i0;
    ps[a.t1] = pd[a.t1] = i1;
    const size_t ls = ((size_t)ps[0] * ny + ps[1]) * nz + ps[2], ld = ((size_t)pd[0] * ny + pd[1]) * nz + pd[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint32_t comp = q ? a.t1 : a.t0;
        const float re = field.ld(comp * S + ls);
        if (!a.has_imag) {
            field.st(comp * S + ld, a.cs * re);
        } else {
            float* ip = imag + a.imag_off + (size_t)q * a.plane + (size_t)i0 * a.f1 + i1;
            const float im = *ip;
            field.st(comp * S + ld, a.cs * re - a.sn * im);
            *ip = a.sn * re + a.cs * im;
        }
    }
}

// Steady-state energy accumulation (src/webgpu-engine.mjs:586-625): v^2 of
// each probe added to period 1 while current < period, else to period 2.
__global__ void k_ss_accumulate(FieldView volt, const uint32_t* __restrict__ gidx, uint32_t n,
                                uint32_t period, uint32_t current, float* __restrict__ e1,
                                float* __restrict__ e2) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float v = volt.ld(gidx[k]);
    if (current < period) e1[k] += v * v;
    else e2[k] += v * v;
}

// Padding cells of UPML regions widened to 4-cell lane boundaries: their
// flux is the field's value (the padding's identity coefficients then give
// the core update), set after the regions are configured and after uploads.
__global__ void k_pml_pad_sync(FieldView field, float* __restrict__ flux, const uint32_t* __restrict__ gidx,
                               const uint32_t* __restrict__ fidx, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    flux[fidx[k]] = field.ld(gidx[k]);
}

// Probe gather into out[2*k], out[2*k+1] (src/webgpu-engine.mjs:2593-2665).
template <class FV>
__device__ __forceinline__ void gather_one(const FV& volt, const FV& curr, const uint32_t* __restrict__ idx,
                                           float* __restrict__ out, uint32_t k) {
    const uint32_t j = idx[k];
    out[2 * k] = (float)volt.ld(j);
    out[2 * k + 1] = (float)curr.ld(j);
}

template <class FV>
__global__ void k_gather(FV volt, FV curr, const uint32_t* __restrict__ idx, float* __restrict__ out, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    gather_one(volt, curr, idx, out, k);
}

}  // namespace efs
