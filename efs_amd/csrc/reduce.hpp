// reduce.hpp -- field energy (computeEnergy, src/webgpu-engine.mjs:2974-3158).
//
// Sum of volt^2 + curr^2 over the (Nx-1)(Ny-1)(Nz-1) interior, 3 components.
// One wavefront walks one (x, y) row at a time with lanes along z
// (coalesced); squares are summed in FP32 per row, rows in FP64; the wave
// total is formed with cross-lane shuffles in FP64 and a second single-
// workgroup kernel adds the per-workgroup partials in a fixed order, so the
// result is deterministic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep.hpp"

namespace efs {

constexpr int kEnergyThreads = 256;
constexpr int kEnergyMaxBlocks = 2048;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// T: field storage; with FP16 each component's squares are scaled by its
// (power-of-two) scale squared, so the sum is of the same values.  Squares
// and row sums are FP32, or FP64 for FP64 storage.
template <typename T>
struct EnergyArith {
    typedef float type;
    static constexpr bool kHalf = Fs<T>::kHalf;
};
template <>
struct EnergyArith<double> {
    typedef double type;
    static constexpr bool kHalf = false;
};

template <typename T>
__global__ __launch_bounds__(kEnergyThreads) void k_energy_partial(const T* __restrict__ volt,
                                                                   const T* __restrict__ curr,
                                                                   uint32_t nx, uint32_t ny, uint32_t nz,
                                                                   uint32_t nzs, uint64_t S,
                                                                   const float* __restrict__ fscale,
                                                                   double* __restrict__ partial) {
    __shared__ double s_w[kEnergyThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t nwaves = gridDim.x * (kEnergyThreads / 64);
    const size_t N = S;  // component stride
    const uint32_t rows = (nx - 1) * (ny - 1);
    double acc = 0.0;
    for (uint32_t row = blockIdx.x * (kEnergyThreads / 64) + wv; row < rows; row += nwaves) {
        const uint32_t x = row / (ny - 1), y = row - x * (ny - 1);
        const size_t base = ((size_t)x * ny + y) * nzs;  // nzs: device row length
        typedef typename EnergyArith<T>::type E;
        E s = 0;
        for (uint32_t z = lane; z + 1 < nz; z += 64) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                E v = (E)volt[n * N + base + z], c = (E)curr[n * N + base + z];
                if (EnergyArith<T>::kHalf) {
                    v *= fscale[n];
                    c *= fscale[3 + n];
                }
                s += v * v + c * c;
            }
        }
        acc += (double)s;
    }
    acc = wave_sum(acc);
    if (lane == 0) s_w[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kEnergyThreads / 64; ++k) t += s_w[k];
        partial[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kEnergyThreads) void k_energy_final(const double* __restrict__ partial, uint32_t n,
                                                                 double* __restrict__ out) {
    __shared__ double s_w[kEnergyThreads / 64];
    double acc = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += kEnergyThreads) acc += partial[k];
    acc = wave_sum(acc);
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    if (lane == 0) s_w[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kEnergyThreads / 64; ++k) t += s_w[k];
        out[0] = t;
    }
}

}  // namespace efs

namespace efs {

// Range control of FP16 field storage.  Every few steps k_absmax16 takes the
// largest |stored| per component (6: volt 0..2, curr 0..2), k_range_plan
// picks a power-of-two shift that brings it back near 2^kRangeTarget when it
// left [2^kRangeLo, 2^kRangeHi), updates the component's scale by the
// inverse factor, and k_range_apply multiplies the stored values (exact).
// A scale never gets finer than a floor, so that a component still near zero
// is not scaled for values far below the field's working range (the
// arriving wave would overflow before the next check): while a source still
// injects, the floor is smin (the scale the source or the uploaded fields
// set); afterwards the field can only redistribute and decay, and the floor
// is 2^-kRangeSibling of the largest component of the same array, which
// follows the decay.
// Everything stays on the device: no host synchronisation in iterate().
constexpr int kRangeLo = 6, kRangeHi = 12, kRangeTarget = 10, kRangeSibling = 14;

// Block-level maximum first (waves, then LDS), one atomic per workgroup:
// per-wave atomics on the 6 words serialise at the L2.
__device__ __forceinline__ float absmax_block(float m) {
    __shared__ float s_m[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_down(m, off, 64);
        m = (o != o || m != m) ? __builtin_nanf("") : fmaxf(m, o);
    }
    if ((threadIdx.x & 63u) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            const float o = s_m[w];
            m = (o != o || m != m) ? __builtin_nanf("") : fmaxf(m, o);
        }
    }
    return m;  // valid in thread 0
}

__global__ __launch_bounds__(256) void k_absmax16(const f16* __restrict__ volt, const f16* __restrict__ curr,
                                                  uint64_t S, uint32_t N, unsigned int* __restrict__ amax) {
    const uint32_t k = blockIdx.y;  // component 0..5
    const f16* f = (k < 3 ? volt : curr) + (uint64_t)(k % 3) * S;
    float m = 0.f;
    for (uint32_t i = 4 * (blockIdx.x * blockDim.x + threadIdx.x); i < N; i += 4 * gridDim.x * blockDim.x) {
        const v4 v = __builtin_convertvector(*reinterpret_cast<const h4*>(f + i), v4);
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        if (v.x != v.x || v.y != v.y || v.z != v.z || v.w != v.w) m = __builtin_nanf("");
    }
    m = absmax_block(m);
    // non-negative floats order like their bit patterns; NaN sorts above inf
    if (threadIdx.x == 0) atomicMax(amax + k, __float_as_uint(m));
}

__global__ void k_range_plan(unsigned int* __restrict__ amax, float* __restrict__ fscale,
                             const float* __restrict__ smin, float* __restrict__ factor, uint32_t source_active) {
    if (threadIdx.x != 0) return;
    float m[6], arr[2] = {0.f, 0.f};  // largest value (real units) of each array
    for (int k = 0; k < 6; ++k) {
        m[k] = __uint_as_float(amax[k]);
        amax[k] = 0u;
        if (m[k] < __builtin_huge_valf()) arr[k / 3] = fmaxf(arr[k / 3], m[k] * fscale[k]);
    }
    for (int k = 0; k < 6; ++k) {
        float f = 1.0f;
        if (m[k] > 0.0f && m[k] < __builtin_huge_valf()) {
            const int e = ilogbf(m[k]);
            if (e < kRangeLo || e >= kRangeHi) {
                int shift = kRangeTarget - e;  // > 0: finer scale, stored values grow
                if (shift > 0) {
                    const int floor_e = source_active ? ilogbf(smin[k]) : ilogbf(arr[k / 3]) - kRangeSibling;
                    shift = min(shift, max(0, ilogbf(fscale[k]) - floor_e));
                }
                f = ldexpf(1.0f, shift);
                fscale[k] = ldexpf(fscale[k], -shift);
            }
        }
        factor[k] = f;
    }
}

__global__ __launch_bounds__(256) void k_range_apply(f16* __restrict__ volt, f16* __restrict__ curr, uint64_t S,
                                                     uint32_t N, const float* __restrict__ factor) {
    const uint32_t k = blockIdx.y;
    const float f = factor[k];
    if (f == 1.0f) return;
    f16* p = (k < 3 ? volt : curr) + (uint64_t)(k % 3) * S;
    for (uint32_t i = 4 * (blockIdx.x * blockDim.x + threadIdx.x); i < N; i += 4 * gridDim.x * blockDim.x) {
        h4* q = reinterpret_cast<h4*>(p + i);
        *q = __builtin_convertvector(clamp16(__builtin_convertvector(*q, v4) * f), h4);
    }
}

}  // namespace efs
