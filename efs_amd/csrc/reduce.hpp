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

namespace efs {

constexpr int kEnergyThreads = 256;
constexpr int kEnergyMaxBlocks = 2048;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kEnergyThreads) void k_energy_partial(const float* __restrict__ volt,
                                                                   const float* __restrict__ curr,
                                                                   uint32_t nx, uint32_t ny, uint32_t nz,
                                                                   uint64_t S, double* __restrict__ partial) {
    __shared__ double s_w[kEnergyThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t nwaves = gridDim.x * (kEnergyThreads / 64);
    const size_t N = S;  // component stride
    const uint32_t rows = (nx - 1) * (ny - 1);
    double acc = 0.0;
    for (uint32_t row = blockIdx.x * (kEnergyThreads / 64) + wv; row < rows; row += nwaves) {
        const uint32_t x = row / (ny - 1), y = row - x * (ny - 1);
        const size_t base = ((size_t)x * ny + y) * nz;
        float s = 0.f;
        for (uint32_t z = lane; z + 1 < nz; z += 64) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const float v = volt[n * N + base + z], c = curr[n * N + base + z];
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
