// Streaming-rate reference for reading the sweeps' real traffic against:
// out_c = a_c + 0.1 * b_c, c < k, over 3k FP32 arrays of n floats (2k
// reads + k writes per element), larger than the last-level cache, timed
// with HIP events.  A measurement aid of the benchmark (no counterpart in the
// reference); it is built into its own library, libefs_stream_ref.so, so the
// engine library carries none of it.
//
//   int efs_stream_ref(int device, uint64_t n, uint32_t k, uint32_t grid,
//                      uint32_t reps, double* ms)
//     n: floats per array (a multiple of 4); k: 1..3 output streams; grid:
//     workgroups of 256 threads (grid-stride loop); ms: average time of one
//     call.  Returns 0, or -1 with a message on stderr.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

typedef float v4 __attribute__((ext_vector_type(4)));

// buf: a_0..a_{k-1}, b_0..b_{k-1}, o_0..o_{k-1}, n4 vectors each
template <int K>
__global__ __launch_bounds__(256) void k_stream_add(v4* __restrict__ buf, uint64_t n4) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x) {
        v4 r[K];
#pragma unroll
        for (int c = 0; c < K; ++c) r[c] = buf[c * n4 + i] + 0.1f * buf[(K + c) * n4 + i];
#pragma unroll
        for (int c = 0; c < K; ++c) buf[(2 * K + c) * n4 + i] = r[c];
    }
}

}  // namespace

extern "C" int efs_stream_ref(int device, uint64_t n, uint32_t k, uint32_t grid_blocks, uint32_t reps,
                              double* ms) {
    if (!ms || n == 0 || n % 4 != 0 || reps == 0 || k < 1 || k > 3 || grid_blocks == 0) {
        fprintf(stderr, "efs_stream_ref: bad arguments\n");
        return -1;
    }
    float* buf = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -1;
    float t = 0.f;
    const uint64_t n4 = n / 4;
    const dim3 grid(grid_blocks), block(256);
    const size_t bytes = 3 * (size_t)k * n * sizeof(float);
    auto launch = [&] {
        if (k == 1)
            hipLaunchKernelGGL(k_stream_add<1>, grid, block, 0, s, (v4*)buf, n4);
        else if (k == 2)
            hipLaunchKernelGGL(k_stream_add<2>, grid, block, 0, s, (v4*)buf, n4);
        else
            hipLaunchKernelGGL(k_stream_add<3>, grid, block, 0, s, (v4*)buf, n4);
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMalloc((void**)&buf, bytes) != hipSuccess) {
        fprintf(stderr, "efs_stream_ref: HIP set-up failed\n");
        goto done;
    }
    if (hipMemsetAsync(buf, 0, bytes, s) != hipSuccess) goto done;
    for (int w = 0; w < 3; ++w) launch();
    if (hipEventRecord(e0, s) != hipSuccess) goto done;
    for (uint32_t r = 0; r < reps; ++r) launch();
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&t, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess) {
        fprintf(stderr, "efs_stream_ref: launch or timing failed\n");
        goto done;
    }
    *ms = (double)t / reps;
    rc = 0;
done:
    if (buf) hipFree(buf);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    return rc;
}
