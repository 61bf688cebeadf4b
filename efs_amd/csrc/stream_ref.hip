// Streaming-rate references for reading the sweeps' real traffic against,
// timed with HIP events over arrays larger than the last-level cache.  A
// measurement aid of the benchmark (no counterpart in the reference); it is
// built into its own library, libefs_stream_ref.so, so the engine library
// carries none of it.
//
//   int efs_stream_ref(int device, uint64_t n, uint32_t k, uint32_t grid,
//                      uint32_t reps, double* ms)
//     out_c = a_c + 0.1 * b_c, c < k, over 3k FP32 arrays of n floats (2k
//     reads + k writes per element).  n: floats per array (a multiple of 4);
//     k: 1..3 output streams; grid: workgroups of 256 threads (grid-stride
//     loop); ms: average time of one call.
//
//   int efs_stream_kind(int device, int kind, uint64_t n, uint32_t grid,
//                       uint32_t reps, double* ms)
//     kind 0: copy o = a with 4 independent 16-byte loads in flight per
//             thread (grid-stride), plain loads and stores;
//     kind 1: the same copy with non-temporal loads and stores;
//     kind 2: read only: a sum over a (one float per thread written, so the
//             loads are not removed);
//     kind 3: hipMemcpyAsync device to device (the runtime's own copy).
//     Bytes per call: 8n for kinds 0, 1, 3; 4n for kind 2.
//
// Both return 0, or -1 with a message on stderr.
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

constexpr int kU = 4;  // independent 16-byte loads in flight per thread

template <bool NT>
__global__ __launch_bounds__(256) void k_copy(const v4* __restrict__ a, v4* __restrict__ o, uint64_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; i + (kU - 1) * stride < n4; i += kU * stride) {
        v4 r[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) r[u] = NT ? __builtin_nontemporal_load(a + i + u * stride) : a[i + u * stride];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (NT)
                __builtin_nontemporal_store(r[u], o + i + u * stride);
            else
                o[i + u * stride] = r[u];
        }
    }
    for (; i < n4; i += stride) o[i] = a[i];
}

__global__ __launch_bounds__(256) void k_read(const v4* __restrict__ a, uint64_t n4, float* __restrict__ sink) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    v4 s = 0.f;
    for (; i + (kU - 1) * stride < n4; i += kU * stride) {
        v4 r[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) r[u] = a[i + u * stride];
#pragma unroll
        for (int u = 0; u < kU; ++u) s += r[u];
    }
    for (; i < n4; i += stride) s += a[i];
    sink[blockIdx.x * (uint64_t)blockDim.x + threadIdx.x] = s.x + s.y + s.z + s.w;
}

struct Timer {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = false;
    explicit Timer(int device) {
        ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    }
    ~Timer() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
    // average ms of `reps` calls of f after 3 warm calls
    template <class F>
    bool time(F f, uint32_t reps, double* ms) {
        for (int w = 0; w < 3; ++w) f();
        float t = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess) return false;
        for (uint32_t r = 0; r < reps; ++r) f();
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess)
            return false;
        *ms = (double)t / reps;
        return true;
    }
};

}  // namespace

extern "C" int efs_stream_ref(int device, uint64_t n, uint32_t k, uint32_t grid_blocks, uint32_t reps,
                              double* ms) {
    if (!ms || n == 0 || n % 4 != 0 || reps == 0 || k < 1 || k > 3 || grid_blocks == 0) {
        fprintf(stderr, "efs_stream_ref: bad arguments\n");
        return -1;
    }
    Timer T(device);
    float* buf = nullptr;
    const uint64_t n4 = n / 4;
    const dim3 grid(grid_blocks), block(256);
    const size_t bytes = 3 * (size_t)k * n * sizeof(float);
    int rc = -1;
    if (!T.ok || hipMalloc((void**)&buf, bytes) != hipSuccess || hipMemsetAsync(buf, 0, bytes, T.s) != hipSuccess) {
        fprintf(stderr, "efs_stream_ref: HIP set-up failed\n");
    } else {
        auto launch = [&] {
            if (k == 1)
                hipLaunchKernelGGL(k_stream_add<1>, grid, block, 0, T.s, (v4*)buf, n4);
            else if (k == 2)
                hipLaunchKernelGGL(k_stream_add<2>, grid, block, 0, T.s, (v4*)buf, n4);
            else
                hipLaunchKernelGGL(k_stream_add<3>, grid, block, 0, T.s, (v4*)buf, n4);
        };
        if (T.time(launch, reps, ms))
            rc = 0;
        else
            fprintf(stderr, "efs_stream_ref: launch or timing failed\n");
    }
    if (buf) (void)hipFree(buf);
    return rc;
}

extern "C" int efs_stream_kind(int device, int kind, uint64_t n, uint32_t grid_blocks, uint32_t reps, double* ms) {
    if (!ms || n == 0 || n % 4 != 0 || reps == 0 || kind < 0 || kind > 3 || grid_blocks == 0) {
        fprintf(stderr, "efs_stream_kind: bad arguments\n");
        return -1;
    }
    Timer T(device);
    float *a = nullptr, *o = nullptr, *sink = nullptr;
    const uint64_t n4 = n / 4;
    const dim3 grid(grid_blocks), block(256);
    int rc = -1;
    if (!T.ok || hipMalloc((void**)&a, n * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&o, n * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&sink, (size_t)grid_blocks * 256 * sizeof(float)) != hipSuccess ||
        hipMemsetAsync(a, 0, n * sizeof(float), T.s) != hipSuccess) {
        fprintf(stderr, "efs_stream_kind: HIP set-up failed\n");
    } else {
        auto launch = [&] {
            if (kind == 0)
                hipLaunchKernelGGL(k_copy<false>, grid, block, 0, T.s, (const v4*)a, (v4*)o, n4);
            else if (kind == 1)
                hipLaunchKernelGGL(k_copy<true>, grid, block, 0, T.s, (const v4*)a, (v4*)o, n4);
            else if (kind == 2)
                hipLaunchKernelGGL(k_read, grid, block, 0, T.s, (const v4*)a, n4, sink);
            else
                (void)hipMemcpyAsync(o, a, n * sizeof(float), hipMemcpyDeviceToDevice, T.s);
        };
        if (T.time(launch, reps, ms))
            rc = 0;
        else
            fprintf(stderr, "efs_stream_kind: launch or timing failed\n");
    }
    if (a) (void)hipFree(a);
    if (o) (void)hipFree(o);
    if (sink) (void)hipFree(sink);
    return rc;
}
