// grid_barrier.hpp -- grid-wide barrier for persistent kernels on gfx950.
//
// A persistent kernel keeps every workgroup resident (cooperative launch)
// and separates its phases with a barrier over the whole grid.  On gfx950
// each XCD has its own L2: stores of one XCD reach another only after a
// write-back of the writer's L2 (agent-scope release, buffer_wbl2) and an
// invalidate on the reader's side (agent-scope acquire, buffer_inv).  The
// write-back walks the whole L2, so one per workgroup (the textbook
// grid barrier: fence, atomic, spin, fence in every workgroup) serialises
// dozens of walks per XCD.  Here the workgroups of one XCD meet first: each
// drains its own stores to L2 (s_waitcnt vmcnt(0) in every wave, then the
// workgroup barrier) and counts itself in its XCD's counter; the last one
// of the XCD writes the L2 back once and counts the XCD in the global
// counter.  Every workgroup then waits for all XCDs and invalidates its
// caches (agent-scope acquire, buffer_inv sc1).  Measured on one tiny grid
// (a phase of a few microseconds), per barrier: ~65 us with a write-back in
// every workgroup, ~12 us with one per XCD; a kernel boundary costs ~6 us.
// An invalidate of the L2 once per XCD with only the L1 invalidated in the
// other workgroups (buffer_inv sc0) is not enough: it gave stale values.
//
// The XCD of a workgroup is read from the hardware (HW_REG_XCC_ID), and the
// number of workgroups per XCD is counted once at kernel start, so nothing
// depends on how the dispatcher spreads workgroups over the XCDs.
//
// Every wait is bounded: after kSpinLimit polls the waiting workgroup sets
// the error word and every later wait returns at once, so a grid that is
// not co-resident ends with an error instead of hanging.  Counters only
// grow within a launch; the host clears the state (a fill kernel) before
// every launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efs {

constexpr uint32_t kMaxXcc = 8;
constexpr uint32_t kBarStride = 16;  // 64 B between counters
constexpr uint32_t kSpinLimit = 1u << 22;

struct GridBarState {
    uint32_t reg[kMaxXcc * kBarStride];   // workgroups per XCD (counted at start)
    uint32_t xcnt[kMaxXcc * kBarStride];  // arrivals per XCD (monotonic)
    uint32_t reg_all;                     // workgroups registered
    uint32_t pad0[kBarStride - 1];
    uint32_t glob;                        // XCDs released (monotonic)
    uint32_t pad1[kBarStride - 1];
    uint32_t err;                         // nonzero: a wait timed out
    uint32_t pad2[kBarStride - 1];
};
static_assert(sizeof(GridBarState) % 4 == 0, "words");
constexpr uint32_t kGridBarWords = sizeof(GridBarState) / 4;

__device__ __forceinline__ uint32_t bar_ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t bar_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t hw_xcc_id() {
    uint32_t x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x & 7u;
}

// Per-workgroup view of the barrier (kept in registers by thread 0).
struct GridBar {
    GridBarState* st;
    uint32_t xcc;      // this workgroup's XCD
    uint32_t nwg_x;    // workgroups on it
    uint32_t nact;     // XCDs with at least one workgroup
    uint32_t gen;      // barriers passed
    bool flat;         // every workgroup writes its L2 back (reference form)

    // Bounded wait for *p >= want; false (and the error word set) on timeout
    // or when another workgroup timed out.
    __device__ bool wait_ge(const uint32_t* p, uint32_t want) {
        for (uint32_t n = 0; n < kSpinLimit; ++n) {
            if (bar_ld(p) >= want) return true;
            if (bar_ld(&st->err)) return false;
            __builtin_amdgcn_s_sleep(2);
        }
        bar_add(&st->err, 1u);
        return false;
    }

    // Registration (kernel start; all threads of the workgroup call it).
    // Returns false when the grid did not become resident in time.
    __device__ bool init(GridBarState* s, bool flat_form) {
        __shared__ uint32_t sh[3];
        st = s;
        flat = flat_form;
        gen = 0;
        if (threadIdx.x == 0) {
            xcc = hw_xcc_id();
            bar_add(&st->reg[xcc * kBarStride], 1u);
            bar_add(&st->reg_all, 1u);
            const bool ok = wait_ge(&st->reg_all, gridDim.x);
            uint32_t act = 0;
            for (uint32_t k = 0; k < kMaxXcc; ++k) act += bar_ld(&st->reg[k * kBarStride]) ? 1u : 0u;
            sh[0] = ok ? 1u : 0u;
            sh[1] = bar_ld(&st->reg[xcc * kBarStride]);
            sh[2] = act;
        }
        __syncthreads();
        const bool ok = sh[0] != 0;
        nwg_x = sh[1];
        nact = sh[2];
        __syncthreads();
        return ok;
    }

    // Grid-wide barrier with release/acquire of every global store made
    // before it (all threads call it).  False on timeout.
    __device__ bool sync() {
        __shared__ uint32_t ok_sh;
        // every wave's stores acknowledged by L2 before the workgroup counts
        // itself (the workgroup-scope fence does not wait for them)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        ++gen;
        if (threadIdx.x == 0) {
            bool ok;
            if (flat) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                bar_add(&st->glob, 1u);
                ok = wait_ge(&st->glob, gen * gridDim.x);
            } else {
                const uint32_t old = bar_add(&st->xcnt[xcc * kBarStride], 1u);
                if (old + 1u == gen * nwg_x) {
                    // last of its XCD: one write-back covers the XCD's stores
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                    bar_add(&st->glob, 1u);
                }
                ok = wait_ge(&st->glob, gen * nact);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ok_sh = ok ? 1u : 0u;
        }
        __syncthreads();
        return ok_sh != 0;
    }
};

// Clears a barrier state before a launch (no memset nodes in graphs).
__global__ void k_gridbar_clear(uint32_t* __restrict__ w) {
    for (uint32_t i = threadIdx.x; i < kGridBarWords; i += blockDim.x) w[i] = 0u;
}

}  // namespace efs
