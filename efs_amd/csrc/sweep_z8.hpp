// sweep_z8.hpp -- interior E/H marches with 8 z cells per lane.
//
// The same update as e4_march / h4_march in kMarchInterior mode (grids with
// no UPML: the batched form), for lanes that own 8 consecutive z cells.  An
// FP16 lane then loads and stores each field stream 16 bytes at a time (one
// global_load_dwordx4) instead of 8, so a wave keeps twice the bytes in
// flight per memory instruction.  Rows must be padded to a multiple of 8.
// Every cell takes the same fused multiply-adds in the same order as the
// 4-cell marches, so the bits are the same; the z-face Mur entries of the
// E sweep (ZMur) are applied exactly as there.
#pragma once
#include "sweep.hpp"

namespace efs {

typedef float v8 __attribute__((ext_vector_type(8)));
typedef f16 h8 __attribute__((ext_vector_type(8)));

// Field access in stored units, 8 cells (16-B FP16 / 32-B FP32 accesses).
template <typename T>
struct Fs8;
template <>
struct Fs8<float> {
    __device__ static __forceinline__ v8 ldr(const float* p) { return *reinterpret_cast<const v8*>(p); }
    __device__ static __forceinline__ void str(float* p, v8 v) { *reinterpret_cast<v8*>(p) = v; }
    __device__ static __forceinline__ float storedr1(float v) { return v; }
};
template <>
struct Fs8<f16> {
    __device__ static __forceinline__ v8 ldr(const f16* p) {
        return __builtin_convertvector(*reinterpret_cast<const h8*>(p), v8);
    }
    __device__ static __forceinline__ void str(f16* p, v8 v) {
        const v8 c = __builtin_elementwise_min(__builtin_elementwise_max(v, (v8)(-kHalfMax)), (v8)kHalfMax);
        *reinterpret_cast<h8*>(p) = __builtin_convertvector(c, h8);
    }
    __device__ static __forceinline__ float storedr1(float v) { return (float)(f16)clamp16(v); }
};

__device__ __forceinline__ v8 fma8(v8 a, v8 b, v8 c) { return __builtin_elementwise_fma(a, b, c); }

// {s, v0, ..., v6}: the z-1 neighbours of the lane's cells
__device__ __forceinline__ v8 shift_dn8(v8 v, float s) {
    v8 r = __builtin_shufflevector(v, v, 0, 0, 1, 2, 3, 4, 5, 6);
    r[0] = s;
    return r;
}
// {v1, ..., v7, s}: the z+1 neighbours
__device__ __forceinline__ v8 shift_up8(v8 v, float s) {
    v8 r = __builtin_shufflevector(v, v, 1, 2, 3, 4, 5, 6, 7, 7);
    r[7] = s;
    return r;
}

__device__ __forceinline__ v8 sel8(uint32_t bits, v8 t, v8 f) {
    v8 r;
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = ((bits >> k) & 1u) ? t[k] : f[k];
    return r;
}

// Coefficients of 8 cells: 16-bit indices (16 B per component) into the
// cached [3][tabn] table, or the two FP32 arrays.
template <int COMP>
__device__ __forceinline__ void coef8_fetch(const Coef& cf, size_t S, uint32_t i, uint32_t q[12], v8 a[3],
                                            v8 b[3]) {
    if (COMP == kCoefIdx16) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            typedef uint32_t u4 __attribute__((ext_vector_type(4)));
            const u4 w = __builtin_nontemporal_load(reinterpret_cast<const u4*>(cf.idx16 + n * S + i));
            q[4 * n] = w.x;
            q[4 * n + 1] = w.y;
            q[4 * n + 2] = w.z;
            q[4 * n + 3] = w.w;
        }
    } else {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            a[n] = __builtin_nontemporal_load(reinterpret_cast<const v8*>(cf.a + n * S + i));
            b[n] = __builtin_nontemporal_load(reinterpret_cast<const v8*>(cf.b + n * S + i));
        }
    }
}

template <int COMP>
__device__ __forceinline__ void coef8_decode(const uint32_t q[12], v8 a[3], v8 b[3], const Coef& cf) {
    if (COMP == kCoefIdx16) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const float2* t = cf.tab + (size_t)n * cf.tabn;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t w = q[4 * n + k / 2];
                const float2 p = t[(k & 1) ? (w >> 16) : (w & 0xffffu)];
                a[n][k] = p.x;
                b[n][k] = p.y;
            }
        }
    }
}

// E update of 8 cells (lane column (y, z0), z0 a multiple of 8) over x in
// [x0, x1); see e4_march.
template <int COMP, typename T>
__device__ __forceinline__ void e8_march(T* __restrict__ volt, const T* __restrict__ curr, const Coef& cf,
                                         const SweepGeom& g, const Lane4 w, const Scales& sc, const ZMur& zm) {
    typedef Fs<T> L;
    typedef Fs8<T> L8;
    const size_t S = g.S;
    T* const E[3] = {volt, volt + S, volt + 2 * S};
    const T* const H0 = curr;
    const T* const H1 = curr + S;
    const T* const H2 = curr + 2 * S;
    const uint32_t plane = g.cols;
    const uint32_t sy = w.y > 0 ? g.nz : 0u;
    const uint32_t sz = w.z0 > 0 ? 1u : 0u;
    uint32_t i = w.x0 * plane + w.y * g.nz + w.z0;
    bool carry = false;
    v8 hy_m = 0.f, hz_m = 0.f;
    bool zany = false;
    uint32_t zo = 0, zs = 1, zxl[2] = {0x80000000u, 0x80000000u}, zxs[2] = {0u, 0u};
    float zc[2] = {0.f, 0.f};
    if (zm.on) {
        const bool f0 = w.z0 == 0;
        const bool f1 = !f0 && zm.nzr >= 6 && zm.nzr - 1 >= w.z0 && zm.nzr - 1 < w.z0 + 8 && (zm.nzr - 1) % 8 != 0;
        if (f1) {
            zo = zm.nzr - 1 - w.z0;
            zs = zo - 1;
        }
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            const int32_t x0 = f1 ? zm.x0[2 + n] : zm.x0[n], x1 = f1 ? zm.x1[2 + n] : zm.x1[n];
            const int32_t y0 = f1 ? zm.y0[2 + n] : zm.y0[n], y1 = f1 ? zm.y1[2 + n] : zm.y1[n];
            if ((f0 || f1) && (int32_t)w.y >= y0 && (int32_t)w.y <= y1 && x0 <= x1) {
                zxl[n] = (uint32_t)x0;
                zxs[n] = (uint32_t)(x1 - x0);
                zc[n] = f1 ? zm.c[2 + n] : zm.c[n];
                zany = true;
            }
        }
    }
    for (uint32_t x = w.x0; x < w.x1; ++x, i += plane) {
        if (!carry) {
            const uint32_t ib = x > 0 ? i - plane : i;
            hy_m = L8::ldr(H1 + ib);
            hz_m = L8::ldr(H2 + ib);
        }
        const v8 hx = L8::ldr(H0 + i), hy = L8::ldr(H1 + i), hz = L8::ldr(H2 + i);
        const v8 hz_ym = L8::ldr(H2 + i - sy), hx_ym = L8::ldr(H0 + i - sy);
        const float hy_z = L::ldr1(H1 + i - sz), hx_z = L::ldr1(H0 + i - sz);
        v8 V[3], a[3], b[3];
        uint32_t q[12];
#pragma unroll
        for (int n = 0; n < 3; ++n) V[n] = L8::ldr(E[n] + i);
        coef8_fetch<COMP>(cf, S, i, q, a, b);
        coef8_decode<COMP>(q, a, b, cf);
        const v8 hy_zm = shift_dn8(hy, hy_z);
        const v8 hx_zm = shift_dn8(hx, hx_z);
        const float s0 = L::sc(sc.other[0]), s1 = L::sc(sc.other[1]), s2 = L::sc(sc.other[2]);
        const v8 cur[3] = {fma8((v8)s1, hy_zm, fma8((v8)(-s1), hy, s2 * (hz - hz_ym))),
                           fma8((v8)s2, hz_m, fma8((v8)(-s2), hz, s0 * (hx - hx_zm))),
                           fma8((v8)s0, hx_ym, fma8((v8)(-s0), hx, s1 * (hy - hy_m)))};
        hy_m = hy;
        hz_m = hz;
        carry = true;
        v8 vn[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) vn[n] = fma8(a[n], V[n], (b[n] * cur[n]) * L::sc(sc.own_inv[n]));
        if (zany) {
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                if (x - zxl[n] > zxs[n]) continue;
                const float pre = __builtin_fmaf(-zc[n], V[n][zo], V[n][zs]);
                const float st = L8::storedr1(vn[n][zs]);
                vn[n][zo] = __builtin_fmaf(zc[n], st, pre);
            }
        }
#pragma unroll
        for (int n = 0; n < 3; ++n) L8::str(E[n] + i, vn[n]);
    }
}

// H update of 8 cells over x in [x0, x1); see h4_march.  Cells outside the
// core bounds (x = Nx-1, y = Ny-1, the row's last cell) keep I.
template <int COMP, typename T>
__device__ __forceinline__ void h8_march(T* __restrict__ curr, const T* __restrict__ volt, const Coef& cf,
                                         const SweepGeom& g, const Lane4 w, const Scales& sc) {
    typedef Fs<T> L;
    typedef Fs8<T> L8;
    const size_t S = g.S;
    T* const Ic[3] = {curr, curr + S, curr + 2 * S};
    const T* const E0 = volt;
    const T* const E1 = volt + S;
    const T* const E2 = volt + 2 * S;
    const uint32_t plane = g.cols;
    const bool row_core = w.y + 1 < g.ny;
    const bool last = w.z0 + 8 >= g.nz;
    if (!row_core) return;
    const uint32_t sy = g.nz;
    const uint32_t sz = last ? 7u : 8u;
    uint32_t i = w.x0 * plane + w.y * g.nz + w.z0;
    bool carry = false;
    v8 ex = 0.f, ey = 0.f, ez = 0.f;
    for (uint32_t x = w.x0; x < w.x1; ++x, i += plane) {
        const bool core = x + 1 < g.nx;
        if (!core) {
            carry = false;
            continue;
        }
        if (!carry) {
            ex = L8::ldr(E0 + i);
            ey = L8::ldr(E1 + i);
            ez = L8::ldr(E2 + i);
        }
        const uint32_t cb = last ? 0x7fu : 0xffu;
        const v8 ex_p = L8::ldr(E0 + i + plane), ey_p = L8::ldr(E1 + i + plane), ez_p = L8::ldr(E2 + i + plane);
        const v8 ez_yp = L8::ldr(E2 + i + sy), ex_yp = L8::ldr(E0 + i + sy);
        const float ey_z = L::ldr1(E1 + i + sz), ex_z = L::ldr1(E0 + i + sz);
        v8 I[3], a[3], b[3];
        uint32_t q[12];
#pragma unroll
        for (int n = 0; n < 3; ++n) I[n] = L8::ldr(Ic[n] + i);
        coef8_fetch<COMP>(cf, S, i, q, a, b);
        coef8_decode<COMP>(q, a, b, cf);
        const v8 ey_zp = shift_up8(ey, ey_z);
        const v8 ex_zp = shift_up8(ex, ex_z);
        const float s0 = L::sc(sc.other[0]), s1 = L::sc(sc.other[1]), s2 = L::sc(sc.other[2]);
        const v8 cur[3] = {fma8((v8)s1, ey_zp, fma8((v8)(-s1), ey, s2 * (ez - ez_yp))),
                           fma8((v8)s2, ez_p, fma8((v8)(-s2), ez, s0 * (ex - ex_zp))),
                           fma8((v8)s0, ex_yp, fma8((v8)(-s0), ex, s1 * (ey - ey_p)))};
        ex = ex_p;
        ey = ey_p;
        ez = ez_p;
        carry = true;
#pragma unroll
        for (int n = 0; n < 3; ++n)
            L8::str(Ic[n] + i, sel8(cb, fma8(a[n], I[n], (b[n] * cur[n]) * L::sc(sc.own_inv[n])), I[n]));
    }
}

}  // namespace efs
