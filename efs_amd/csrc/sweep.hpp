// sweep.hpp -- fused Yee E and H sweeps for gfx950 (CDNA4).
//
// Restates update_voltages / update_currents (src/webgpu-engine.mjs:7-122,
// CPU form src/webgpu-fdtd.mjs:64-155) with the UPML pre/post passes
// (update_pml modes 0-3, src/webgpu-engine.mjs:124-204) folded in.
//
// Mapping.  Fields keep the reference's NIJK layout, where for one x plane
// and one component the (y, z) cells form a contiguous run of Ny*Nz floats.
// A workgroup owns 256 consecutive cells of that run ("columns" c = y*Nz+z,
// fully coalesced whatever Nz is) and marches along x, the slowest axis,
// over a chunk of planes.  The x-neighbour needed by the stencil stays in
// registers from the previous plane, so each H value (E sweep) or E value
// (H sweep) is fetched from HBM once; the y and z neighbours belong to
// adjacent lanes/waves of the same workgroup and are served by L1/L2.
// Workgroup ids are remapped so each XCD (own L2) gets a contiguous range
// of tiles: neighbouring column tiles share an L2.
//
// UPML fold (per component, per PML cell; F = flux, exact algebra of
// pre -> core -> post):
//   F' = vv_c*F + vi*curl                      (the core update acts on F)
//   V' = (vv_p*V - vvfo*F) + vvfn*F'
// and for H the same with ii/iv, iifo/iifn.  H cells outside the core
// bounds (last x/y/z plane) get the pre/post pair without the core update:
//   I' = (ii_p*I - iifo*F) + iifn*F,  F unchanged.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efs {

constexpr int kMaxPml = 16;
constexpr int kSweepThreads = 256;

struct PmlRegion {
    uint32_t x0, x1, y0, y1, z0, z1;  // [x0,x1) x [y0,y1) x [z0,z1)
    uint32_t ny, nz;                  // local extents in y, z
    uint32_t cs;                      // nx*ny*nz (component stride)
    uint32_t pad;
    uint64_t coef;                    // element offset of this region's 6 coefficient arrays
    uint64_t flux;                    // element offset of this region's flux arrays (3*cs each)
};

struct PmlTable {
    uint32_t n;
    uint32_t pad;
    PmlRegion r[kMaxPml];
};

struct SweepGeom {
    uint32_t nx, ny, nz;
    uint32_t cols;        // ny*nz
    uint32_t col_tiles;   // ceil(cols / 256)
    uint32_t chunks;      // x chunks
    uint32_t chunk_len;   // planes per chunk
    uint32_t blocks;      // col_tiles * chunks
};

// XCD-contiguous remap of the hardware workgroup id: dispatch hands out
// workgroup ids round-robin over the 8 XCDs; this gives XCD k the logical
// range [k*q + min(k,r), ...) so adjacent tiles land on one XCD.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    const uint32_t xcd = bid & 7u, k = bid >> 3;
    const uint32_t q = nblocks >> 3, r = nblocks & 7u;
    return (xcd < r) ? xcd * (q + 1u) + k : r * (q + 1u) + (xcd - r) * q + k;
}

// Bit mask of the PML regions whose (y, z) extent contains this column.
__device__ __forceinline__ uint32_t pml_column_mask(const PmlTable& t, uint32_t y, uint32_t z) {
    uint32_t m = 0;
    for (uint32_t k = 0; k < t.n; ++k)
        if (y >= t.r[k].y0 && y < t.r[k].y1 && z >= t.r[k].z0 && z < t.r[k].z1) m |= 1u << k;
    return m;
}

__device__ __forceinline__ int pml_find(const PmlTable& t, uint32_t mask, uint32_t x) {
    while (mask) {
        const int k = __builtin_ctz(mask);
        mask &= mask - 1u;
        if (x >= t.r[k].x0 && x < t.r[k].x1) return k;
    }
    return -1;
}

// E sweep: V = vv*V + vi*curl(H), backward neighbour clamped to self at 0.
template <bool HAS_PML>
__global__ __launch_bounds__(kSweepThreads) void sweep_e(float* __restrict__ volt,
                                                         const float* __restrict__ curr,
                                                         const float* __restrict__ vv,
                                                         const float* __restrict__ vi, SweepGeom g,
                                                         PmlTable pml,
                                                         const float* __restrict__ pml_coef,
                                                         float* __restrict__ pml_vflux) {
    const uint32_t L = xcd_remap(blockIdx.x, g.blocks);
    const uint32_t tile = L % g.col_tiles, chunk = L / g.col_tiles;
    const uint32_t c = tile * kSweepThreads + threadIdx.x;
    if (c >= g.cols) return;
    const uint32_t y = c / g.nz, z = c - y * g.nz;
    const uint32_t x0 = chunk * g.chunk_len;
    const uint32_t x1 = min(g.nx, x0 + g.chunk_len);
    if (x0 >= x1) return;

    const size_t plane = g.cols;
    const size_t N = (size_t)g.nx * plane;
    const size_t sy = y > 0 ? g.nz : 0;
    const size_t sz = z > 0 ? 1 : 0;
    uint32_t mask = 0;
    if (HAS_PML) mask = pml_column_mask(pml, y, z);

    size_t i = (size_t)x0 * plane + c;
    float hy_m, hz_m;  // Hy, Hz at x-1
    if (x0 > 0) {
        hy_m = curr[N + i - plane];
        hz_m = curr[2 * N + i - plane];
    } else {
        hy_m = curr[N + i];
        hz_m = curr[2 * N + i];
    }
    for (uint32_t x = x0; x < x1; ++x, i += plane) {
        const float hx = curr[i], hy = curr[N + i], hz = curr[2 * N + i];
        const float hz_ym = curr[2 * N + i - sy], hx_ym = curr[i - sy];
        const float hy_zm = curr[N + i - sz], hx_zm = curr[i - sz];
        const float cx = hz - hz_ym - hy + hy_zm;
        const float cy = hx - hx_zm - hz + hz_m;
        const float cz = hy - hy_m - hx + hx_ym;
        hy_m = hy;
        hz_m = hz;
        const float curl[3] = {cx, cy, cz};
        int k = -1;
        if (HAS_PML && mask) k = pml_find(pml, mask, x);
        if (!HAS_PML || k < 0) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const size_t j = n * N + i;
                volt[j] = vv[j] * volt[j] + vi[j] * curl[n];
            }
        } else {
            const PmlRegion& R = pml.r[k];
            const uint32_t p0 = ((x - R.x0) * R.ny + (y - R.y0)) * R.nz + (z - R.z0);
            const size_t T = 3 * (size_t)R.cs;
            const float* cf = pml_coef + R.coef;
            float* fl = pml_vflux + R.flux;
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const size_t j = n * N + i;
                const size_t p = (size_t)n * R.cs + p0;
                const float F = fl[p];
                const float Fn = vv[j] * F + vi[j] * curl[n];
                const float V = volt[j];
                const float f = cf[p] * V - cf[T + p] * F;  // vv_p*V - vvfo*F
                volt[j] = f + cf[2 * T + p] * Fn;           // + vvfn*F'
                fl[p] = Fn;
            }
        }
    }
}

// H sweep: I = ii*I + iv*curl(V) on [0,Nx-1) x [0,Ny-1) x [0,Nz-1).
template <bool HAS_PML>
__global__ __launch_bounds__(kSweepThreads) void sweep_h(float* __restrict__ curr,
                                                         const float* __restrict__ volt,
                                                         const float* __restrict__ ii,
                                                         const float* __restrict__ iv, SweepGeom g,
                                                         PmlTable pml,
                                                         const float* __restrict__ pml_coef,
                                                         float* __restrict__ pml_cflux) {
    const uint32_t L = xcd_remap(blockIdx.x, g.blocks);
    const uint32_t tile = L % g.col_tiles, chunk = L / g.col_tiles;
    const uint32_t c = tile * kSweepThreads + threadIdx.x;
    if (c >= g.cols) return;
    const uint32_t y = c / g.nz, z = c - y * g.nz;
    const uint32_t x0 = chunk * g.chunk_len;
    const uint32_t x1 = min(g.nx, x0 + g.chunk_len);
    if (x0 >= x1) return;

    const size_t plane = g.cols;
    const size_t N = (size_t)g.nx * plane;
    const bool col_core = (y + 1 < g.ny) && (z + 1 < g.nz);
    uint32_t mask = 0;
    if (HAS_PML) mask = pml_column_mask(pml, y, z);
    if (!col_core && !mask) return;

    size_t i = (size_t)x0 * plane + c;
    float ex = 0.f, ey = 0.f, ez = 0.f;  // E at x (carried from the previous plane)
    if (col_core) {
        ex = volt[i];
        ey = volt[N + i];
        ez = volt[2 * N + i];
    }
    for (uint32_t x = x0; x < x1; ++x, i += plane) {
        const bool core = col_core && (x + 1 < g.nx);
        int k = -1;
        if (HAS_PML && mask) k = pml_find(pml, mask, x);
        if (core) {
            const float ex_p = volt[i + plane], ey_p = volt[N + i + plane], ez_p = volt[2 * N + i + plane];
            const float ez_yp = volt[2 * N + i + g.nz], ey_zp = volt[N + i + 1];
            const float ex_zp = volt[i + 1], ex_yp = volt[i + g.nz];
            const float curl[3] = {ez - ez_yp - ey + ey_zp, ex - ex_zp - ez + ez_p, ey - ey_p - ex + ex_yp};
            ex = ex_p;
            ey = ey_p;
            ez = ez_p;
            if (!HAS_PML || k < 0) {
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    const size_t j = n * N + i;
                    curr[j] = ii[j] * curr[j] + iv[j] * curl[n];
                }
            } else {
                const PmlRegion& R = pml.r[k];
                const uint32_t p0 = ((x - R.x0) * R.ny + (y - R.y0)) * R.nz + (z - R.z0);
                const size_t T = 3 * (size_t)R.cs;
                const float* cf = pml_coef + R.coef + 3 * T;  // ii, iifo, iifn
                float* fl = pml_cflux + R.flux;
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    const size_t j = n * N + i;
                    const size_t p = (size_t)n * R.cs + p0;
                    const float F = fl[p];
                    const float Fn = ii[j] * F + iv[j] * curl[n];
                    const float I = curr[j];
                    const float f = cf[p] * I - cf[T + p] * F;
                    curr[j] = f + cf[2 * T + p] * Fn;
                    fl[p] = Fn;
                }
            }
        } else if (HAS_PML && k >= 0) {
            const PmlRegion& R = pml.r[k];
            const uint32_t p0 = ((x - R.x0) * R.ny + (y - R.y0)) * R.nz + (z - R.z0);
            const size_t T = 3 * (size_t)R.cs;
            const float* cf = pml_coef + R.coef + 3 * T;
            const float* fl = pml_cflux + R.flux;
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const size_t j = n * N + i;
                const size_t p = (size_t)n * R.cs + p0;
                const float F = fl[p];
                const float f = cf[p] * curr[j] - cf[T + p] * F;
                curr[j] = f + cf[2 * T + p] * F;
            }
        }
    }
}

}  // namespace efs
