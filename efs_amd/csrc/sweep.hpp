// sweep.hpp -- fused Yee E and H sweeps for gfx950 (CDNA4).
//
// Restates update_voltages / update_currents (src/webgpu-engine.mjs:7-122,
// CPU form src/webgpu-fdtd.mjs:64-155) with the UPML pre/post passes
// (update_pml modes 0-3, src/webgpu-engine.mjs:124-204) folded in.
//
// Mapping.  Fields keep the reference's NIJK order; for one x plane and one
// component the (y, z) cells form a contiguous run of Ny*Nz floats.  A lane
// owns 4 consecutive z cells (16-byte accesses), a workgroup a run of 1024
// cells of the plane, and marches along x, the slowest axis, over a chunk
// of planes.  The x-neighbour needed by the stencil stays in registers from
// the previous plane, so each H value (E sweep) or E value (H sweep) comes
// from HBM once; y and z neighbours belong to adjacent lanes or waves of
// the same or the neighbouring workgroup and are served by L1/L2.
// Work is cut into workgroups of one column tile and a chunk of planes by
// a host-built table of equal-cost blocks (see SweepBlock); workgroup ids
// are remapped so each XCD (own L2) gets a contiguous run of the table.
// Components are S floats apart (S >= Nx*Ny*Nz, padded).
//
// UPML cells are stored compactly per region (region-local NIJK, as the
// reference hands them over).  Default form, one launch per sweep
// (sweep_*4f): the grid launch updates every cell and branches per lane and
// plane into the UPML update; the branch is uniform over whole waves in x-
// and y-slabs and diverges only on the few lanes of a z-slab.  Every row of
// the main arrays is then read whole, once.  Alternative form (kept for
// comparison and tests): an interior launch that skips UPML cells
// (sweep_*4m) plus a launch over the region boxes (sweep_*4u).  Both need
// region z-bounds that are multiples of 4; other grids use the one-cell-
// per-lane sweeps (sweep_*1).
//
// Coefficients: two FP32 arrays (a, b) per component, or -- when every
// component has at most 256 distinct (a, b) pairs -- one byte per cell
// indexing a 256-entry table held in LDS.  Values and arithmetic are
// identical; only the bytes moved differ (3 B instead of 24 B per cell and
// sweep).
//
// UPML fold (per component, per PML cell; F = flux, exact algebra of
// pre -> core -> post):
//   F' = a*F + b*curl                          (the core update acts on F)
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
constexpr int kTab = 256;
#ifndef EFS_SWEEP_WAVES
#define EFS_SWEEP_WAVES 4  // waves per SIMD the interior sweeps are compiled for (128 registers)
#endif
#ifndef EFS_SWEEP_WAVES_PML
#define EFS_SWEEP_WAVES_PML 3  // UPML sweeps (168 registers)
#endif

typedef float v4 __attribute__((ext_vector_type(4)));

struct PmlRegion {
    uint32_t x0, x1, y0, y1, z0, z1;  // [x0,x1) x [y0,y1) x [z0,z1)
    uint32_t ny, nz;                  // local extents in y, z
    uint32_t cs;                      // nx*ny*nz (component stride of the region arrays)
    uint32_t blk0;                    // first workgroup of this region in the UPML launch
    uint32_t tiles;                   // lane tiles per plane (ceil(ny*nz/4 / 256))
    uint32_t chunk_len;               // x planes per workgroup
    uint32_t blocks;                  // tiles * x chunks
    uint32_t pad;
    uint64_t coef;                    // element offset of the region's 6 coefficient arrays
    uint64_t flux;                    // element offset of the region's flux arrays (3*cs each)
    uint64_t pidx;                    // byte offset of the region's UPML coefficient indices
                                      // (E components, then H components, cs bytes each)
};

struct PmlTable {
    uint32_t n;
    uint32_t blocks;  // total workgroups of the UPML launch
    PmlRegion r[kMaxPml];
};

// One workgroup of a grid launch: a column tile (256 lanes over
// contiguous (y, z) columns) and the x planes [x0, x1) it marches over.
// A block may take part of its tile (4-wide launches): the lanes
// [first, first + count) of it, with count in flag bits 4-15 (0: the whole
// tile) and first in bits 16-31.
struct SweepBlock {
    uint32_t tile, x0, x1;
    uint32_t flags;  // kBlockUpml: the block contains UPML cells; its lanes of the tile
};
constexpr uint32_t kBlockUpml = 1u;
__host__ __device__ constexpr uint32_t block_lane_first(uint32_t flags) { return flags >> 16; }
__host__ __device__ constexpr uint32_t block_lane_count(uint32_t flags) { return (flags >> 4) & 0xfffu; }
__host__ __device__ constexpr uint32_t block_lanes(uint32_t first, uint32_t count) {
    return (first << 16) | (count << 4);
}

// Grid launches read their work from a host-built block table: each column
// tile's x range is cut into chunks of equal bytes moved (UPML cells move
// flux and UPML coefficients too, so tiles and planes full of them get
// shorter chunks), and the table is ordered by x, so the contiguous runs of
// workgroups each XCD receives form x slabs of equal work.
struct SweepGeom {
    uint32_t nx, ny, nz;
    uint32_t cols;        // ny*nz
    uint32_t blocks;      // entries of tab
    uint32_t tile_lanes;  // lanes of a column tile (4-wide launches): 256, or whole rows of lanes
    uint64_t S;           // component stride in floats
    const SweepBlock* tab;
};

// NF2FF z-face capture by the H sweep.  Points of the NF2FF surface on a
// z plane read one cache line per (x, y) row and array in NIJK, for 2 of
// its values.  The H sweep already holds those rows: for each plane zf (zf
// and zf+1 inside one 4-cell lane) and each row (x, y) of the rectangle
// [x0, x1] x [y0, y1], the lane that owns them stores 12 floats into
// slab[(f * nx_c + x - x0) * ny_c + y - y0][12]:
//   E_x, E_y, E_z at zf and zf+1 (6), then H_x(zf), H_y(zf), H_z(zf),
//   H_z(zf+1) as stored by the sweep, 2 unused.
// The NF2FF sample kernel then reads those slabs instead of the rows.
// Unroll factor of the x march (build-time tuning knob; 1 = no unrolling;
// 2 measured no faster in FP32 and spills the FP16 fused H sweep).
#ifndef EFS_MARCH_UNROLL
#define EFS_MARCH_UNROLL 1
#endif
#define EFS_PRAGMA_STR2(x) #x
#define EFS_PRAGMA_STR(x) EFS_PRAGMA_STR2(x)
#define EFS_MARCH_UNROLL_PRAGMA _Pragma(EFS_PRAGMA_STR(unroll EFS_MARCH_UNROLL))

// Mur on the z = 0 and z = Nz-1 faces done inside the E march (batched
// form).  Slot s = 2f + n: face f (0: z = 0, shifted cell z = 1; 1: z =
// Nz-1, shifted cell z = Nz-2, both in one lane; Nz >= 6 keeps the two faces
// in different lanes) and tangential component n (0: x, 1: y), over the rows
// x0[s] <= x <= x1[s], y0[s] <= y <= y1[s] (empty: x0 > x1), coefficient
// c[s].  The march turns the slots of its row into per-lane registers before
// the loop, so the scalar copies are dead inside it.  The lane forms the pre value from the
// cells it loaded before the update and the post value from the shifted cell
// as stored, with the Mur kernels' fused multiply-adds: the same numbers as
// pre / post / apply in their own passes, for entries whose cells nothing
// else touches between sweep and apply.
struct ZMur {
    uint32_t on = 0;
    uint32_t nzr = 0;  // the real Nz (rows are padded)
    float c[4] = {0.f, 0.f, 0.f, 0.f};
    int32_t x0[4] = {1, 1, 1, 1}, x1[4] = {0, 0, 0, 0};
    int32_t y0[4] = {1, 1, 1, 1}, y1[4] = {0, 0, 0, 0};
};

struct NfCapture {
    float* slab;
    uint32_t on;          // 0: off
    uint32_t nf;          // planes: 1 or 2
    uint32_t z[2];
    uint32_t x0, x1, y0, y1;  // inclusive rectangle of rows
    uint32_t nx_c, ny_c;  // x1 - x0 + 1, y1 - y0 + 1
};
constexpr int kCapFloats = 12;

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

__device__ __forceinline__ v4 ld4(const float* p) { return *reinterpret_cast<const v4*>(p); }
__device__ __forceinline__ v4 ld4nt(const float* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
}
__device__ __forceinline__ float ldnt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st4(float* p, v4 v) { *reinterpret_cast<v4*>(p) = v; }

// Field storage.  FP32, or FP16 holding value * 2^-e per component: the
// scale s = 2^e is exact, loads widen to FP32 and multiply by s, stores
// multiply by 1/s and round to nearest.  The update arithmetic is FP32 in
// both cases.
typedef _Float16 f16;
typedef f16 h4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Fs;
template <>
struct Fs<float> {
    static constexpr bool kHalf = false;
    __device__ static __forceinline__ v4 ld(const float* p, float) { return ld4(p); }
    __device__ static __forceinline__ float ld1(const float* p, float) { return *p; }
    __device__ static __forceinline__ void st(float* p, v4 v, float) { st4(p, v); }
    // the value a later load returns for v stored with (s, 1/s)
    __device__ static __forceinline__ v4 stored(v4 v, float, float) { return v; }
    __device__ static __forceinline__ float stored1(float v, float, float) { return v; }
    // raw (stored-unit) forms for the 4-wide marches: value = raw * sc(s)
    __device__ static __forceinline__ float sc(float) { return 1.0f; }
    __device__ static __forceinline__ v4 ldr(const float* p) { return ld4(p); }
    __device__ static __forceinline__ float ldr1(const float* p) { return *p; }
    __device__ static __forceinline__ void str(float* p, v4 v) { st4(p, v); }
    __device__ static __forceinline__ v4 storedr(v4 v) { return v; }
    __device__ static __forceinline__ float storedr1(float v) { return v; }
};
// FP16 stores saturate at the largest finite half: a value that outgrew its
// scale between two range checks is clamped (and the next check rescales)
// instead of becoming infinite.
constexpr float kHalfMax = 65504.0f;
__device__ __forceinline__ v4 clamp16(v4 v) {
    return __builtin_elementwise_min(__builtin_elementwise_max(v, (v4)(-kHalfMax)), (v4)kHalfMax);
}
__device__ __forceinline__ float clamp16(float v) { return fminf(fmaxf(v, -kHalfMax), kHalfMax); }

template <>
struct Fs<f16> {
    static constexpr bool kHalf = true;
    __device__ static __forceinline__ v4 ld(const f16* p, float s) {
        return __builtin_convertvector(*reinterpret_cast<const h4*>(p), v4) * s;
    }
    __device__ static __forceinline__ float ld1(const f16* p, float s) { return (float)*p * s; }
    __device__ static __forceinline__ void st(f16* p, v4 v, float inv) {
        *reinterpret_cast<h4*>(p) = __builtin_convertvector(clamp16(v * inv), h4);
    }
    __device__ static __forceinline__ v4 stored(v4 v, float s, float inv) {
        return __builtin_convertvector(__builtin_convertvector(clamp16(v * inv), h4), v4) * s;
    }
    __device__ static __forceinline__ float stored1(float v, float s, float inv) {
        return (float)(f16)clamp16(v * inv) * s;
    }
    // raw (stored-unit) forms: the marches keep fields as stored values and
    // apply the power-of-two scales inside the curl, where they are exact,
    // so each load and store saves its multiply and the results are the
    // bits of the scaled form
    __device__ static __forceinline__ float sc(float s) { return s; }
    __device__ static __forceinline__ v4 ldr(const f16* p) {
        return __builtin_convertvector(*reinterpret_cast<const h4*>(p), v4);
    }
    __device__ static __forceinline__ float ldr1(const f16* p) { return (float)*p; }
    __device__ static __forceinline__ void str(f16* p, v4 v) {
        *reinterpret_cast<h4*>(p) = __builtin_convertvector(clamp16(v), h4);
    }
    __device__ static __forceinline__ v4 storedr(v4 v) {
        return __builtin_convertvector(__builtin_convertvector(clamp16(v), h4), v4);
    }
    __device__ static __forceinline__ float storedr1(float v) { return (float)(f16)clamp16(v); }
};

// A field array (volt or curr, 3 components S apart) as the point kernels
// see it: FP32, or FP16 with per-component scales (value = stored *
// scale[n], n = index / S; powers of two, so the scaling is exact).
struct FieldView {
    typedef float R;  // arithmetic type of the point kernels
    void* p;
    const float* scale;  // 3 scales of this array (FP16 storage)
    uint64_t S;
    uint32_t half;

    // component c, index j inside the component
    // (products not contracted into the caller's sums: a load returns the
    // value a store-and-reload would)
    __device__ __forceinline__ float ldc(uint32_t c, size_t j) const {
#pragma clang fp contract(off)
        if (!half) return static_cast<const float*>(p)[c * S + j];
        return (float)static_cast<const f16*>(p)[c * S + j] * scale[c];
    }
    __device__ __forceinline__ float ld(size_t i) const {
#pragma clang fp contract(off)
        if (!half) return static_cast<const float*>(p)[i];
        return (float)static_cast<const f16*>(p)[i] * scale[i / S];
    }
    __device__ __forceinline__ void st(size_t i, float v) const {
        if (!half) {
            static_cast<float*>(p)[i] = v;
            return;
        }
        static_cast<f16*>(p)[i] = (f16)clamp16(v / scale[i / S]);
    }
    // field[i] += v, safe when indices repeat: float atomics, or for FP16 a
    // compare-and-swap loop on the aligned 32-bit word holding the half.
    __device__ __forceinline__ void add(size_t i, float v) const {
        if (!half) {
            atomicAdd(static_cast<float*>(p) + i, v);
            return;
        }
        const float sc = scale[i / S];
        f16* hp = static_cast<f16*>(p) + i;
        unsigned int* w = reinterpret_cast<unsigned int*>(reinterpret_cast<uintptr_t>(hp) & ~(uintptr_t)3);
        const unsigned int sh = (reinterpret_cast<uintptr_t>(hp) & 2) ? 16u : 0u;
        unsigned int old = *w, assumed;
        do {
            assumed = old;
            const unsigned short bits = (unsigned short)(assumed >> sh);
            f16 h;
            __builtin_memcpy(&h, &bits, 2);
            const f16 nh = (f16)clamp16(((float)h * sc + v) / sc);
            unsigned short nb;
            __builtin_memcpy(&nb, &nh, 2);
            const unsigned int nw = (assumed & ~(0xffffu << sh)) | ((unsigned int)nb << sh);
            old = atomicCAS(w, assumed, nw);
        } while (old != assumed);
    }
};

// FP64 field storage (the precision reference run of the same engine):
// values and the point kernels' arithmetic in double.
struct FieldView64 {
    typedef double R;
    double* p;
    uint64_t S;
    __device__ __forceinline__ double ldc(uint32_t c, size_t j) const { return p[c * S + j]; }
    __device__ __forceinline__ double ld(size_t i) const { return p[i]; }
    __device__ __forceinline__ void st(size_t i, double v) const { p[i] = v; }
    __device__ __forceinline__ void add(size_t i, double v) const { atomicAdd(p + i, v); }
};

// Per-component scales of the two field arrays (FP16 storage): fscale[n]
// for volt component n, fscale[3+n] for curr.  own = the array the sweep
// writes, other = the array it reads.
struct Scales {
    float own[3], own_inv[3], other[3];
};

template <typename T>
__device__ __forceinline__ Scales load_scales(const float* __restrict__ fscale, bool own_is_curr) {
    Scales sc;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        if (Fs<T>::kHalf) {
            sc.own[n] = fscale[(own_is_curr ? 3 : 0) + n];
            sc.own_inv[n] = 1.0f / sc.own[n];  // exact: powers of two
            sc.other[n] = fscale[(own_is_curr ? 0 : 3) + n];
        } else {
            sc.own[n] = sc.own_inv[n] = sc.other[n] = 1.0f;
        }
    }
    return sc;
}

// The same scales read with device-coherent loads, for a kernel that
// rewrites them itself between two grid barriers (a uniform plain load may
// be served by the scalar cache, which a barrier does not invalidate).
template <typename T>
__device__ __forceinline__ Scales load_scales_coherent(const float* fscale, bool own_is_curr) {
    Scales sc;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        if (Fs<T>::kHalf) {
            const float* po = fscale + (own_is_curr ? 3 : 0) + n;
            const float* pt = fscale + (own_is_curr ? 0 : 3) + n;
            // uniform values: moved to scalar registers
            sc.own[n] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(
                                                      int, __hip_atomic_load(po, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))));
            sc.own_inv[n] = 1.0f / sc.own[n];
            sc.other[n] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(
                                                        int, __hip_atomic_load(pt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))));
        } else {
            sc.own[n] = sc.own_inv[n] = sc.other[n] = 1.0f;
        }
    }
    return sc;
}

// Explicit fused multiply-adds fix the rounding of every update, so the
// 1-wide and 4-wide sweeps (and the interior/UPML split) give identical bits:
//   core  V' = fma(a, V, b*curl)
//   UPML  F' = fma(a, F, b*curl),  V' = fma(vvfn, F', fma(vv_p, V, -(vvfo*F)))
__device__ __forceinline__ float fma1(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma1(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ v4 fma4(v4 a, v4 b, v4 c) { return __builtin_elementwise_fma(a, b, c); }

__device__ __forceinline__ v4 sel4(uint32_t bits, v4 t, v4 f) {
    return (v4){(bits & 1u) ? t.x : f.x, (bits & 2u) ? t.y : f.y, (bits & 4u) ? t.z : f.z,
                (bits & 8u) ? t.w : f.w};
}

// ---------------------------------------------------------------- coefficients

struct Coef {
    const float* a;      // FP32 arrays (3 components, stride S)
    const float* b;
    const uint8_t* idx;  // byte indices (3 components, stride S)
    const float2* tab;   // [3][256] pairs (global; staged into LDS)
    // UPML coefficients in the same form: per region byte indices (PmlRegion
    // pidx) into a [3][256] table of (vv, vvfo, vvfn, 0) -- or ii, iifo,
    // iifn for the H sweep -- when each component has at most 256 triples.
    const uint8_t* pidx;
    const float4* ptab;
    // 16-bit form (kCoefIdx16): one index per cell and component into a
    // [3][tabn] table of (a, b) pairs read through the caches (graded
    // meshes have a few hundred distinct pairs per component, more than
    // the byte form holds).
    const uint16_t* idx16;
    uint32_t tabn;
};

// Coefficient forms of the sweeps (template argument COMP).
enum : int {
    kCoefF32 = 0,    // two FP32 arrays
    kCoefIdx8 = 1,   // byte index, 256-entry table staged in LDS
    kCoefIdx16 = 2,  // 16-bit index, table in global memory (cached)
};

template <int COMP>
__device__ __forceinline__ void coef_stage(const Coef& cf, float2* s_tab) {
    if (COMP == kCoefIdx8) {
        for (int t = threadIdx.x; t < 3 * kTab; t += kSweepThreads) s_tab[t] = cf.tab[t];
        __syncthreads();
    }
}

// UPML table staging, only in workgroups whose block has UPML cells (the
// flag is uniform over the workgroup, so the barrier is too).
template <bool PCOMP>
__device__ __forceinline__ void upml_stage(const Coef& cf, float4* s_ptab, uint32_t flags) {
    if (PCOMP && (flags & kBlockUpml)) {
        for (int t = threadIdx.x; t < 3 * kTab; t += kSweepThreads) s_ptab[t] = cf.ptab[t];
        __syncthreads();
    }
}

// Two steps so a caller can issue every global load of an iteration before
// the first wait: coef_fetch loads the byte indices (or the FP32 arrays),
// coef_decode turns indices into values through the LDS table.
template <int COMP>
__device__ __forceinline__ void coef_fetch(const Coef& cf, size_t S, uint32_t i, uint32_t q[6], v4 a[3], v4 b[3]) {
    if (COMP == kCoefIdx8) {
#pragma unroll
        for (int n = 0; n < 3; ++n)
            q[n] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(cf.idx + n * S + i));
    } else if (COMP == kCoefIdx16) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            typedef uint32_t u2 __attribute__((ext_vector_type(2)));
            const u2 w = __builtin_nontemporal_load(reinterpret_cast<const u2*>(cf.idx16 + n * S + i));
            q[2 * n] = w.x;
            q[2 * n + 1] = w.y;
        }
    } else {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            a[n] = ld4nt(cf.a + n * S + i);
            b[n] = ld4nt(cf.b + n * S + i);
        }
    }
}

template <int COMP>
__device__ __forceinline__ void coef_decode(const float2* s_tab, const uint32_t q[6], v4 a[3], v4 b[3],
                                            const Coef& cf) {
    if (COMP == kCoefIdx16) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const float2* t = cf.tab + (size_t)n * cf.tabn;
            const float2 p0 = t[q[2 * n] & 0xffffu], p1 = t[q[2 * n] >> 16], p2 = t[q[2 * n + 1] & 0xffffu],
                         p3 = t[q[2 * n + 1] >> 16];
            a[n] = (v4){p0.x, p1.x, p2.x, p3.x};
            b[n] = (v4){p0.y, p1.y, p2.y, p3.y};
        }
    } else if (COMP == kCoefIdx8) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const float2* t = s_tab + n * kTab;
            const float2 p0 = t[q[n] & 255u], p1 = t[(q[n] >> 8) & 255u], p2 = t[(q[n] >> 16) & 255u],
                         p3 = t[q[n] >> 24];
            a[n] = (v4){p0.x, p1.x, p2.x, p3.x};
            b[n] = (v4){p0.y, p1.y, p2.y, p3.y};
        }
    }
}

// ---------------------------------------------------------------- 1 cell/lane

// One cell of component n with curl formed; k = UPML region or -1; core =
// whether the core update applies (always for E); HF selects the H-side
// UPML coefficients (ii, iifo, iifn follow vv, vvfo, vvfn in each region).
// V is the field type: float, or double for FP64 storage (coefficients stay
// the FP32 values they were given, widened; the arithmetic is in V).
template <bool HF, typename V>
__device__ __forceinline__ void cell1(V* __restrict__ fld, const float* __restrict__ a,
                                      const float* __restrict__ b, size_t j, bool core, V curl, int k,
                                      const PmlTable& pml, const float* __restrict__ pcoef,
                                      V* __restrict__ pflux, uint32_t n, uint32_t x, uint32_t y, uint32_t z) {
    if (k < 0) {
        if (core) fld[j] = fma1((V)ldnt(a + j), fld[j], (V)ldnt(b + j) * curl);
        return;
    }
    const PmlRegion& R = pml.r[k];
    const size_t p = (size_t)n * R.cs + ((size_t)(x - R.x0) * R.ny + (y - R.y0)) * R.nz + (z - R.z0);
    const size_t T = 3 * (size_t)R.cs;
    const float* cf = pcoef + R.coef + (HF ? 3 * T : 0);
    V* fl = pflux + R.flux;
    const V F = fl[p];
    const V Fn = core ? fma1((V)ldnt(a + j), F, (V)ldnt(b + j) * curl) : F;
    fld[j] = fma1((V)ldnt(cf + 2 * T + p), Fn, fma1((V)ldnt(cf + p), fld[j], -((V)ldnt(cf + T + p) * F)));
    fl[p] = Fn;
}

template <bool HAS_PML, typename V = float>
__global__ __launch_bounds__(kSweepThreads) void sweep_e1(V* __restrict__ volt,
                                                          const V* __restrict__ curr,
                                                          const float* __restrict__ vv,
                                                          const float* __restrict__ vi, SweepGeom g,
                                                          PmlTable pml, const float* __restrict__ pcoef,
                                                          V* __restrict__ pflux) {
    const SweepBlock d = g.tab[xcd_remap(blockIdx.x, g.blocks)];
    const uint32_t c = d.tile * kSweepThreads + threadIdx.x;
    if (c >= g.cols) return;
    const uint32_t y = c / g.nz, z = c - y * g.nz;
    const uint32_t x0 = d.x0;
    const uint32_t x1 = d.x1;
    if (x0 >= x1) return;
    const size_t plane = g.cols, S = g.S;
    const size_t sy = y > 0 ? g.nz : 0;
    const size_t sz = z > 0 ? 1 : 0;
    const uint32_t mask = HAS_PML ? pml_column_mask(pml, y, z) : 0u;

    size_t i = (size_t)x0 * plane + c;
    const size_t ib = x0 > 0 ? i - plane : i;
    V hy_m = curr[S + ib], hz_m = curr[2 * S + ib];  // Hy, Hz at x-1 (self at x = 0)
    for (uint32_t x = x0; x < x1; ++x, i += plane) {
        const V hx = curr[i], hy = curr[S + i], hz = curr[2 * S + i];
        const V hz_ym = curr[2 * S + i - sy], hx_ym = curr[i - sy];
        const V hy_zm = curr[S + i - sz], hx_zm = curr[i - sz];
        const V curl[3] = {hz - hz_ym - hy + hy_zm, hx - hx_zm - hz + hz_m, hy - hy_m - hx + hx_ym};
        hy_m = hy;
        hz_m = hz;
        const int k = (HAS_PML && mask) ? pml_find(pml, mask, x) : -1;
#pragma unroll
        for (int n = 0; n < 3; ++n)
            cell1<false, V>(volt, vv, vi, n * S + i, true, curl[n], k, pml, pcoef, pflux, n, x, y, z);
    }
}

template <bool HAS_PML, typename V = float>
__global__ __launch_bounds__(kSweepThreads) void sweep_h1(V* __restrict__ curr,
                                                          const V* __restrict__ volt,
                                                          const float* __restrict__ ii,
                                                          const float* __restrict__ iv, SweepGeom g,
                                                          PmlTable pml, const float* __restrict__ pcoef,
                                                          V* __restrict__ pflux) {
    const SweepBlock d = g.tab[xcd_remap(blockIdx.x, g.blocks)];
    const uint32_t c = d.tile * kSweepThreads + threadIdx.x;
    if (c >= g.cols) return;
    const uint32_t y = c / g.nz, z = c - y * g.nz;
    const uint32_t x0 = d.x0;
    const uint32_t x1 = d.x1;
    if (x0 >= x1) return;
    const size_t plane = g.cols, S = g.S;
    const bool col_core = (y + 1 < g.ny) && (z + 1 < g.nz);
    const uint32_t mask = HAS_PML ? pml_column_mask(pml, y, z) : 0u;
    if (!col_core && !mask) return;

    size_t i = (size_t)x0 * plane + c;
    V ex = 0, ey = 0, ez = 0;  // E at x, carried from the previous plane
    if (col_core) {
        ex = volt[i];
        ey = volt[S + i];
        ez = volt[2 * S + i];
    }
    for (uint32_t x = x0; x < x1; ++x, i += plane) {
        const bool core = col_core && (x + 1 < g.nx);
        const int k = (HAS_PML && mask) ? pml_find(pml, mask, x) : -1;
        if (!core && k < 0) continue;
        V curl[3] = {0, 0, 0};
        if (core) {
            const V ex_p = volt[i + plane], ey_p = volt[S + i + plane], ez_p = volt[2 * S + i + plane];
            const V ez_yp = volt[2 * S + i + g.nz], ey_zp = volt[S + i + 1];
            const V ex_zp = volt[i + 1], ex_yp = volt[i + g.nz];
            curl[0] = ez - ez_yp - ey + ey_zp;
            curl[1] = ex - ex_zp - ez + ez_p;
            curl[2] = ey - ey_p - ex + ex_yp;
            ex = ex_p;
            ey = ey_p;
            ez = ez_p;
        }
#pragma unroll
        for (int n = 0; n < 3; ++n)
            cell1<true, V>(curr, ii, iv, n * S + i, core, curl[n], k, pml, pcoef, pflux, n, x, y, z);
    }
}

// ---------------------------------------------------------------- 4 cells/lane

// Per-lane work of the 4-wide sweeps: the lane's (y, z0) column and its x
// range.  The grid launches derive it from the grid tiling, the region
// launch from the region boxes.
struct Lane4 {
    uint32_t y, z0, x0, x1;
    int region;  // UPML region of the lane (region launch), -1 otherwise
};

// What a march does with the UPML cells it meets.
enum : int {
    kMarchInterior = 0,  // grid launch, UPML planes skipped (another launch does them)
    kMarchRegion = 1,    // region launch: every plane is in region w.region
    kMarchFused = 2,     // grid launch doing both: per plane, core or UPML update
};

// A lane's position in the compact UPML arrays as it marches along x.
// The region lookup (a walk over the column's region mask, reading the
// region table) runs only when x reaches `until`, i.e. a few times per
// march; on the other planes the pointers advance by one region plane.
// This keeps the table reads off the per-plane path, where they would
// delay the UPML loads by a round trip.
struct PmlCursor {
    int k;            // region of the current plane, -1 = none
    uint32_t until;   // the state holds for planes < until
    float* flux;      // flux of component 0 at the lane's first cell
    const float* coef;  // first of the three coefficient arrays (component 0)
    const uint8_t* pq;  // coefficient indices of component 0 (compressed UPML)
    uint32_t cs;      // component stride of the region arrays
    uint32_t step;    // cells per region plane (ny*nz of the region)
};

template <bool HF>
__device__ __forceinline__ void cursor_seek(PmlCursor& u, const PmlTable& pml, uint32_t m, uint32_t x, uint32_t y,
                                            uint32_t z0, const float* __restrict__ pcoef,
                                            float* __restrict__ pflux, const uint8_t* __restrict__ pidx) {
    u.k = -1;
    u.until = 0xffffffffu;
    while (m) {
        const int k = __builtin_ctz(m);
        m &= m - 1u;
        const PmlRegion& R = pml.r[k];
        if (x >= R.x0 && x < R.x1) {
            const size_t p0 = ((size_t)(x - R.x0) * R.ny + (y - R.y0)) * R.nz + (z0 - R.z0);
            u.k = k;
            u.until = R.x1;
            u.cs = R.cs;
            u.step = R.ny * R.nz;
            u.flux = pflux + R.flux + p0;
            u.coef = pcoef + R.coef + (HF ? 9 * (size_t)R.cs : 0) + p0;
            u.pq = pidx + R.pidx + (HF ? 3 * (size_t)R.cs : 0) + p0;
            return;
        }
        if (R.x0 > x) u.until = min(u.until, R.x0);
    }
}

// Position the cursor on plane x (called for consecutive x).
template <bool HF>
__device__ __forceinline__ void cursor_at(PmlCursor& u, const PmlTable& pml, uint32_t m, uint32_t x, uint32_t y,
                                          uint32_t z0, const float* __restrict__ pcoef,
                                          float* __restrict__ pflux, const uint8_t* __restrict__ pidx) {
    if (x >= u.until) {
        cursor_seek<HF>(u, pml, m, x, y, z0, pcoef, pflux, pidx);
    } else if (u.k >= 0) {
        u.flux += u.step;
        u.coef += u.step;
        u.pq += u.step;
    }
}

// UPML operands of one lane and plane: flux always; coefficients either
// from the FP32 region arrays or, compressed, as byte indices decoded later.
template <bool PCOMP>
__device__ __forceinline__ void upml_fetch(const PmlCursor& u, v4 F[3], v4 P0[3], v4 P1[3], v4 P2[3],
                                           uint32_t pq[3]) {
    const size_t T = 3 * (size_t)u.cs;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        F[n] = ld4(u.flux + n * u.cs);
        if (PCOMP) {
            pq[n] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(u.pq + n * u.cs));
        } else {
            P0[n] = ld4nt(u.coef + n * u.cs);
            P1[n] = ld4nt(u.coef + T + n * u.cs);
            P2[n] = ld4nt(u.coef + 2 * T + n * u.cs);
        }
    }
}

template <bool PCOMP>
__device__ __forceinline__ void upml_decode(const float4* s_ptab, const uint32_t pq[3], v4 P0[3], v4 P1[3],
                                            v4 P2[3]) {
    if (PCOMP) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const float4* t = s_ptab + n * kTab;
            const float4 c0 = t[pq[n] & 255u], c1 = t[(pq[n] >> 8) & 255u], c2 = t[(pq[n] >> 16) & 255u],
                         c3 = t[pq[n] >> 24];
            P0[n] = (v4){c0.x, c1.x, c2.x, c3.x};
            P1[n] = (v4){c0.y, c1.y, c2.y, c3.y};
            P2[n] = (v4){c0.z, c1.z, c2.z, c3.z};
        }
    }
}

// Cells of a block's first two planes kept in LDS for sweep_epilogue:
// [plane][component][4 * lane + cell of the lane], the stored (raw) values.
constexpr uint32_t kEpLdsPlanes = 2, kEpLdsCells = 4 * kSweepThreads;
constexpr uint32_t kEpLdsFloats = kEpLdsPlanes * 3 * kEpLdsCells;

// E update of 4 cells over x in [x0, x1); MODE says how UPML cells are
// treated (column mask m gives the regions the lane's column crosses).
// ep_lds (null: none): where the first two planes' new values go for the
// block's epilogue.
template <int MODE, int COMP, bool PCOMP, typename T>
__device__ __forceinline__ void e4_march(T* __restrict__ volt, const T* __restrict__ curr, const Coef& cf,
                                         const float2* s_tab, const float4* s_ptab, const SweepGeom& g,
                                         const PmlTable& pml,
                                         const float* __restrict__ pcoef, float* __restrict__ pflux, Lane4 w,
                                         uint32_t m, const Scales& sc, const NfCapture&, const ZMur& zm = ZMur{},
                                         float* __restrict__ ep_lds = nullptr) {
    typedef Fs<T> L;
    const size_t S = g.S;
    T* const E[3] = {volt, volt + S, volt + 2 * S};
    const T* const H0 = curr;
    const T* const H1 = curr + S;
    const T* const H2 = curr + 2 * S;
    const uint32_t plane = g.cols;
    const uint32_t sy = w.y > 0 ? g.nz : 0u;  // backward neighbours clamp to self at index 0
    const uint32_t sz = w.z0 > 0 ? 1u : 0u;
    uint32_t i = w.x0 * plane + w.y * g.nz + w.z0;
    bool carry = false;  // hy_m/hz_m hold H at x-1
    v4 hy_m = 0.f, hz_m = 0.f;
    // z-face Mur of this lane's row: normal / shifted cell at lane offsets
    // zo / zs; component n applies when x - zxl[n] <= zxs[n] (unsigned; an
    // idle component has zxl = 2^31, zxs = 0)
    bool zany = false;
    uint32_t zo = 0, zs = 1, zxl[2] = {0x80000000u, 0x80000000u}, zxs[2] = {0u, 0u};
    float zc[2] = {0.f, 0.f};
    if (zm.on) {
        // slots picked with selects between the two faces' (uniform) values,
        // not by a lane-dependent index into the struct
        const bool f0 = w.z0 == 0;
        const bool f1 = !f0 && zm.nzr >= 6 && zm.nzr - 1 >= w.z0 && zm.nzr - 1 < w.z0 + 4 && (zm.nzr - 1) % 4 != 0;
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
    PmlCursor u{-1, m ? 0u : 0xffffffffu, nullptr, nullptr, nullptr, 0u, 0u};
    auto keep = [&](uint32_t x, int n, const v4& out) {
        if (ep_lds && x - w.x0 < kEpLdsPlanes)
            *reinterpret_cast<v4*>(ep_lds + ((x - w.x0) * 3 + n) * kEpLdsCells + 4 * threadIdx.x) = L::storedr(out);
    };
    EFS_MARCH_UNROLL_PRAGMA
    for (uint32_t x = w.x0; x < w.x1; ++x, i += plane) {
        cursor_at<false>(u, pml, m, x, w.y, w.z0, pcoef, pflux, cf.pidx);
        const bool upml = MODE == kMarchRegion || (MODE == kMarchFused && u.k >= 0);
        if (MODE == kMarchInterior && u.k >= 0) {
            carry = false;
            continue;
        }
        // fields in stored units (raw); value = raw * L::sc(scale)
        if (!carry) {
            const uint32_t ib = x > 0 ? i - plane : i;
            hy_m = L::ldr(H1 + ib);
            hz_m = L::ldr(H2 + ib);
        }
        const v4 hx = L::ldr(H0 + i), hy = L::ldr(H1 + i), hz = L::ldr(H2 + i);
        const v4 hz_ym = L::ldr(H2 + i - sy), hx_ym = L::ldr(H0 + i - sy);
        const float hy_z = L::ldr1(H1 + i - sz), hx_z = L::ldr1(H0 + i - sz);  // z-1
        v4 V[3], a[3], b[3];
        uint32_t q[6];
#pragma unroll
        for (int n = 0; n < 3; ++n) V[n] = L::ldr(E[n] + i);
        coef_fetch<COMP>(cf, S, i, q, a, b);
        v4 F[3], P0[3], P1[3], P2[3];
        uint32_t pq[3];
        if (upml) upml_fetch<PCOMP>(u, F, P0, P1, P2, pq);
        coef_decode<COMP>(s_tab, q, a, b, cf);
        const v4 hy_zm = (v4){hy_z, hy.x, hy.y, hy.z};
        const v4 hx_zm = (v4){hx_z, hx.x, hx.y, hx.z};
        // curl in real units: ((hz - hz_ym) - hy) + hy_zm and its two
        // rotations, each raw term scaled exactly by its component's scale
        const float s0 = L::sc(sc.other[0]), s1 = L::sc(sc.other[1]), s2 = L::sc(sc.other[2]);
        const v4 cur[3] = {fma4((v4)s1, hy_zm, fma4((v4)(-s1), hy, s2 * (hz - hz_ym))),
                           fma4((v4)s2, hz_m, fma4((v4)(-s2), hz, s0 * (hx - hx_zm))),
                           fma4((v4)s0, hx_ym, fma4((v4)(-s0), hx, s1 * (hy - hy_m)))};
        hy_m = hy;
        hz_m = hz;
        carry = true;
        if (!upml) {
            // raw update: a * V_raw + (b * curl) / s_own (exact scaling)
            v4 vn[3];
#pragma unroll
            for (int n = 0; n < 3; ++n) vn[n] = fma4(a[n], V[n], (b[n] * cur[n]) * L::sc(sc.own_inv[n]));
            if (zany) {
                // Mur on scaled values commutes with the exact scaling
#pragma unroll
                for (int n = 0; n < 2; ++n) {
                    if (x - zxl[n] > zxs[n]) continue;
                    const float pre = __builtin_fmaf(-zc[n], V[n][zo], V[n][zs]);
                    const float st = L::storedr1(vn[n][zs]);
                    vn[n][zo] = __builtin_fmaf(zc[n], st, pre);
                }
            }
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                L::str(E[n] + i, vn[n]);
                keep(x, n, vn[n]);
            }
        } else {
            upml_decode<PCOMP>(s_ptab, pq, P0, P1, P2);
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const v4 Fn = fma4(a[n], F[n], b[n] * cur[n]);
                const v4 Vs = V[n] * L::sc(sc.own[n]);
                const v4 out = fma4(P2[n], Fn, fma4(P0[n], Vs, -(P1[n] * F[n]))) * L::sc(sc.own_inv[n]);
                L::str(E[n] + i, out);
                keep(x, n, out);
                st4(u.flux + n * u.cs, Fn);
            }
        }
    }
}

// H update of 4 cells over x in [x0, x1) (see e4_march).  Cells outside the
// core bounds (x = Nx-1, y = Ny-1, z = Nz-1) keep I unless in UPML.
template <int MODE, int COMP, bool PCOMP, typename T>
__device__ __forceinline__ void h4_march(T* __restrict__ curr, const T* __restrict__ volt, const Coef& cf,
                                         const float2* s_tab, const float4* s_ptab, const SweepGeom& g,
                                         const PmlTable& pml,
                                         const float* __restrict__ pcoef, float* __restrict__ pflux, Lane4 w,
                                         uint32_t m, const Scales& sc, const NfCapture& cap, const ZMur& = ZMur{},
                                         float* = nullptr) {
    typedef Fs<T> L;
    const size_t S = g.S;
    T* const Ic[3] = {curr, curr + S, curr + 2 * S};
    // NF2FF capture: planes of the capture inside this lane (bit f), when
    // the lane's row lies in the rectangle
    uint32_t capz = 0;
    if (cap.on && w.y - cap.y0 <= cap.y1 - cap.y0)
        for (uint32_t f = 0; f < cap.nf; ++f)
            if (cap.z[f] >= w.z0 && cap.z[f] + 1 < w.z0 + 4) capz |= 1u << f;
    // the lane's capture plane and its offset in the lane (zf and zf+1 at
    // capo, capo + 1 <= 3)
    const uint32_t capf = (capz & 1u) ? 0u : 1u;
    const uint32_t capo = capz ? cap.z[capf] - w.z0 : 0u;
    const T* const E0 = volt;
    const T* const E1 = volt + S;
    const T* const E2 = volt + 2 * S;
    const uint32_t plane = g.cols;
    const bool row_core = w.y + 1 < g.ny;
    const bool last = w.z0 + 4 >= g.nz;  // the lane holds the row's last z cell (no core update there)
    if (MODE == kMarchInterior && !row_core) return;
    if (MODE == kMarchFused && !row_core && !m) return;
    const uint32_t sy = row_core ? g.nz : 0u;  // clamped forward neighbours stay in bounds
    const uint32_t sz = last ? 3u : 4u;
    uint32_t i = w.x0 * plane + w.y * g.nz + w.z0;
    bool carry = false;  // ex/ey/ez hold E at x
    v4 ex = 0.f, ey = 0.f, ez = 0.f;
    PmlCursor u{-1, m ? 0u : 0xffffffffu, nullptr, nullptr, nullptr, 0u, 0u};
    EFS_MARCH_UNROLL_PRAGMA
    for (uint32_t x = w.x0; x < w.x1; ++x, i += plane) {
        const bool core = row_core && (x + 1 < g.nx);
        cursor_at<true>(u, pml, m, x, w.y, w.z0, pcoef, pflux, cf.pidx);
        const bool upml = MODE == kMarchRegion || (MODE == kMarchFused && u.k >= 0);
        if (MODE != kMarchRegion && !upml && (!core || u.k >= 0)) {
            carry = false;
            continue;
        }
        // fields in stored units (raw); value = raw * L::sc(scale)
        if (!carry) {
            ex = L::ldr(E0 + i);
            ey = L::ldr(E1 + i);
            ez = L::ldr(E2 + i);
        }
        // per-cell core flags: cells 0..2 follow `core`, cell 3 also needs z+1 < Nz
        const uint32_t cb = core ? (last ? 7u : 15u) : 0u;
        const uint32_t sx = core ? plane : 0u;
        const v4 ex_p = L::ldr(E0 + i + sx), ey_p = L::ldr(E1 + i + sx), ez_p = L::ldr(E2 + i + sx);
        const v4 ez_yp = L::ldr(E2 + i + sy), ex_yp = L::ldr(E0 + i + sy);
        const float ey_z = L::ldr1(E1 + i + sz), ex_z = L::ldr1(E0 + i + sz);  // z+1
        v4 I[3], a[3], b[3];
        uint32_t q[6];
#pragma unroll
        for (int n = 0; n < 3; ++n) I[n] = L::ldr(Ic[n] + i);
        coef_fetch<COMP>(cf, S, i, q, a, b);
        v4 F[3], P0[3], P1[3], P2[3];  // P: ii, iifo, iifn
        uint32_t pq[3];
        if (upml) upml_fetch<PCOMP>(u, F, P0, P1, P2, pq);
        coef_decode<COMP>(s_tab, q, a, b, cf);
        const v4 ey_zp = (v4){ey.y, ey.z, ey.w, ey_z};
        const v4 ex_zp = (v4){ex.y, ex.z, ex.w, ex_z};
        // curl in real units: ((ez - ez_yp) - ey) + ey_zp and its two
        // rotations, each raw term scaled exactly by its component's scale
        const float s0 = L::sc(sc.other[0]), s1 = L::sc(sc.other[1]), s2 = L::sc(sc.other[2]);
        const v4 cur[3] = {fma4((v4)s1, ey_zp, fma4((v4)(-s1), ey, s2 * (ez - ez_yp))),
                           fma4((v4)s2, ez_p, fma4((v4)(-s2), ez, s0 * (ex - ex_zp))),
                           fma4((v4)s0, ex_yp, fma4((v4)(-s0), ex, s1 * (ey - ey_p)))};
        const v4 e_x[3] = {ex, ey, ez};  // E at this plane (raw), for the capture
        ex = ex_p;
        ey = ey_p;
        ez = ez_p;
        carry = core;
        v4 In[3];  // new I, raw
        if (!upml) {
            // raw update: a * I_raw + (b * curl) / s_own (exact scaling)
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                In[n] = sel4(cb, fma4(a[n], I[n], (b[n] * cur[n]) * L::sc(sc.own_inv[n])), I[n]);
                L::str(Ic[n] + i, In[n]);
            }
        } else {
            upml_decode<PCOMP>(s_ptab, pq, P0, P1, P2);
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const v4 Fn = sel4(cb, fma4(a[n], F[n], b[n] * cur[n]), F[n]);
                const v4 Is = I[n] * L::sc(sc.own[n]);
                In[n] = fma4(P2[n], Fn, fma4(P0[n], Is, -(P1[n] * F[n]))) * L::sc(sc.own_inv[n]);
                L::str(Ic[n] + i, In[n]);
                st4(u.flux + n * u.cs, Fn);
            }
        }
        if (capz && x - cap.x0 <= cap.x1 - cap.x0) {
            // a lane holds at most one capture plane (the host keeps two
            // planes in different lanes): the planes of all lanes share one
            // block, so a wave runs it once
            {
                const uint32_t f = capf, o = capo;
                const bool o1 = o == 1u, o2 = o == 2u;
                // elements o and o + 1 of a lane vector
                auto lo = [&](const v4& v) { return o2 ? v.z : (o1 ? v.y : v.x); };
                auto hi = [&](const v4& v) { return o2 ? v.w : (o1 ? v.z : v.y); };
                float* q = cap.slab +
                           ((size_t)(f * cap.nx_c + (x - cap.x0)) * cap.ny_c + (w.y - cap.y0)) * kCapFloats;
                // real units: E (raw) and H as stored, times their scales
                const v4 hx = L::storedr(In[0]) * L::sc(sc.own[0]);
                const v4 hy = L::storedr(In[1]) * L::sc(sc.own[1]);
                const v4 hz = L::storedr(In[2]) * L::sc(sc.own[2]);
                const float t0 = L::sc(sc.other[0]), t1 = L::sc(sc.other[1]), t2 = L::sc(sc.other[2]);
                const v4 q0 = {lo(e_x[0]) * t0, hi(e_x[0]) * t0, lo(e_x[1]) * t1, hi(e_x[1]) * t1};
                const v4 q1 = {lo(e_x[2]) * t2, hi(e_x[2]) * t2, lo(hx), lo(hy)};
                const v4 q2 = {lo(hz), hi(hz), 0.f, 0.f};
                st4(q, q0);
                st4(q + 4, q1);
                st4(q + 8, q2);
            }
        }
    }
}

// Interior launch: grid tiling of the (y, z) plane, 1024 cells per workgroup.
// bid: the sweep block (the workgroup index after any side-work groups).
template <bool HAS_PML>
// flags: the block's flags (uniform over the workgroup, set on every path).
__device__ __forceinline__ bool interior_lane(const SweepGeom& g, const PmlTable& pml, Lane4& w, uint32_t& m,
                                              uint32_t& flags, uint32_t bid) {
    const SweepBlock d = g.tab[xcd_remap(bid, g.blocks)];
    flags = d.flags;
    m = 0u;
    const uint32_t n = block_lane_count(d.flags);
    const uint32_t c0 = (d.tile * g.tile_lanes + block_lane_first(d.flags) + threadIdx.x) * 4u;
    if (threadIdx.x >= (n ? n : g.tile_lanes) || c0 >= g.cols) return false;
    w.y = c0 / g.nz;
    w.z0 = c0 - w.y * g.nz;
    w.x0 = d.x0;
    w.x1 = d.x1;
    w.region = -1;
    m = HAS_PML ? pml_column_mask(pml, w.y, w.z0) : 0u;  // region z-bounds are multiples of 4
    return w.x0 < w.x1;
}

// UPML launch: workgroups [blk0, blk0 + blocks) of region k tile its local
// (y, z) plane (4 cells per lane) and march a chunk of its x range.
__device__ __forceinline__ bool pml_lane(const PmlTable& pml, Lane4& w) {
    const uint32_t b = xcd_remap(blockIdx.x, pml.blocks);
    int k = 0;
    while (k + 1 < (int)pml.n && b >= pml.r[k + 1].blk0) ++k;
    const PmlRegion& R = pml.r[k];
    const uint32_t lb = b - R.blk0;
    const uint32_t tile = lb % R.tiles, chunk = lb / R.tiles;
    const uint32_t gz = R.nz / 4;
    const uint32_t lane = tile * kSweepThreads + threadIdx.x;
    if (lane >= R.ny * gz) return false;
    const uint32_t ly = lane / gz;
    w.y = R.y0 + ly;
    w.z0 = R.z0 + 4 * (lane - ly * gz);
    w.x0 = R.x0 + chunk * R.chunk_len;
    w.x1 = min(R.x1, w.x0 + R.chunk_len);
    w.region = k;
    return w.x0 < w.x1;
}

// side: point work of extra workgroups at the start of the grid (SideWork,
// defined with the point kernels).  Each grid entry point comes in two
// kernels: NAME, without side work (the code of the sweep alone), and
// NAME_s, whose first side.nb workgroups run side_work and whose E-sweep
// blocks end with sweep_epilogue (Mur post and apply and the excitation of
// the block's own cells; the extra code costs the fused FP32 UPML H sweeps
// a spill, so grids that bring neither launch NAME).
#define EFS_SWEEP4_PARAMS(T)                                                                                    \
    T *__restrict__ dst, const T *__restrict__ src, Coef cf, SweepGeom g, PmlTable pml,                        \
        const float *__restrict__ pcoef, float *__restrict__ pflux, const float *__restrict__ fscale, NfCapture cap, \
        SideWork side
#define EFS_SWEEP4_ARGS dst, src, cf, g, pml, pcoef, pflux, fscale, cap, side
#define EFS_SWEEP4_SIDE(HF)                      \
    uint32_t bid = blockIdx.x;                   \
    if (SIDE) {                                  \
        if (blockIdx.x < side.nb) {              \
            side_work<HF>(side, blockIdx.x);     \
            return;                              \
        }                                        \
        bid -= side.nb;                          \
    }
// the two kernels of a body NAME##_body<SIDE>
#define EFS_SWEEP4_KERNELS(NAME, T, WAVES)                                                                         \
    __global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void NAME(     \
        EFS_SWEEP4_PARAMS(T)) {                                                                                    \
        NAME##_body<false>(EFS_SWEEP4_ARGS);                                                                       \
    }                                                                                                              \
    __global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void NAME##_s( \
        EFS_SWEEP4_PARAMS(T)) {                                                                                    \
        NAME##_body<true>(EFS_SWEEP4_ARGS);                                                                        \
    }

// Interior (4 waves/SIMD) and UPML (3 waves/SIMD) entry points, each with
// FP32 arrays or table-compressed (suffix c) coefficients; the interior
// kernels with suffix m skip the UPML regions.  FP32 field storage.
#define EFS_SWEEP4_INTERIOR(NAME, MARCH, HF, HAS_PML, COMP)                                                       \
    template <bool SIDE>                                                                                           \
    __device__ __forceinline__ void NAME##_body(EFS_SWEEP4_PARAMS(float)) {                                        \
        EFS_SWEEP4_SIDE(HF)                                                                                        \
        __shared__ float2 s_tab[COMP == kCoefIdx8 ? 3 * kTab : 1];                                                 \
        __shared__ float s_ep[SIDE && !HF ? kEpLdsFloats : 1];                                                     \
        coef_stage<COMP>(cf, s_tab);                                                                               \
        Lane4 w;                                                                                                   \
        uint32_t m, flags;                                                                                         \
        if (interior_lane<HAS_PML>(g, pml, w, m, flags, bid)) {                                                    \
            const Scales sc = load_scales<float>(fscale, HF);                                                      \
            MARCH<kMarchInterior, COMP, false, float>(dst, src, cf, s_tab, nullptr, g, pml, pcoef, pflux, w, m, sc, \
                                                      cap, ZMur{}, SIDE && !HF && side.ep_on ? s_ep : nullptr);    \
        }                                                                                                          \
        if (SIDE && !HF) sweep_epilogue(side, xcd_remap(bid, g.blocks), s_ep);                                     \
    }                                                                                                              \
    EFS_SWEEP4_KERNELS(NAME, float, EFS_SWEEP_WAVES)

#define EFS_SWEEP4_UPML(NAME, MARCH, HF, COMP)                                                                     \
    __global__ __launch_bounds__(kSweepThreads) __attribute__((                                                    \
        amdgpu_waves_per_eu(EFS_SWEEP_WAVES_PML, EFS_SWEEP_WAVES_PML))) void NAME(EFS_SWEEP4_PARAMS(float)) {     \
        __shared__ float2 s_tab[COMP == kCoefIdx8 ? 3 * kTab : 1];                                                 \
        coef_stage<COMP>(cf, s_tab);                                                                               \
        Lane4 w;                                                                                                   \
        if (!pml_lane(pml, w)) return;                                                                             \
        const Scales sc = load_scales<float>(fscale, HF);                                                          \
        MARCH<kMarchRegion, COMP, false, float>(dst, src, cf, s_tab, nullptr, g, pml, pcoef, pflux, w,             \
                                                1u << w.region, sc, NfCapture{});                                  \
    }

// Fused: the grid launch does the UPML cells too (3 waves/SIMD, the UPML
// register budget); PCOMP: UPML coefficients as byte indices + LDS table;
// T: field storage (float or f16).
#define EFS_SWEEP4_FUSED(NAME, MARCH, HF, COMP, PCOMP, T)                                                          \
    template <bool SIDE>                                                                                           \
    __device__ __forceinline__ void NAME##_body(EFS_SWEEP4_PARAMS(T)) {                                            \
        EFS_SWEEP4_SIDE(HF)                                                                                        \
        __shared__ float2 s_tab[COMP == kCoefIdx8 ? 3 * kTab : 1];                                                 \
        __shared__ float4 s_ptab[PCOMP ? 3 * kTab : 1];                                                            \
        __shared__ float s_ep[SIDE && !HF ? kEpLdsFloats : 1];                                                     \
        coef_stage<COMP>(cf, s_tab);                                                                               \
        Lane4 w;                                                                                                   \
        uint32_t m, flags;                                                                                         \
        const bool live = interior_lane<true>(g, pml, w, m, flags, bid);                                           \
        upml_stage<PCOMP>(cf, s_ptab, flags);                                                                      \
        if (live) {                                                                                                \
            const Scales sc = load_scales<T>(fscale, HF);                                                          \
            MARCH<kMarchFused, COMP, PCOMP, T>(dst, src, cf, s_tab, s_ptab, g, pml, pcoef, pflux, w, m, sc, cap,   \
                                               ZMur{}, SIDE && !HF && side.ep_on ? s_ep : nullptr);                \
        }                                                                                                          \
        if (SIDE && !HF) sweep_epilogue(side, xcd_remap(bid, g.blocks), s_ep);                                     \
    }                                                                                                              \
    EFS_SWEEP4_KERNELS(NAME, T, EFS_SWEEP_WAVES_PML)

// Grid launch without UPML, FP16 field storage (FP32 storage uses the
// interior kernels above, which are the same code for HAS_PML = false).
#define EFS_SWEEP4_PLAIN16(NAME, MARCH, HF, COMP)                                                                  \
    template <bool SIDE>                                                                                           \
    __device__ __forceinline__ void NAME##_body(EFS_SWEEP4_PARAMS(f16)) {                                          \
        EFS_SWEEP4_SIDE(HF)                                                                                        \
        __shared__ float2 s_tab[COMP == kCoefIdx8 ? 3 * kTab : 1];                                                 \
        __shared__ float s_ep[SIDE && !HF ? kEpLdsFloats : 1];                                                     \
        coef_stage<COMP>(cf, s_tab);                                                                               \
        Lane4 w;                                                                                                   \
        uint32_t m, flags;                                                                                         \
        if (interior_lane<false>(g, pml, w, m, flags, bid)) {                                                      \
            const Scales sc = load_scales<f16>(fscale, HF);                                                        \
            MARCH<kMarchInterior, COMP, false, f16>(dst, src, cf, s_tab, nullptr, g, pml, pcoef, pflux, w, m, sc,  \
                                                    cap, ZMur{}, SIDE && !HF && side.ep_on ? s_ep : nullptr);      \
        }                                                                                                          \
        if (SIDE && !HF) sweep_epilogue(side, xcd_remap(bid, g.blocks), s_ep);                                     \
    }                                                                                                              \
    EFS_SWEEP4_KERNELS(NAME, f16, EFS_SWEEP_WAVES)

}  // namespace efs
