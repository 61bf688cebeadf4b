// nf2ff.hpp -- NF2FF surface sampling, K-step DFT fold and far field.
//
// The reference accumulates one frequency every step with an FP32 phase
// (NF2FF accumulate WGSL, src/webgpu-engine.mjs:637-800).  Here the
// cell-interpolated surface E/H of K consecutive steps are written to a
// buffer B[K][rows] and folded into F frequency bins at once:
//     acc[row][2f + {0: re, 1: im}] += sum_k B[k][row] * W[h][k][2f + {0, 1}]
// a real [rows x K] x [K x 2F] product (the complex product with re/im
// interleaved in the columns) on the matrix cores, whose twiddles
//     W[h][k][f] = exp(-j*omega_f*t_k) * w(n_k) * windowNorm * dT
// (t_k = n_k*dT for E rows, h = 0; (n_k+1/2)*dT for H rows, h = 1) are
// generated in FP64.  Row c*P + p holds E component c of point p, row
// rowsE + c*P + p H component c; rowsE = 3P rounded up to kFoldRows, so a
// fold workgroup's rows are all E or all H.
//
// Compensated accumulation (default): each fold adds its block sum D into
// the running acc with an error-free transformation (Knuth's TwoSum),
//     s = acc + D,  e = (acc - (s - b')) + (D - b'),  b' = s - acc,
// keeping acc = s and adding the rounding error e into a second array cmp,
// so acc + cmp carries the running DFT to about FP32 rounding of the
// total instead of an error that grows with the number of folds (a 30k-step
// run folds ~470 times).  Readers take acc + cmp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep.hpp"

namespace efs {

constexpr int kFoldRows = 512;  // rows per fold workgroup: 4 waves x 128
constexpr int kFoldCols = 32;   // columns per fold workgroup (16 frequencies)

struct Nf2ffGeom {
    uint32_t nx, ny, nz, P;
    uint32_t el_off[6];  // offsets of primal x,y,z / dual x,y,z edge lengths
    uint32_t rowsE;      // first H row
    uint32_t nzs;        // device row length (Nz padded to a multiple of 4)
    uint64_t S;          // component stride of the field arrays
};

// Cell interpolation of E (mean of 4 edges / primal length; 0 on the last
// grid line) and H (lerp of 2 edges / dual length along its own axis),
// as in the NF2FF accumulate WGSL.  V(c, dx, dy, dz) returns volt component
// c at (ix+dx, iy+dy, iz+dz); I(n, d) returns curr component n at the point
// or (d = 1) one cell further along axis n.  One body for both sources of
// the values keeps their arithmetic identical; the E and H halves are
// separate so that they can run at different places in the step.
// R: arithmetic type (float; double for FP64 field storage).
template <class R, class VA>
__device__ __forceinline__ void nf_point_e(uint32_t ix, uint32_t iy, uint32_t iz, const float* __restrict__ el,
                                           const Nf2ffGeom& g, VA V, R out[3]) {
    // no contraction: the same bits whichever accessor supplies the values
#pragma clang fp contract(off)
    R E0 = 0, E1 = 0, E2 = 0;
    if (!(ix >= g.nx - 1 || iy >= g.ny - 1 || iz >= g.nz - 1)) {
        E0 = (V(0, 0, 0, 0) + V(0, 0, 1, 0) + V(0, 0, 1, 1) + V(0, 0, 0, 1)) / ((R)4 * (R)el[g.el_off[0] + ix]);
        E1 = (V(1, 0, 0, 0) + V(1, 0, 0, 1) + V(1, 1, 0, 1) + V(1, 1, 0, 0)) / ((R)4 * (R)el[g.el_off[1] + iy]);
        E2 = (V(2, 0, 0, 0) + V(2, 1, 0, 0) + V(2, 1, 1, 0) + V(2, 0, 1, 0)) / ((R)4 * (R)el[g.el_off[2] + iz]);
    }
    out[0] = E0;
    out[1] = E1;
    out[2] = E2;
}

template <class R, class IA>
__device__ __forceinline__ void nf_point_h(uint32_t ix, uint32_t iy, uint32_t iz, const float* __restrict__ el,
                                           const Nf2ffGeom& g, IA I, R out[3]) {
#pragma clang fp contract(off)
    const uint32_t pos[3] = {ix, iy, iz};
    const uint32_t lim[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        R h = 0;
        const R d0 = el[g.el_off[3 + n] + pos[n]];
        if (d0 > (R)0) {
            h = I(n, 0) / d0;
            if (pos[n] < lim[n] - 1) {
                const R d1 = el[g.el_off[3 + n] + pos[n] + 1];
                const R dr = d0 / (d0 + d1);
                const R up = I(n, 1) / d1;
                h = h * ((R)1 - dr) + up * dr;
            }
        }
        out[n] = h;
    }
}

template <class R, class VA, class IA>
__device__ __forceinline__ void nf_point(uint32_t ix, uint32_t iy, uint32_t iz, const float* __restrict__ el,
                                         const Nf2ffGeom& g, VA V, IA I, R out[6]) {
    nf_point_e<R>(ix, iy, iz, el, g, V, out);
    nf_point_h<R>(ix, iy, iz, el, g, I, out + 3);
}

// One thread per surface point.  Points on a captured z plane (see
// NfCapture) read the slabs the H sweep wrote; the others read the fields.
// PART: kSampleAll, the whole sample of the point; kSampleE, the E half
// of a point that reads the fields (captured points: nothing), from the
// kernel of the H sweep, which does not write E; kSampleRest, the H half
// of such a point and both halves of a captured one, after the H sweep
// wrote H and the slabs.  The three give the same rows.
constexpr int kSampleAll = 0, kSampleE = 1, kSampleRest = 2;

template <class FV, int PART = kSampleAll>
__device__ __forceinline__ void nf2ff_sample_point(const FV& volt, const FV& curr, const uint32_t* __restrict__ pts,
                                                   const float* __restrict__ el, const Nf2ffGeom& g,
                                                   typename FV::R* __restrict__ row_out, const NfCapture& cap,
                                                   uint32_t p) {
    typedef typename FV::R R;
    if (p >= g.P) return;
    const uint32_t ix = pts[3 * p], iy = pts[3 * p + 1], iz = pts[3 * p + 2];
    R o[6];
    int f = -1;
    if (cap.on && ix - cap.x0 < cap.x1 - cap.x0 && iy - cap.y0 < cap.y1 - cap.y0)
        f = iz == cap.z[0] ? 0 : (cap.nf > 1 && iz == cap.z[1] ? 1 : -1);
    const size_t P = g.P, RE = g.rowsE;
    if (f >= 0) {
        if (PART == kSampleE) return;
        const float* q = cap.slab + ((size_t)(f * cap.nx_c + (ix - cap.x0)) * cap.ny_c + (iy - cap.y0)) * kCapFloats;
        const size_t rx = (size_t)cap.ny_c * kCapFloats;  // next x row
        auto V = [&](int c, int dx, int dy, int dz) { return (R)q[dx * rx + dy * kCapFloats + 2 * c + dz]; };
        auto I = [&](int n, int d) {
            return (R)(n == 0 ? q[d * rx + 6] : (n == 1 ? q[d * kCapFloats + 7] : q[8 + d]));
        };
        nf_point<R>(ix, iy, iz, el, g, V, I, o);
    } else {
        const size_t Ny = g.ny, Nz = g.nzs;
        const size_t b = ix * Ny * Nz + iy * Nz + iz;
        const size_t sx = Ny * Nz, sy = Nz;
        const size_t step[3] = {sx, sy, 1};
        auto V = [&](int c, int dx, int dy, int dz) { return (R)volt.ldc(c, b + dx * sx + dy * sy + dz); };
        auto I = [&](int n, int d) { return (R)curr.ldc(n, b + (d ? step[n] : 0)); };
        if (PART != kSampleRest) nf_point_e<R>(ix, iy, iz, el, g, V, o);
        if (PART != kSampleE) nf_point_h<R>(ix, iy, iz, el, g, I, o + 3);
        if (PART == kSampleE) {
            row_out[p] = o[0];
            row_out[P + p] = o[1];
            row_out[2 * P + p] = o[2];
            return;
        }
        if (PART == kSampleRest) {
            row_out[RE + p] = o[3];
            row_out[RE + P + p] = o[4];
            row_out[RE + 2 * P + p] = o[5];
            return;
        }
    }
    row_out[p] = o[0];
    row_out[P + p] = o[1];
    row_out[2 * P + p] = o[2];
    row_out[RE + p] = o[3];
    row_out[RE + P + p] = o[4];
    row_out[RE + 2 * P + p] = o[5];
}

template <class FV>
__global__ void k_nf2ff_sample(FV volt, FV curr,
                               const uint32_t* __restrict__ pts, const float* __restrict__ el, Nf2ffGeom g,
                               typename FV::R* __restrict__ row_out /* buf + slot*rows */, NfCapture cap) {
    // points are in grid order; consecutive workgroups share cells, so they
    // stay on one XCD (own L2)
    const uint32_t p = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    nf2ff_sample_point(volt, curr, pts, el, g, row_out, cap, p);
}

// FP64 twiddles for k in [0,kused): tw[(h*K + k)*NC + 2f + {0: re, 1: im}],
// h = 0 (E) or 1 (H, half-step later), zero for f >= F.  The Hann weight
// uses the integer step index for both, as the reference does.
// T: float (the matrix-core fold) or double (the FP64 fold).
template <typename T>
__global__ void k_nf2ff_twiddle(T* __restrict__ tw, uint32_t K, uint32_t kused, uint32_t F, uint32_t NC,
                                uint32_t ts0, const double* __restrict__ omega, double dT, uint32_t maxTS,
                                uint32_t window, double wnorm) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * kused * NC) return;
    const uint32_t c = t % NC, k = (t / NC) % kused, h = t / (NC * kused);
    const uint32_t f = c / 2;
    T v = 0;
    if (f < F) {
        const double n = (double)ts0 + (double)k;
        double w = 1.0;
        if (window == 1 && maxTS > 1) w = 0.5 * (1.0 - cos(2.0 * M_PI * n / (double)(maxTS - 1)));
        const double s = w * wnorm * dT;
        const double ph = omega[f] * (n + (h ? 0.5 : 0.0)) * dT;
        double sn, cs;
        sincos(ph, &sn, &cs);
        v = (T)((c & 1u) ? -sn * s : cs * s);
    }
    tw[((size_t)h * K + k) * NC + c] = v;
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Fold on the matrix cores (v_mfma_f32_32x32x2f32: D[32x32] += A[32x2] B[2x32]).
// Workgroup: kFoldRows rows x 32 columns (column tile blockIdx.y); wave w
// owns rows r0 + 4m + j (m < 32, j < 4) of r0 = base + 128w as four 32-row
// tiles j, so one 16-byte load per lane, buf[k][r0 + 4m .. +3], feeds the
// four products of a k-pair.  Operand layouts of the instruction: lane l
// gives A[m = l%32][k = l/32] and B[k = l/32][n = l%32]; D[m][n] is in lane
// n + 32*((m/4)%2), accumulator register (m%4) + 4*(m/8).  The block's sums
// start at zero and are added to acc (the running DFT) at the end; every
// acc element belongs to exactly one lane.
// TwoSum of a and b: a + b = s + e exactly (round to nearest, no
// contraction).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

__global__ __launch_bounds__(256) void k_nf2ff_fold(const float* __restrict__ buf, uint32_t rows, uint32_t rowsE,
                                                    uint32_t K, uint32_t kused, const float* __restrict__ tw,
                                                    uint32_t NC, float* __restrict__ acc, float* __restrict__ cmp) {
    extern __shared__ float s_w[];  // [kpad][32]: this tile's twiddle columns, zero rows past kused
    const uint32_t base = blockIdx.x * kFoldRows;
    const uint32_t h = base >= rowsE ? 1u : 0u;
    const uint32_t c0 = blockIdx.y * kFoldCols;
    const uint32_t kp = (kused + 1u) & ~1u;
    for (uint32_t t = threadIdx.x; t < kp * kFoldCols; t += blockDim.x) {
        const uint32_t k = t / kFoldCols, c = t % kFoldCols;
        s_w[t] = k < kused ? tw[((size_t)h * K + k) * NC + c0 + c] : 0.f;
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t r0 = base + wave * 128u;
    const uint32_t col = lane & 31u, kh = lane >> 5;
    f32x16 D[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) D[j] = (f32x16)0.f;
    const float* a_ptr = buf + r0 + 4u * col;
    auto pair = [&](v4 a, float b) {
        D[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b, D[0], 0, 0, 0);
        D[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b, D[1], 0, 0, 0);
        D[2] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b, D[2], 0, 0, 0);
        D[3] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b, D[3], 0, 0, 0);
    };
    // 8 k-pairs per batch: all loads issued before the first product
    const uint32_t kfull = kused & ~1u;  // k-pairs with both steps inside [0, kused)
    uint32_t k = 0;
    for (; k + 16 <= kfull; k += 16) {
        v4 a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = ld4(a_ptr + (size_t)(k + 2 * q + kh) * rows);
        __builtin_amdgcn_sched_barrier(0);  // keep the 8 loads ahead of the products
#pragma unroll
        for (int q = 0; q < 8; ++q) pair(a[q], s_w[(k + 2 * q + kh) * kFoldCols + col]);
    }
    for (; k < kfull; k += 2) pair(ld4(a_ptr + (size_t)(k + kh) * rows), s_w[(k + kh) * kFoldCols + col]);
    if (k < kp) {  // odd kused: the second step of the last pair is zero
        const uint32_t kk = k + kh;
        pair(kk < kused ? ld4(a_ptr + (size_t)kk * rows) : (v4)0.f, s_w[kk * kFoldCols + col]);
    }
    float* const acc_t = acc + c0;  // uniform base; per-lane 32-bit element offsets
    const uint32_t o0 = (r0 + 16u * kh) * NC + col;
    if (cmp) {  // uniform: compensated form
        float* const cmp_t = cmp + c0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v[16], w[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const uint32_t o = o0 + (4u * (8u * (i / 4) + (i % 4)) + j) * NC;
                v[i] = acc_t[o];
                w[i] = cmp_t[o];
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const uint32_t o = o0 + (4u * (8u * (i / 4) + (i % 4)) + j) * NC;
                float s, e;
                two_sum(v[i], D[j][i], s, e);
                acc_t[o] = s;
                cmp_t[o] = w[i] + e;
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = acc_t[o0 + (4u * (8u * (i / 4) + (i % 4)) + j) * NC];
#pragma unroll
        for (int i = 0; i < 16; ++i) acc_t[o0 + (4u * (8u * (i / 4) + (i % 4)) + j) * NC] = v[i] + D[j][i];
    }
}

// FP64 fold of the FP64 storage form (its precision reference run): the
// same product in double on the vector ALUs, one thread per row and 32
// accumulator columns (16 frequencies) of column tile blockIdx.y, twiddles
// of the row block's half (E or H; rowsE is a multiple of 512, so a block
// of 256 rows is one or the other) in LDS.
constexpr int kFold64Rows = 256;
__global__ __launch_bounds__(kFold64Rows) void k_nf2ff_fold64(const double* __restrict__ buf, uint32_t rows,
                                                              uint32_t rowsE, uint32_t K, uint32_t kused,
                                                              const double* __restrict__ tw, uint32_t NC,
                                                              double* __restrict__ acc) {
    extern __shared__ double s_w64[];  // [kused][kFoldCols]
    const uint32_t base = blockIdx.x * kFold64Rows;
    const uint32_t h = base >= rowsE ? 1u : 0u;
    const uint32_t c0 = blockIdx.y * kFoldCols;
    for (uint32_t t = threadIdx.x; t < kused * kFoldCols; t += blockDim.x)
        s_w64[t] = tw[((size_t)h * K + t / kFoldCols) * NC + c0 + t % kFoldCols];
    __syncthreads();
    const uint32_t row = base + threadIdx.x;
    if (row >= rows) return;
    double d[kFoldCols];
#pragma unroll
    for (int c = 0; c < kFoldCols; ++c) d[c] = 0.0;
    for (uint32_t k = 0; k < kused; ++k) {
        const double b = buf[(size_t)k * rows + row];
#pragma unroll
        for (int c = 0; c < kFoldCols; ++c) d[c] = __builtin_fma(b, s_w64[k * kFoldCols + c], d[c]);
    }
    double* a = acc + (size_t)row * NC + c0;
#pragma unroll
    for (int c = 0; c < kFoldCols; ++c) a[c] += d[c];
}

// Far field for one frequency bin (NF2FF far-field WGSL,
// src/webgpu-engine.mjs:802-986): thread per (theta, phi); surface points
// are staged through LDS; sums are carried in FP64.
struct FarParams {
    uint32_t P, nT, nP, pad;
    float k, Z0, cx, cy, cz, fac_re, fac_im, pad2;
    // mirror (mdir >= 0): points P .. 2P-1 are the surface reflected in the
    // plane x[mdir] = mpos, fields times ef / hf, normal sign times msign
    // for faces perpendicular to the plane
    int32_t mdir;
    float mpos, msign;
    float ef[3], hf[3];
};

constexpr int kFarTile = 256;

// acc_f = acc + 2f: the frequency's (re, im) column pair, NC floats per row.
// Grid: x over angles (one thread each), y over chunks of `chunk` surface
// points (a multiple of kFarTile); each workgroup writes the 8 FP64 radiation
// integrals (N_theta, N_phi, L_theta, L_phi; re, im) of its angles over its
// points to partial[(chunk index * nA + angle) * 8], and k_farfield_final
// adds the chunks in order, so the result does not depend on scheduling.
// Splitting the points keeps the launch wide when there are few angles
// (an RCS cut of 64 angles is one workgroup of angles).
template <typename A>
__global__ __launch_bounds__(256) void k_farfield(const A* __restrict__ acc_f, const A* __restrict__ cmp_f,
                                                  uint32_t NC, uint32_t rowsE,
                                                  const float* __restrict__ meta, const uint32_t* __restrict__ perm,
                                                  const float* __restrict__ theta,
                                                  const float* __restrict__ phi, FarParams q, uint32_t chunk,
                                                  double* __restrict__ partial) {
    __shared__ float s_pt[kFarTile][18];  // pos xyz, nd, sign, area, E re/im x3, H re/im x3
    const uint32_t nA = q.nT * q.nP;
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = a < nA;
    float sT = 0.f, cT = 1.f, sP = 0.f, cP = 1.f;
    if (active) {
        sincosf(theta[a / q.nP], &sT, &cT);
        sincosf(phi[a % q.nP], &sP, &cP);
    }
    const uint32_t Pv = q.mdir >= 0 ? 2 * q.P : q.P;  // points of both passes
    const uint32_t p0 = blockIdx.y * chunk, p1 = min(Pv, p0 + chunk);
    double Ntr = 0, Nti = 0, Npr = 0, Npi = 0, Ltr = 0, Lti = 0, Lpr = 0, Lpi = 0;
    for (uint32_t base = p0; base < p1; base += kFarTile) {
        __syncthreads();
        const uint32_t pv = base + threadIdx.x;
        if (threadIdx.x < kFarTile && pv < p1) {
            const bool img = pv >= q.P;
            const uint32_t p = img ? pv - q.P : pv;
            // meta in internal point order, or (perm) in the caller's order
            const float* m = meta + 8 * (size_t)(perm ? perm[p] : p);
            float* s = s_pt[threadIdx.x];
            float pos[3] = {m[0], m[1], m[2]};
            s[3] = m[3];
            s[4] = m[4];
            s[5] = m[5];
            if (img) {
                pos[q.mdir] = 2.0f * q.mpos - pos[q.mdir];
                if ((int)s[3] == q.mdir) s[4] *= q.msign;
            }
            s[0] = pos[0] - q.cx;
            s[1] = pos[1] - q.cy;
            s[2] = pos[2] - q.cz;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const size_t oe = ((size_t)c * q.P + p) * NC, oh = ((size_t)rowsE + (size_t)c * q.P + p) * NC;
                A e0 = acc_f[oe], e1 = acc_f[oe + 1], h0 = acc_f[oh], h1 = acc_f[oh + 1];
                if (cmp_f) {  // compensated accumulators: value = acc + cmp
                    e0 += cmp_f[oe]; e1 += cmp_f[oe + 1];
                    h0 += cmp_f[oh]; h1 += cmp_f[oh + 1];
                }
                const float fe = img ? q.ef[c] : 1.0f, fh = img ? q.hf[c] : 1.0f;
                s[6 + 2 * c] = (float)e0 * fe;
                s[7 + 2 * c] = (float)e1 * fe;
                s[12 + 2 * c] = (float)h0 * fh;
                s[13 + 2 * c] = (float)h1 * fh;
            }
        }
        __syncthreads();
        if (!active) continue;
        const uint32_t cnt = min((uint32_t)kFarTile, p1 - base);
        for (uint32_t j = 0; j < cnt; ++j) {
            const float* s = s_pt[j];
            const int nd = (int)s[3];
            const float sg = s[4], area = s[5];
            float Jx[2] = {0.f, 0.f}, Jy[2] = {0.f, 0.f}, Jz[2] = {0.f, 0.f};
            float Mx[2] = {0.f, 0.f}, My[2] = {0.f, 0.f}, Mz[2] = {0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const float Ex = s[6 + r], Ey = s[8 + r], Ez = s[10 + r];
                const float Hx = s[12 + r], Hy = s[14 + r], Hz = s[16 + r];
                if (nd == 0) {
                    Jy[r] = -sg * Hz; Jz[r] = sg * Hy; My[r] = sg * Ez; Mz[r] = -sg * Ey;
                } else if (nd == 1) {
                    Jx[r] = sg * Hz; Jz[r] = -sg * Hx; Mx[r] = -sg * Ez; Mz[r] = sg * Ex;
                } else {
                    Jx[r] = -sg * Hy; Jy[r] = sg * Hx; Mx[r] = sg * Ey; My[r] = -sg * Ex;
                }
            }
            const float rc = s[0] * sT * cP + s[1] * sT * sP + s[2] * cT;
            float sn, cs;
            sincosf(q.k * rc, &sn, &cs);
            const float aer = area * cs, aei = area * sn;
            float Jt[2], Jp[2], Mt[2], Mp[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                Jt[r] = Jx[r] * cT * cP + Jy[r] * cT * sP - Jz[r] * sT;
                Jp[r] = Jy[r] * cP - Jx[r] * sP;
                Mt[r] = Mx[r] * cT * cP + My[r] * cT * sP - Mz[r] * sT;
                Mp[r] = My[r] * cP - Mx[r] * sP;
            }
            Ntr += aer * Jt[0] - aei * Jt[1]; Nti += aer * Jt[1] + aei * Jt[0];
            Npr += aer * Jp[0] - aei * Jp[1]; Npi += aer * Jp[1] + aei * Jp[0];
            Ltr += aer * Mt[0] - aei * Mt[1]; Lti += aer * Mt[1] + aei * Mt[0];
            Lpr += aer * Mp[0] - aei * Mp[1]; Lpi += aer * Mp[1] + aei * Mp[0];
        }
    }
    if (!active) return;
    double* o = partial + ((size_t)blockIdx.y * nA + a) * 8;
    o[0] = Ntr; o[1] = Nti; o[2] = Npr; o[3] = Npi;
    o[4] = Ltr; o[5] = Lti; o[6] = Lpr; o[7] = Lpi;
}

// Sum of the chunk partials per angle, then E_theta, E_phi and the radiation
// intensity (5 floats per angle).
template <typename O>
__global__ __launch_bounds__(256) void k_farfield_final(const double* __restrict__ partial, uint32_t nchunks,
                                                        FarParams q, O* __restrict__ out) {
    const uint32_t nA = q.nT * q.nP;
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nA) return;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t c = 0; c < nchunks; ++c) {
        const double* p = partial + ((size_t)c * nA + a) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] += p[i];
    }
    const double Ntr = v[0], Nti = v[1], Npr = v[2], Npi = v[3], Ltr = v[4], Lti = v[5], Lpr = v[6], Lpi = v[7];
    const double fr = q.fac_re, fi = q.fac_im, Z0 = q.Z0;
    const double a1r = Lpr + Z0 * Ntr, a1i = Lpi + Z0 * Nti;
    const double Etr = -(fr * a1r - fi * a1i), Eti = -(fr * a1i + fi * a1r);
    const double a2r = Ltr - Z0 * Npr, a2i = Lti - Z0 * Npi;
    const double Epr = fr * a2r - fi * a2i, Epi = fr * a2i + fi * a2r;
    O* o = out + 5 * (size_t)a;
    o[0] = (O)((Etr * Etr + Eti * Eti + Epr * Epr + Epi * Epi) / (2.0 * Z0));
    o[1] = (O)Etr;
    o[2] = (O)Eti;
    o[3] = (O)Epr;
    o[4] = (O)Epi;
}

}  // namespace efs
