// nf2ff.hpp -- NF2FF surface sampling, K-step DFT fold and far field.
//
// The reference accumulates one frequency every step with an FP32 phase
// (NF2FF accumulate WGSL, src/webgpu-engine.mjs:637-800).  Here the
// cell-interpolated surface E/H of K consecutive steps are written to a
// buffer B[K][6P] and folded into F frequency bins at once:
//     acc[f][re|im][row] += sum_k B[k][row] * W[k][f]
// a complex [6P x K] x [K x F] product whose twiddles
//     W[k][f] = exp(-j*omega_f*t_k) * w(n_k) * windowNorm * dT
// (t_k = n_k*dT for E rows, (n_k+1/2)*dT for H rows) are generated in FP64.
// Row r = c*P + p for E component c, (3+c)*P + p for H component c.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efs {

struct Nf2ffGeom {
    uint32_t nx, ny, nz, P;
    uint32_t el_off[6];  // offsets of primal x,y,z / dual x,y,z edge lengths
    uint64_t S;          // component stride of the field arrays
};

// Cell interpolation of E (mean of 4 edges / primal length; 0 on the last
// grid line) and H (lerp of 2 edges / dual length along its own axis),
// as in the NF2FF accumulate WGSL.
__global__ void k_nf2ff_sample(const float* __restrict__ volt, const float* __restrict__ curr,
                               const uint32_t* __restrict__ pts, const float* __restrict__ el, Nf2ffGeom g,
                               float* __restrict__ row_out /* buf + slot*6P */) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const uint32_t ix = pts[3 * p], iy = pts[3 * p + 1], iz = pts[3 * p + 2];
    const size_t Ny = g.ny, Nz = g.nz;
    const size_t N = g.S;
    const size_t b = ix * Ny * Nz + iy * Nz + iz;
    const size_t sx = Ny * Nz, sy = Nz;
    float E0 = 0.f, E1 = 0.f, E2 = 0.f;
    if (!(ix >= g.nx - 1 || iy >= g.ny - 1 || iz >= g.nz - 1)) {
        E0 = (volt[b] + volt[b + sy] + volt[b + sy + 1] + volt[b + 1]) / (4.0f * el[g.el_off[0] + ix]);
        E1 = (volt[N + b] + volt[N + b + 1] + volt[N + b + sx + 1] + volt[N + b + sx]) /
             (4.0f * el[g.el_off[1] + iy]);
        E2 = (volt[2 * N + b] + volt[2 * N + b + sx] + volt[2 * N + b + sx + sy] + volt[2 * N + b + sy]) /
             (4.0f * el[g.el_off[2] + iz]);
    }
    const uint32_t pos[3] = {ix, iy, iz};
    const uint32_t lim[3] = {g.nx, g.ny, g.nz};
    const size_t step[3] = {sx, sy, 1};
    float H[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        float h = 0.f;
        const float d0 = el[g.el_off[3 + n] + pos[n]];
        if (d0 > 0.f) {
            h = curr[n * N + b] / d0;
            if (pos[n] < lim[n] - 1) {
                const float d1 = el[g.el_off[3 + n] + pos[n] + 1];
                const float dr = d0 / (d0 + d1);
                const float up = curr[n * N + b + step[n]] / d1;
                h = h * (1.0f - dr) + up * dr;
            }
        }
        H[n] = h;
    }
    const size_t P = g.P;
    row_out[p] = E0;
    row_out[P + p] = E1;
    row_out[2 * P + p] = E2;
    row_out[3 * P + p] = H[0];
    row_out[4 * P + p] = H[1];
    row_out[5 * P + p] = H[2];
}

// FP64 twiddles for k in [0,kused): tw[(h*K + k)*F + f] = {re, im}, h = 0 (E)
// or 1 (H, half-step later).  The Hann weight uses the integer step index
// for both, as the reference does.
__global__ void k_nf2ff_twiddle(float2* __restrict__ tw, uint32_t K, uint32_t kused, uint32_t F, uint32_t ts0,
                                const double* __restrict__ omega, double dT, uint32_t maxTS,
                                uint32_t window, double wnorm) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * kused * F) return;
    const uint32_t f = t % F, k = (t / F) % kused, h = t / (F * kused);
    const double n = (double)ts0 + (double)k;
    double w = 1.0;
    if (window == 1 && maxTS > 1) w = 0.5 * (1.0 - cos(2.0 * M_PI * n / (double)(maxTS - 1)));
    const double s = w * wnorm * dT;
    const double ph = omega[f] * (n + (h ? 0.5 : 0.0)) * dT;
    double sn, cs;
    sincos(ph, &sn, &cs);
    tw[((size_t)h * K + k) * F + f] = make_float2((float)(cs * s), (float)(-sn * s));
}

// Fold: one thread per row, FB frequencies per workgroup (grid.y over F).
template <int FB>
__global__ __launch_bounds__(256) void k_nf2ff_fold(const float* __restrict__ buf, uint32_t rows,
                                                    uint32_t rowsE, uint32_t K, uint32_t kused,
                                                    const float2* __restrict__ tw, uint32_t F,
                                                    float* __restrict__ acc) {
    extern __shared__ float2 s_tw[];  // [2][kused][FB]
    const uint32_t f0 = blockIdx.y * FB;
    for (uint32_t t = threadIdx.x; t < 2 * kused * FB; t += blockDim.x) {
        const uint32_t f = t % FB, k = (t / FB) % kused, h = t / (FB * kused);
        s_tw[t] = (f0 + f < F) ? tw[((size_t)h * K + k) * F + f0 + f] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float2* t = s_tw + (r >= rowsE ? kused * FB : 0);
    float re[FB], im[FB];
#pragma unroll
    for (int f = 0; f < FB; ++f) re[f] = im[f] = 0.f;
    for (uint32_t k = 0; k < kused; ++k) {
        const float v = buf[(size_t)k * rows + r];
#pragma unroll
        for (int f = 0; f < FB; ++f) {
            const float2 w = t[k * FB + f];
            re[f] += v * w.x;
            im[f] += v * w.y;
        }
    }
#pragma unroll
    for (int f = 0; f < FB; ++f) {
        if (f0 + f < F) {
            acc[((size_t)(f0 + f) * 2) * rows + r] += re[f];
            acc[((size_t)(f0 + f) * 2 + 1) * rows + r] += im[f];
        }
    }
}

// Far field for one frequency bin (NF2FF far-field WGSL,
// src/webgpu-engine.mjs:802-986): thread per (theta, phi); surface points
// are staged through LDS; sums are carried in FP64.
struct FarParams {
    uint32_t P, nT, nP, pad;
    float k, Z0, cx, cy, cz, fac_re, fac_im, pad2;
};

constexpr int kFarTile = 256;

__global__ __launch_bounds__(256) void k_farfield(const float* __restrict__ acc_f /* [2][6P] */,
                                                  const float* __restrict__ meta, const float* __restrict__ theta,
                                                  const float* __restrict__ phi, FarParams q,
                                                  float* __restrict__ out) {
    __shared__ float s_pt[kFarTile][18];  // pos xyz, nd, sign, area, E re/im x3, H re/im x3
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = a < q.nT * q.nP;
    float sT = 0.f, cT = 1.f, sP = 0.f, cP = 1.f;
    if (active) {
        sincosf(theta[a / q.nP], &sT, &cT);
        sincosf(phi[a % q.nP], &sP, &cP);
    }
    double Ntr = 0, Nti = 0, Npr = 0, Npi = 0, Ltr = 0, Lti = 0, Lpr = 0, Lpi = 0;
    const size_t rows = 6 * (size_t)q.P;
    for (uint32_t base = 0; base < q.P; base += kFarTile) {
        __syncthreads();
        const uint32_t p = base + threadIdx.x;
        if (threadIdx.x < kFarTile && p < q.P) {
            float* s = s_pt[threadIdx.x];
            s[0] = meta[8 * (size_t)p] - q.cx;
            s[1] = meta[8 * (size_t)p + 1] - q.cy;
            s[2] = meta[8 * (size_t)p + 2] - q.cz;
            s[3] = meta[8 * (size_t)p + 3];
            s[4] = meta[8 * (size_t)p + 4];
            s[5] = meta[8 * (size_t)p + 5];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                s[6 + 2 * c] = acc_f[(size_t)c * q.P + p];
                s[7 + 2 * c] = acc_f[rows + (size_t)c * q.P + p];
                s[12 + 2 * c] = acc_f[(size_t)(3 + c) * q.P + p];
                s[13 + 2 * c] = acc_f[rows + (size_t)(3 + c) * q.P + p];
            }
        }
        __syncthreads();
        if (!active) continue;
        const uint32_t cnt = min((uint32_t)kFarTile, q.P - base);
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
    const double fr = q.fac_re, fi = q.fac_im, Z0 = q.Z0;
    const double a1r = Lpr + Z0 * Ntr, a1i = Lpi + Z0 * Nti;
    const double Etr = -(fr * a1r - fi * a1i), Eti = -(fr * a1i + fi * a1r);
    const double a2r = Ltr - Z0 * Npr, a2i = Lti - Z0 * Npi;
    const double Epr = fr * a2r - fi * a2i, Epi = fr * a2i + fi * a2r;
    float* o = out + 5 * (size_t)a;
    o[0] = (float)((Etr * Etr + Eti * Eti + Epr * Epr + Epi * Epi) / (2.0 * Z0));
    o[1] = (float)Etr;
    o[2] = (float)Eti;
    o[3] = (float)Epr;
    o[4] = (float)Epi;
}

}  // namespace efs
