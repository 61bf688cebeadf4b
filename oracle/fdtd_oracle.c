/*
 * fdtd_oracle.c -- CPU restatement of the reference FDTD time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * engine in efs_amd/csrc and the CPU baseline timed by bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the product path never links or calls it.
 *
 * What it restates (reference files cited per function):
 *   - CPUFDTDEngine in src/webgpu-fdtd.mjs:13-1202 (core Yee updates, UPML
 *     pre/post, excitation, ADE, TFSF, lumped RLC, Mur, step order), and
 *   - the WebGPUEngine hooks the CPU class lacks, from src/webgpu-engine.mjs:
 *     NF2FF accumulation (637-800, 2729-2838), far-field (802-986,
 *     2871-2967), energy (3039-3117), buffered probe gather (2615-2687).
 *
 * Two arithmetic variants are built from this one source:
 *   ORC_F64=0  "f32": fields and coefficients stored as float, every
 *              expression evaluated in double and rounded on store.  This is
 *              exactly what the reference's JavaScript does (Float32Array
 *              storage, IEEE double arithmetic), so results are bit-identical
 *              to CPUFDTDEngine when compiled with -ffp-contract=off.
 *   ORC_F64=1  "f64": the same algorithm with double storage -- the FP64
 *              oracle that FP32/FP16 engine results are measured against.
 *
 * Cells are independent inside each core sweep (E reads only H, H reads
 * only E), so the x loop is split across OpenMP threads without changing a
 * single bit of the result; sparse extension loops stay sequential because
 * their index lists may repeat a cell.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>

/* Loops over fewer cells than this run on one thread: on small grids a
 * thread team per loop costs more than the loop (each cell's result is the
 * same either way). */
#ifndef OMP_MIN_CELLS
#define OMP_MIN_CELLS 65536
#endif
#endif

#if ORC_F64
typedef double real;
#define PFX(name) orc64_##name
#else
typedef float real;
#define PFX(name) orc32_##name
#endif

#define ST(v) ((real)(v))

typedef struct {
    uint32_t start[3], num[3];
    size_t total; /* 3*nx*ny*nz */
    real *vv, *vvfo, *vvfn, *ii, *iifo, *iifn;
    real *vflux, *cflux;
} pml_region;

typedef struct {
    uint32_t dir, ncells, lorentz;
    uint32_t *pos;
    real *v_int, *v_ext, *v_lor, *i_int, *i_ext, *i_lor;
    real *volt_ade, *volt_lor, *curr_ade, *curr_lor;
} ade_dir;

typedef struct {
    uint32_t field_idx, delay_int;
    double delay_frac, amp; /* plain JS numbers in the reference */
} tfsf_point;

typedef struct {
    uint32_t field_idx, direction, type_flag;
    double i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2; /* plain JS numbers */
} rlc_elem;

typedef struct {
    uint32_t nx, ny, nz;
    size_t N;
    real *vv, *vi, *ii, *iv;
    real *volt, *curr;
    uint32_t numTS;

    /* excitation */
    uint32_t exc_n, exc_len, exc_period;
    real *exc_signal, *exc_amp;
    uint32_t *exc_delay, *exc_dir, *exc_pos;

    /* UPML */
    uint32_t npml;
    pml_region *pml;

    /* Mur */
    uint32_t mur_n;
    real *mur_coeff, *mur_saved;
    uint32_t *mur_normal, *mur_shifted;

    /* ADE */
    uint32_t nade;
    ade_dir *ade;
    int ade_apply; /* 0 = reference (no write-back); 1 = V -= volt_ADE */

    /* TFSF */
    int has_tfsf;
    real *tfsf_signal;
    uint32_t tfsf_len, tfsf_period, tfsf_nv, tfsf_ni;
    tfsf_point *tfsf_v, *tfsf_i;

    /* RLC */
    uint32_t rlc_n;
    rlc_elem *rlc;
    real *rlc_vdn, *rlc_jn, *rlc_vil;

    /* NF2FF accumulation (multi-frequency generalisation of one omega) */
    uint32_t nf_P, nf_F, nf_maxTS, nf_window;
    uint32_t *nf_points;
    double *nf_omega;
    double nf_dT, nf_wnorm;
    double *nf_el[6]; /* primal x,y,z then dual x,y,z */
    uint32_t nf_el_len[6];
    double *nf_accE, *nf_accH; /* [F][6P] */

    /* buffered probe gather */
    uint32_t g_n, g_max, g_interval, g_count;
    uint32_t *g_idx;
    real *g_ring;

    /* steady-state detection (CPU class, src/webgpu-fdtd.mjs:1037-1095) */
    uint32_t ss_n, ss_period, ss_current;
    int ss_recording;
    double ss_threshold;
    uint32_t *ss_idx;
    real *ss_e1, *ss_e2;

    /* periodic boundaries (GPU engine only, src/webgpu-engine.mjs:475-584,
     * 1976-2155): per axis the Bloch phase as the f32 cos/sin the engine
     * uploads, and 4 face-sized imaginary-part buffers when sin != 0 */
    uint32_t pbc_n;
    uint32_t pbc_axis[3];
    float pbc_cos[3], pbc_sin[3];
    real *pbc_imag[3];
} orc;

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    return p;
}

static real *dup_real(const double *src, size_t n) {
    real *d = (real *)xcalloc(n, sizeof(real));
    if (src)
        for (size_t i = 0; i < n; ++i) d[i] = ST(src[i]);
    return d;
}

static uint32_t *dup_u32(const uint32_t *src, size_t n) {
    uint32_t *d = (uint32_t *)xcalloc(n, sizeof(uint32_t));
    if (src) memcpy(d, src, n * sizeof(uint32_t));
    return d;
}

static inline size_t IDX(const orc *e, int n, size_t x, size_t y, size_t z) {
    return (size_t)n * e->N + x * e->ny * e->nz + y * e->nz + z;
}

/* ------------------------------------------------------------------ */
/* lifetime                                                            */
/* ------------------------------------------------------------------ */

orc *PFX(create)(uint32_t nx, uint32_t ny, uint32_t nz, const double *vv, const double *vi,
                 const double *ii, const double *iv) {
    orc *e = (orc *)xcalloc(1, sizeof(orc));
    e->nx = nx; e->ny = ny; e->nz = nz;
    e->N = (size_t)nx * ny * nz;
    e->vv = dup_real(vv, 3 * e->N);
    e->vi = dup_real(vi, 3 * e->N);
    e->ii = dup_real(ii, 3 * e->N);
    e->iv = dup_real(iv, 3 * e->N);
    e->volt = (real *)xcalloc(3 * e->N, sizeof(real));
    e->curr = (real *)xcalloc(3 * e->N, sizeof(real));
    return e;
}

static void free_pml(orc *e) {
    for (uint32_t r = 0; r < e->npml; ++r) {
        pml_region *p = &e->pml[r];
        free(p->vv); free(p->vvfo); free(p->vvfn); free(p->ii); free(p->iifo); free(p->iifn);
        free(p->vflux); free(p->cflux);
    }
    free(e->pml); e->pml = NULL; e->npml = 0;
}

static void free_ade(orc *e) {
    for (uint32_t k = 0; k < e->nade; ++k) {
        ade_dir *d = &e->ade[k];
        free(d->pos); free(d->v_int); free(d->v_ext); free(d->v_lor);
        free(d->i_int); free(d->i_ext); free(d->i_lor);
        free(d->volt_ade); free(d->volt_lor); free(d->curr_ade); free(d->curr_lor);
    }
    free(e->ade); e->ade = NULL; e->nade = 0;
}

void PFX(destroy)(orc *e) {
    if (!e) return;
    free(e->vv); free(e->vi); free(e->ii); free(e->iv); free(e->volt); free(e->curr);
    free(e->exc_signal); free(e->exc_amp); free(e->exc_delay); free(e->exc_dir); free(e->exc_pos);
    free_pml(e);
    free(e->mur_coeff); free(e->mur_saved); free(e->mur_normal); free(e->mur_shifted);
    free_ade(e);
    free(e->tfsf_signal); free(e->tfsf_v); free(e->tfsf_i);
    free(e->rlc); free(e->rlc_vdn); free(e->rlc_jn); free(e->rlc_vil);
    free(e->nf_points); free(e->nf_omega);
    for (int k = 0; k < 6; ++k) free(e->nf_el[k]);
    free(e->nf_accE); free(e->nf_accH);
    free(e->g_idx); free(e->g_ring);
    free(e->ss_idx); free(e->ss_e1); free(e->ss_e2);
    for (uint32_t k = 0; k < e->pbc_n; ++k) free(e->pbc_imag[k]);
    free(e);
}

void PFX(set_threads)(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int PFX(max_threads)(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

uint32_t PFX(num_ts)(const orc *e) { return e->numTS; }
void PFX(set_num_ts)(orc *e, uint32_t ts) { e->numTS = ts; }

void PFX(get_fields)(const orc *e, double *volt, double *curr) {
    for (size_t i = 0; i < 3 * e->N; ++i) {
        if (volt) volt[i] = e->volt[i];
        if (curr) curr[i] = e->curr[i];
    }
}

void PFX(upload_fields)(orc *e, const double *volt, const double *curr) {
    for (size_t i = 0; i < 3 * e->N; ++i) {
        if (volt) e->volt[i] = ST(volt[i]);
        if (curr) e->curr[i] = ST(curr[i]);
    }
}

/* ------------------------------------------------------------------ */
/* core Yee updates  (src/webgpu-fdtd.mjs:64-106, 113-155)              */
/* ------------------------------------------------------------------ */

/* V = vv*V ; V += vi*(curl H) with the backward neighbour clamped to self at
 * index 0 (shift = 0).  Two separate rounding points, as the JS does. */
void PFX(update_voltages)(orc *e) {
    const long Nx = e->nx, Ny = e->ny, Nz = e->nz;
    real *volt = e->volt;
    const real *curr = e->curr, *vv = e->vv, *vi = e->vi;
#pragma omp parallel for schedule(static) if (Nx * Ny * Nz >= OMP_MIN_CELLS)
    for (long x = 0; x < Nx; ++x) {
        const long sX = x > 0 ? 1 : 0;
        for (long y = 0; y < Ny; ++y) {
            const long sY = y > 0 ? 1 : 0;
            for (long z = 0; z < Nz; ++z) {
                const long sZ = z > 0 ? 1 : 0;
                size_t i;
                i = IDX(e, 0, x, y, z);
                volt[i] = ST((double)volt[i] * (double)vv[i]);
                volt[i] = ST((double)volt[i] + (double)vi[i] *
                             ((double)curr[IDX(e, 2, x, y, z)] - (double)curr[IDX(e, 2, x, y - sY, z)] -
                              (double)curr[IDX(e, 1, x, y, z)] + (double)curr[IDX(e, 1, x, y, z - sZ)]));
                i = IDX(e, 1, x, y, z);
                volt[i] = ST((double)volt[i] * (double)vv[i]);
                volt[i] = ST((double)volt[i] + (double)vi[i] *
                             ((double)curr[IDX(e, 0, x, y, z)] - (double)curr[IDX(e, 0, x, y, z - sZ)] -
                              (double)curr[IDX(e, 2, x, y, z)] + (double)curr[IDX(e, 2, x - sX, y, z)]));
                i = IDX(e, 2, x, y, z);
                volt[i] = ST((double)volt[i] * (double)vv[i]);
                volt[i] = ST((double)volt[i] + (double)vi[i] *
                             ((double)curr[IDX(e, 1, x, y, z)] - (double)curr[IDX(e, 1, x - sX, y, z)] -
                              (double)curr[IDX(e, 0, x, y, z)] + (double)curr[IDX(e, 0, x, y - sY, z)]));
            }
        }
    }
}

/* H update on [0,Nx-1) x [0,Ny-1) x [0,Nz-1); the last planes are untouched. */
void PFX(update_currents)(orc *e) {
    const long Nx = e->nx, Ny = e->ny, Nz = e->nz;
    real *curr = e->curr;
    const real *volt = e->volt, *ii = e->ii, *iv = e->iv;
#pragma omp parallel for schedule(static) if (Nx * Ny * Nz >= OMP_MIN_CELLS)
    for (long x = 0; x < Nx - 1; ++x) {
        for (long y = 0; y < Ny - 1; ++y) {
            for (long z = 0; z < Nz - 1; ++z) {
                size_t i;
                i = IDX(e, 0, x, y, z);
                curr[i] = ST((double)curr[i] * (double)ii[i]);
                curr[i] = ST((double)curr[i] + (double)iv[i] *
                             ((double)volt[IDX(e, 2, x, y, z)] - (double)volt[IDX(e, 2, x, y + 1, z)] -
                              (double)volt[IDX(e, 1, x, y, z)] + (double)volt[IDX(e, 1, x, y, z + 1)]));
                i = IDX(e, 1, x, y, z);
                curr[i] = ST((double)curr[i] * (double)ii[i]);
                curr[i] = ST((double)curr[i] + (double)iv[i] *
                             ((double)volt[IDX(e, 0, x, y, z)] - (double)volt[IDX(e, 0, x, y, z + 1)] -
                              (double)volt[IDX(e, 2, x, y, z)] + (double)volt[IDX(e, 2, x + 1, y, z)]));
                i = IDX(e, 2, x, y, z);
                curr[i] = ST((double)curr[i] * (double)ii[i]);
                curr[i] = ST((double)curr[i] + (double)iv[i] *
                             ((double)volt[IDX(e, 1, x, y, z)] - (double)volt[IDX(e, 1, x + 1, y, z)] -
                              (double)volt[IDX(e, 0, x, y, z)] + (double)volt[IDX(e, 0, x, y + 1, z)]));
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* excitation  (src/webgpu-fdtd.mjs:161-183)                           */
/* ------------------------------------------------------------------ */

void PFX(cfg_excitation)(orc *e, const double *signal, uint32_t len, const uint32_t *delay,
                         const double *amp, const uint32_t *dir, const uint32_t *pos, uint32_t n,
                         uint32_t period) {
    free(e->exc_signal); free(e->exc_amp); free(e->exc_delay); free(e->exc_dir); free(e->exc_pos);
    e->exc_signal = dup_real(signal, len);
    e->exc_len = len;
    e->exc_amp = dup_real(amp, n);
    e->exc_delay = dup_u32(delay, n);
    e->exc_dir = dup_u32(dir, n);
    e->exc_pos = dup_u32(pos, n);
    e->exc_n = n;
    e->exc_period = period;
}

/* openEMS semantics (engine_ext_excitation Apply2Voltages): a position past
 * the signal end is zeroed and reads signal[0]; negative positions clamp to
 * 0; period 0 means "no wrap" (modulo numTS+1). */
void PFX(apply_excitation)(orc *e) {
    if (!e->exc_n || !e->exc_len) return;
    const long numTS = e->numTS;
    const long p = e->exc_period > 0 ? (long)e->exc_period : numTS + 1;
    for (uint32_t n = 0; n < e->exc_n; ++n) {
        long excPos = numTS - (long)e->exc_delay[n];
        if (excPos < 0) excPos = 0;
        excPos %= p;
        if (excPos >= (long)e->exc_len) excPos = 0;
        const size_t g = (size_t)e->exc_dir[n] * e->N + e->exc_pos[n];
        e->volt[g] = ST((double)e->volt[g] + (double)e->exc_amp[n] * (double)e->exc_signal[excPos]);
    }
}

/* ------------------------------------------------------------------ */
/* UPML  (src/webgpu-fdtd.mjs:199-358)                                  */
/* ------------------------------------------------------------------ */

void PFX(clear_pml)(orc *e) { free_pml(e); }

void PFX(add_pml_region)(orc *e, const uint32_t start[3], const uint32_t num[3], const double *vv,
                         const double *vvfo, const double *vvfn, const double *ii,
                         const double *iifo, const double *iifn) {
    e->pml = (pml_region *)realloc(e->pml, (e->npml + 1) * sizeof(pml_region));
    pml_region *p = &e->pml[e->npml++];
    memset(p, 0, sizeof(*p));
    for (int k = 0; k < 3; ++k) { p->start[k] = start[k]; p->num[k] = num[k]; }
    p->total = 3 * (size_t)num[0] * num[1] * num[2];
    p->vv = dup_real(vv, p->total); p->vvfo = dup_real(vvfo, p->total); p->vvfn = dup_real(vvfn, p->total);
    p->ii = dup_real(ii, p->total); p->iifo = dup_real(iifo, p->total); p->iifn = dup_real(iifn, p->total);
    p->vflux = (real *)xcalloc(p->total, sizeof(real));
    p->cflux = (real *)xcalloc(p->total, sizeof(real));
}

/* mode 0: pre-V, 1: post-V, 2: pre-I, 3: post-I (same numbering as the
 * reference's update_pml WGSL, src/webgpu-engine.mjs:185-201). */
void PFX(pml_step)(orc *e, int mode) {
    for (uint32_t r = 0; r < e->npml; ++r) {
        pml_region *R = &e->pml[r];
        const long nx = R->num[0], ny = R->num[1], nz = R->num[2];
        const size_t cs = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static) if (cs >= OMP_MIN_CELLS)
        for (long lx = 0; lx < nx; ++lx) {
            for (long ly = 0; ly < ny; ++ly) {
                for (long lz = 0; lz < nz; ++lz) {
                    for (int n = 0; n < 3; ++n) {
                        const size_t p = n * cs + lx * ny * nz + ly * nz + lz;
                        const size_t g = IDX(e, n, lx + R->start[0], ly + R->start[1], lz + R->start[2]);
                        if (mode == 0) {
                            const double f = (double)R->vv[p] * e->volt[g] - (double)R->vvfo[p] * R->vflux[p];
                            e->volt[g] = R->vflux[p];
                            R->vflux[p] = ST(f);
                        } else if (mode == 1) {
                            const real f = R->vflux[p];
                            R->vflux[p] = e->volt[g];
                            e->volt[g] = ST((double)f + (double)R->vvfn[p] * R->vflux[p]);
                        } else if (mode == 2) {
                            const double f = (double)R->ii[p] * e->curr[g] - (double)R->iifo[p] * R->cflux[p];
                            e->curr[g] = R->cflux[p];
                            R->cflux[p] = ST(f);
                        } else {
                            const real f = R->cflux[p];
                            R->cflux[p] = e->curr[g];
                            e->curr[g] = ST((double)f + (double)R->iifn[p] * R->cflux[p]);
                        }
                    }
                }
            }
        }
    }
}

void PFX(get_pml_flux)(const orc *e, uint32_t r, double *vflux, double *cflux) {
    if (r >= e->npml) return;
    const pml_region *R = &e->pml[r];
    for (size_t i = 0; i < R->total; ++i) {
        if (vflux) vflux[i] = R->vflux[i];
        if (cflux) cflux[i] = R->cflux[i];
    }
}

void PFX(set_pml_flux)(orc *e, uint32_t r, const double *vflux, const double *cflux) {
    if (r >= e->npml) return;
    pml_region *R = &e->pml[r];
    for (size_t i = 0; i < R->total; ++i) {
        if (vflux) R->vflux[i] = ST(vflux[i]);
        if (cflux) R->cflux[i] = ST(cflux[i]);
    }
}

/* ------------------------------------------------------------------ */
/* Lorentz / Drude ADE  (src/webgpu-fdtd.mjs:372-454)                   */
/* ------------------------------------------------------------------ */

void PFX(clear_ade)(orc *e) { free_ade(e); }
void PFX(set_ade_apply)(orc *e, int apply) { e->ade_apply = apply; }

/* One (order, direction) entry; missing coefficient arrays default to 0 as
 * in configureLorentz. */
void PFX(add_ade_dir)(orc *e, uint32_t lorentz, uint32_t dir, uint32_t ncells, const uint32_t *pos,
                      const double *v_int, const double *v_ext, const double *v_lor,
                      const double *i_int, const double *i_ext, const double *i_lor) {
    e->ade = (ade_dir *)realloc(e->ade, (e->nade + 1) * sizeof(ade_dir));
    ade_dir *d = &e->ade[e->nade++];
    memset(d, 0, sizeof(*d));
    d->dir = dir; d->ncells = ncells; d->lorentz = lorentz;
    d->pos = dup_u32(pos, ncells);
    d->v_int = dup_real(v_int, ncells); d->v_ext = dup_real(v_ext, ncells); d->v_lor = dup_real(v_lor, ncells);
    d->i_int = dup_real(i_int, ncells); d->i_ext = dup_real(i_ext, ncells); d->i_lor = dup_real(i_lor, ncells);
    d->volt_ade = (real *)xcalloc(ncells, sizeof(real)); d->volt_lor = (real *)xcalloc(ncells, sizeof(real));
    d->curr_ade = (real *)xcalloc(ncells, sizeof(real)); d->curr_lor = (real *)xcalloc(ncells, sizeof(real));
}

static void ade_update(orc *e, int is_curr) {
    for (uint32_t k = 0; k < e->nade; ++k) {
        ade_dir *d = &e->ade[k];
        real *fld = is_curr ? e->curr : e->volt;
        real *A = is_curr ? d->curr_ade : d->volt_ade;
        real *L = is_curr ? d->curr_lor : d->volt_lor;
        const real *ci = is_curr ? d->i_int : d->v_int;
        const real *ce = is_curr ? d->i_ext : d->v_ext;
        const real *cl = is_curr ? d->i_lor : d->v_lor;
        for (uint32_t i = 0; i < d->ncells; ++i) {
            const double V = fld[(size_t)d->dir * e->N + d->pos[i]];
            if (d->lorentz) {
                L[i] = ST((double)L[i] + (double)cl[i] * A[i]);
                A[i] = ST((double)ci[i] * A[i] + (double)ce[i] * (V - (double)L[i]));
            } else {
                A[i] = ST((double)ci[i] * A[i] + (double)ce[i] * V);
            }
        }
    }
}

void PFX(update_volt_ade)(orc *e) { ade_update(e, 0); }
void PFX(update_curr_ade)(orc *e) { ade_update(e, 1); }

/* Optional openEMS write-back (Engine_Ext_Dispersive::Apply2Voltages/
 * Apply2Current): field -= ADE state.  Off by default because the reference
 * engine never writes the ADE state back (src/webgpu-fdtd.mjs:405-454). */
static void ade_apply(orc *e, int is_curr) {
    for (uint32_t k = 0; k < e->nade; ++k) {
        ade_dir *d = &e->ade[k];
        real *fld = is_curr ? e->curr : e->volt;
        const real *A = is_curr ? d->curr_ade : d->volt_ade;
        for (uint32_t i = 0; i < d->ncells; ++i) {
            const size_t g = (size_t)d->dir * e->N + d->pos[i];
            fld[g] = ST((double)fld[g] - (double)A[i]);
        }
    }
}

void PFX(get_ade_state)(const orc *e, uint32_t k, double *volt_ade, double *volt_lor, double *curr_ade,
                        double *curr_lor) {
    if (k >= e->nade) return;
    const ade_dir *d = &e->ade[k];
    for (uint32_t i = 0; i < d->ncells; ++i) {
        if (volt_ade) volt_ade[i] = d->volt_ade[i];
        if (volt_lor) volt_lor[i] = d->volt_lor[i];
        if (curr_ade) curr_ade[i] = d->curr_ade[i];
        if (curr_lor) curr_lor[i] = d->curr_lor[i];
    }
}

/* ------------------------------------------------------------------ */
/* TFSF  (src/webgpu-fdtd.mjs:683-739)                                  */
/* ------------------------------------------------------------------ */

void PFX(cfg_tfsf)(orc *e, const double *signal, uint32_t len, uint32_t period, uint32_t nv,
                   const uint32_t *v_idx, const uint32_t *v_dint, const double *v_dfrac,
                   const double *v_amp, uint32_t ni, const uint32_t *i_idx, const uint32_t *i_dint,
                   const double *i_dfrac, const double *i_amp) {
    free(e->tfsf_signal); free(e->tfsf_v); free(e->tfsf_i);
    e->tfsf_signal = dup_real(signal, len);
    e->tfsf_len = len; e->tfsf_period = period;
    e->tfsf_nv = nv; e->tfsf_ni = ni;
    e->tfsf_v = (tfsf_point *)xcalloc(nv, sizeof(tfsf_point));
    e->tfsf_i = (tfsf_point *)xcalloc(ni, sizeof(tfsf_point));
    for (uint32_t k = 0; k < nv; ++k) {
        e->tfsf_v[k].field_idx = v_idx[k]; e->tfsf_v[k].delay_int = v_dint[k];
        e->tfsf_v[k].delay_frac = v_dfrac[k]; e->tfsf_v[k].amp = v_amp[k];
    }
    for (uint32_t k = 0; k < ni; ++k) {
        e->tfsf_i[k].field_idx = i_idx[k]; e->tfsf_i[k].delay_int = i_dint[k];
        e->tfsf_i[k].delay_frac = i_dfrac[k]; e->tfsf_i[k].amp = i_amp[k];
    }
    e->has_tfsf = 1;
}

static void tfsf_apply(orc *e, int is_curr) {
    if (!e->has_tfsf) return;
    const uint32_t np = is_curr ? e->tfsf_ni : e->tfsf_nv;
    const tfsf_point *pts = is_curr ? e->tfsf_i : e->tfsf_v;
    real *fld = is_curr ? e->curr : e->volt;
    for (uint32_t k = 0; k < np; ++k) {
        long d = (long)e->numTS - (long)pts[k].delay_int;
        if (d < 0) continue;
        if (e->tfsf_period > 0) d %= (long)e->tfsf_period;
        if (d >= (long)e->tfsf_len - 1) continue;
        const double delta = pts[k].delay_frac;
        const double sig = (1.0 - delta) * e->tfsf_signal[d] + delta * e->tfsf_signal[d + 1];
        fld[pts[k].field_idx] = ST((double)fld[pts[k].field_idx] + pts[k].amp * sig);
    }
}

void PFX(apply_tfsf_voltage)(orc *e) { tfsf_apply(e, 0); }
void PFX(apply_tfsf_current)(orc *e) { tfsf_apply(e, 1); }

/* ------------------------------------------------------------------ */
/* lumped RLC  (src/webgpu-fdtd.mjs:761-846, shift-register form)      */
/* ------------------------------------------------------------------ */

/* coeffs: 9 doubles per element: i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2 */
void PFX(cfg_rlc)(orc *e, uint32_t n, const uint32_t *field_idx, const uint32_t *direction,
                  const uint32_t *type_flag, const double *coeffs) {
    free(e->rlc); free(e->rlc_vdn); free(e->rlc_jn); free(e->rlc_vil);
    e->rlc_n = n;
    e->rlc = (rlc_elem *)xcalloc(n, sizeof(rlc_elem));
    for (uint32_t k = 0; k < n; ++k) {
        rlc_elem *r = &e->rlc[k];
        r->field_idx = field_idx[k]; r->direction = direction[k]; r->type_flag = type_flag[k];
        const double *c = coeffs + 9 * (size_t)k;
        r->i2v = c[0]; r->ilv = c[1]; r->vvd = c[2]; r->vv2 = c[3]; r->vj1 = c[4];
        r->vj2 = c[5]; r->ib0 = c[6]; r->b1 = c[7]; r->b2 = c[8];
    }
    e->rlc_vdn = (real *)xcalloc(3 * (size_t)n, sizeof(real));
    e->rlc_jn = (real *)xcalloc(3 * (size_t)n, sizeof(real));
    e->rlc_vil = (real *)xcalloc(n, sizeof(real));
}

void PFX(pre_voltage_rlc)(orc *e) {
    for (uint32_t n = 0; n < e->rlc_n; ++n) {
        real *V = e->rlc_vdn + 3 * (size_t)n;
        V[2] = V[1];
        V[1] = V[0];
        if (e->rlc[n].type_flag == 0)
            e->rlc_vil[n] = ST((double)e->rlc_vil[n] + e->rlc[n].i2v * e->rlc[n].ilv * (double)V[1]);
    }
}

void PFX(apply_rlc)(orc *e) {
    for (uint32_t n = 0; n < e->rlc_n; ++n) {
        const rlc_elem *r = &e->rlc[n];
        const size_t g = (size_t)r->direction * e->N + r->field_idx;
        real *V = e->rlc_vdn + 3 * (size_t)n;
        real *J = e->rlc_jn + 3 * (size_t)n;
        V[0] = e->volt[g];
        J[2] = J[1];
        J[1] = J[0];
        if (r->type_flag == 1) {
            const double Il = e->rlc_vil[n];
            V[0] = ST(r->vvd * ((double)V[0] - Il + r->vv2 * V[2] + r->vj1 * J[1] + r->vj2 * J[2]));
            J[0] = ST(r->ib0 * ((double)V[0] - (double)V[2]) - r->b1 * r->ib0 * J[1] - r->b2 * r->ib0 * J[2]);
            e->volt[g] = V[0];
        }
    }
}

void PFX(get_rlc_state)(const orc *e, double *vdn, double *jn, double *vil) {
    for (size_t i = 0; i < 3 * (size_t)e->rlc_n; ++i) {
        if (vdn) vdn[i] = e->rlc_vdn[i];
        if (jn) jn[i] = e->rlc_jn[i];
    }
    for (size_t i = 0; i < e->rlc_n; ++i)
        if (vil) vil[i] = e->rlc_vil[i];
}

/* ------------------------------------------------------------------ */
/* Mur ABC  (src/webgpu-fdtd.mjs:863-1023)                              */
/* ------------------------------------------------------------------ */

void PFX(cfg_mur)(orc *e, uint32_t n, const double *coeff, const uint32_t *normal_idx,
                  const uint32_t *shifted_idx) {
    free(e->mur_coeff); free(e->mur_saved); free(e->mur_normal); free(e->mur_shifted);
    e->mur_n = n;
    e->mur_coeff = dup_real(coeff, n);
    e->mur_saved = (real *)xcalloc(n, sizeof(real));
    e->mur_normal = dup_u32(normal_idx, n);
    e->mur_shifted = dup_u32(shifted_idx, n);
}

void PFX(mur_pre)(orc *e) {
    for (uint32_t n = 0; n < e->mur_n; ++n)
        e->mur_saved[n] = ST((double)e->volt[e->mur_shifted[n]] -
                             (double)e->mur_coeff[n] * e->volt[e->mur_normal[n]]);
}

void PFX(mur_post)(orc *e) {
    for (uint32_t n = 0; n < e->mur_n; ++n)
        e->mur_saved[n] = ST((double)e->mur_saved[n] + (double)e->mur_coeff[n] * e->volt[e->mur_shifted[n]]);
}

void PFX(mur_apply)(orc *e) {
    for (uint32_t n = 0; n < e->mur_n; ++n) e->volt[e->mur_normal[n]] = e->mur_saved[n];
}

/* ------------------------------------------------------------------ */
/* NF2FF accumulation  (src/webgpu-engine.mjs:637-800, 2729-2838)      */
/* ------------------------------------------------------------------ */

void PFX(cfg_nf2ff)(orc *e, const uint32_t *points, uint32_t P, const double *omega, uint32_t F,
                    double dT, const double *const el[6], const uint32_t el_len[6], uint32_t maxTS,
                    uint32_t window) {
    free(e->nf_points); free(e->nf_omega); free(e->nf_accE); free(e->nf_accH);
    for (int k = 0; k < 6; ++k) { free(e->nf_el[k]); e->nf_el[k] = NULL; }
    e->nf_P = P; e->nf_F = F; e->nf_dT = dT; e->nf_maxTS = maxTS; e->nf_window = window;
    e->nf_points = dup_u32(points, 3 * (size_t)P);
    e->nf_omega = (double *)xcalloc(F, sizeof(double));
    memcpy(e->nf_omega, omega, F * sizeof(double));
    for (int k = 0; k < 6; ++k) {
        e->nf_el_len[k] = el_len[k];
        e->nf_el[k] = (double *)xcalloc(el_len[k], sizeof(double));
        memcpy(e->nf_el[k], el[k], el_len[k] * sizeof(double));
    }
    e->nf_wnorm = 1.0;
    if (window == 1 && maxTS > 1) {
        double s = 0;
        for (uint32_t n = 0; n < maxTS; ++n) s += 0.5 * (1 - cos(2 * M_PI * n / (maxTS - 1)));
        e->nf_wnorm = 1.0 / s;
    }
    e->nf_accE = (double *)xcalloc((size_t)F * 6 * P, sizeof(double));
    e->nf_accH = (double *)xcalloc((size_t)F * 6 * P, sizeof(double));
}

/* Cell-interpolated E and H at one surface point (CELL_INTERPOLATE). */
void PFX(nf2ff_interp)(const orc *e, uint32_t ix, uint32_t iy, uint32_t iz, double E[3], double H[3]) {
    const uint32_t Nx = e->nx, Ny = e->ny, Nz = e->nz;
    const real *v = e->volt, *c = e->curr;
    const size_t N = e->N;
#define FI(X, Y, Z) ((size_t)(X) * Ny * Nz + (size_t)(Y) * Nz + (Z))
    const size_t b = FI(ix, iy, iz);
    E[0] = E[1] = E[2] = 0;
    H[0] = H[1] = H[2] = 0;
    if (!(ix >= Nx - 1 || iy >= Ny - 1 || iz >= Nz - 1)) {
        E[0] = ((double)v[b] + v[FI(ix, iy + 1, iz)] + v[FI(ix, iy + 1, iz + 1)] + v[FI(ix, iy, iz + 1)]) /
               (4.0 * e->nf_el[0][ix]);
        E[1] = ((double)v[N + b] + v[N + FI(ix, iy, iz + 1)] + v[N + FI(ix + 1, iy, iz + 1)] +
                v[N + FI(ix + 1, iy, iz)]) / (4.0 * e->nf_el[1][iy]);
        E[2] = ((double)v[2 * N + b] + v[2 * N + FI(ix + 1, iy, iz)] + v[2 * N + FI(ix + 1, iy + 1, iz)] +
                v[2 * N + FI(ix, iy + 1, iz)]) / (4.0 * e->nf_el[2][iz]);
    }
    const uint32_t pos[3] = {ix, iy, iz};
    const uint32_t lim[3] = {Nx, Ny, Nz};
    const size_t step[3] = {(size_t)Ny * Nz, Nz, 1};
    for (int n = 0; n < 3; ++n) {
        const double d0 = e->nf_el[3 + n][pos[n]];
        if (d0 > 0.0) {
            double h = c[n * N + b] / d0;
            if (pos[n] < lim[n] - 1) {
                const double d1 = e->nf_el[3 + n][pos[n] + 1];
                const double dr = d0 / (d0 + d1);
                const double up = c[n * N + b + step[n]] / d1;
                h = h * (1.0 - dr) + up * dr;
            }
            H[n] = h;
        }
    }
#undef FI
}

/* One accumulation with phase index ts (the engine passes numTS+1 after the
 * step, src/webgpu-engine.mjs:2427). */
void PFX(nf2ff_accumulate)(orc *e, uint32_t ts) {
    if (!e->nf_P) return;
    double w = 1.0;
    if (e->nf_window == 1 && e->nf_maxTS > 1) w = 0.5 * (1.0 - cos(2.0 * M_PI * ts / (double)(e->nf_maxTS - 1)));
    const double wdT = w * e->nf_wnorm * e->nf_dT;
    const size_t P6 = 6 * (size_t)e->nf_P;
#pragma omp parallel for schedule(static) if (e->nf_P >= OMP_MIN_CELLS / 16)
    for (long p = 0; p < (long)e->nf_P; ++p) {
        double E[3], H[3];
        PFX(nf2ff_interp)(e, e->nf_points[3 * p], e->nf_points[3 * p + 1], e->nf_points[3 * p + 2], E, H);
        for (uint32_t f = 0; f < e->nf_F; ++f) {
            const double pe = e->nf_omega[f] * ts * e->nf_dT;
            const double ph = e->nf_omega[f] * (ts + 0.5) * e->nf_dT;
            const double ce = cos(pe), se = sin(pe), ch = cos(ph), sh = sin(ph);
            double *aE = e->nf_accE + f * P6 + 6 * p;
            double *aH = e->nf_accH + f * P6 + 6 * p;
            for (int n = 0; n < 3; ++n) {
                aE[2 * n] += E[n] * ce * wdT;
                aE[2 * n + 1] += E[n] * (-se) * wdT;
                aH[2 * n] += H[n] * ch * wdT;
                aH[2 * n + 1] += H[n] * (-sh) * wdT;
            }
        }
    }
}

void PFX(read_nf2ff)(const orc *e, double *accE, double *accH) {
    const size_t n = (size_t)e->nf_F * 6 * e->nf_P;
    if (accE) memcpy(accE, e->nf_accE, n * sizeof(double));
    if (accH) memcpy(accH, e->nf_accH, n * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* far field  (src/webgpu-engine.mjs:802-986, 2871-2967)               */
/* ------------------------------------------------------------------ */

/* accE/accH: 6P doubles (one frequency); meta: 8 per point (pos xyz,
 * normalDir, normSign, area, pad, pad); out: 5 per angle.
 * Mirror (mdir >= 0; src/nf2ff.mjs:377-398): the surface is added a second
 * time reflected in the plane x[mdir] = mpos with the image fields of a PEC
 * (mtype 0: tangential E and normal H reversed) or PMC (mtype 1: normal E
 * and tangential H reversed) plane.  The reflected copy of a face
 * perpendicular to the plane faces the other way, so its normal sign is
 * reversed; refrule = 1 keeps it, as the reference does. */
static void farfield_core(const double *accE, const double *accH, const double *meta, uint32_t P,
                          const double *theta, uint32_t nT, const double *phi, uint32_t nP, const double center[3],
                          double freq, double radius, int mdir, double mpos, int mtype, int refrule, double *out) {
    const double C0 = 299792458.0;
    const double Z0 = 376.73031346177066;
    const double k = 2 * M_PI * freq / C0;
    const double fmag = k / (4 * M_PI * radius);
    const double fph = -k * radius;
    const double fre = fmag * (-sin(fph));
    const double fim = fmag * cos(fph);
    double ef[3] = {1, 1, 1}, hf[3] = {1, 1, 1};
    if (mdir >= 0) {
        const int a = (mdir + 1) % 3, b = (mdir + 2) % 3;
        if (mtype == 0) { hf[mdir] = -1; ef[a] = -1; ef[b] = -1; }
        else { ef[mdir] = -1; hf[a] = -1; hf[b] = -1; }
    }
    const int passes = mdir >= 0 ? 2 : 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (long a = 0; a < (long)nT * nP; ++a) {
        const double th = theta[a / nP], ph = phi[a % nP];
        const double sT = sin(th), cT = cos(th), sP = sin(ph), cP = cos(ph);
        double Ntr = 0, Nti = 0, Npr = 0, Npi = 0, Ltr = 0, Lti = 0, Lpr = 0, Lpi = 0;
        for (int pass = 0; pass < passes; ++pass)
        for (uint32_t p = 0; p < P; ++p) {
            const double *m = meta + 8 * (size_t)p;
            const int nd = (int)m[3];
            double s = m[4];
            const double area = m[5];
            double pos[3] = {m[0], m[1], m[2]};
            double E[6], H[6];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 2; ++r) {
                    E[2 * c + r] = accE[6 * (size_t)p + 2 * c + r] * (pass ? ef[c] : 1.0);
                    H[2 * c + r] = accH[6 * (size_t)p + 2 * c + r] * (pass ? hf[c] : 1.0);
                }
            if (pass) {
                pos[mdir] = 2 * mpos - pos[mdir];
                if (nd == mdir && !refrule) s = -s;
            }
            double J[3][2] = {{0, 0}, {0, 0}, {0, 0}}, M[3][2] = {{0, 0}, {0, 0}, {0, 0}};
            for (int r = 0; r < 2; ++r) {
                if (nd == 0) {
                    J[1][r] = -s * H[4 + r]; J[2][r] = s * H[2 + r];
                    M[1][r] = s * E[4 + r]; M[2][r] = -s * E[2 + r];
                } else if (nd == 1) {
                    J[0][r] = s * H[4 + r]; J[2][r] = -s * H[0 + r];
                    M[0][r] = -s * E[4 + r]; M[2][r] = s * E[0 + r];
                } else {
                    J[0][r] = -s * H[2 + r]; J[1][r] = s * H[0 + r];
                    M[0][r] = s * E[2 + r]; M[1][r] = -s * E[0 + r];
                }
            }
            const double rc = (pos[0] - center[0]) * sT * cP + (pos[1] - center[1]) * sT * sP + (pos[2] - center[2]) * cT;
            const double phs = k * rc;
            const double aer = area * cos(phs), aei = area * sin(phs);
            double Jt[2], Jp[2], Mt[2], Mp[2];
            for (int r = 0; r < 2; ++r) {
                Jt[r] = J[0][r] * cT * cP + J[1][r] * cT * sP - J[2][r] * sT;
                Jp[r] = J[1][r] * cP - J[0][r] * sP;
                Mt[r] = M[0][r] * cT * cP + M[1][r] * cT * sP - M[2][r] * sT;
                Mp[r] = M[1][r] * cP - M[0][r] * sP;
            }
            Ntr += aer * Jt[0] - aei * Jt[1]; Nti += aer * Jt[1] + aei * Jt[0];
            Npr += aer * Jp[0] - aei * Jp[1]; Npi += aer * Jp[1] + aei * Jp[0];
            Ltr += aer * Mt[0] - aei * Mt[1]; Lti += aer * Mt[1] + aei * Mt[0];
            Lpr += aer * Mp[0] - aei * Mp[1]; Lpi += aer * Mp[1] + aei * Mp[0];
        }
        const double a1r = Lpr + Z0 * Ntr, a1i = Lpi + Z0 * Nti;
        const double Etr = -(fre * a1r - fim * a1i), Eti = -(fre * a1i + fim * a1r);
        const double a2r = Ltr - Z0 * Npr, a2i = Lti - Z0 * Npi;
        const double Epr = fre * a2r - fim * a2i, Epi = fre * a2i + fim * a2r;
        double *o = out + 5 * a;
        o[0] = (Etr * Etr + Eti * Eti + Epr * Epr + Epi * Epi) / (2.0 * Z0);
        o[1] = Etr; o[2] = Eti; o[3] = Epr; o[4] = Epi;
    }
}

void PFX(farfield)(const double *accE, const double *accH, const double *meta, uint32_t P,
                   const double *theta, uint32_t nT, const double *phi, uint32_t nP,
                   const double center[3], double freq, double radius, double *out) {
    farfield_core(accE, accH, meta, P, theta, nT, phi, nP, center, freq, radius, -1, 0.0, 0, 0, out);
}

void PFX(farfield_mirror)(const double *accE, const double *accH, const double *meta, uint32_t P,
                          const double *theta, uint32_t nT, const double *phi, uint32_t nP,
                          const double center[3], double freq, double radius, int mdir, double mpos, int mtype,
                          int refrule, double *out) {
    farfield_core(accE, accH, meta, P, theta, nT, phi, nP, center, freq, radius, mdir, mpos, mtype, refrule, out);
}

/* ------------------------------------------------------------------ */
/* energy  (src/webgpu-engine.mjs:3039-3117): sum of v^2+c^2 over the    */
/* (Nx-1)(Ny-1)(Nz-1) interior, all 3 components, in FP64.             */
/* ------------------------------------------------------------------ */

double PFX(energy)(const orc *e) {
    const long Nx = e->nx, Ny = e->ny, Nz = e->nz;
    double total = 0;
    double *plane = (double *)xcalloc(Nx > 1 ? Nx - 1 : 1, sizeof(double));
#pragma omp parallel for schedule(static) if (Nx * Ny * Nz >= OMP_MIN_CELLS)
    for (long x = 0; x < Nx - 1; ++x) {
        double s = 0;
        for (long y = 0; y < Ny - 1; ++y)
            for (long z = 0; z < Nz - 1; ++z)
                for (int n = 0; n < 3; ++n) {
                    const double v = e->volt[IDX(e, n, x, y, z)], c = e->curr[IDX(e, n, x, y, z)];
                    s += v * v + c * c;
                }
        plane[x] = s;
    }
    for (long x = 0; x < Nx - 1; ++x) total += plane[x];
    free(plane);
    return total;
}

/* ------------------------------------------------------------------ */
/* buffered probe gather  (src/webgpu-engine.mjs:2615-2687, 2458-2466)  */
/* ------------------------------------------------------------------ */

void PFX(cfg_gather_buffered)(orc *e, const uint32_t *idx, uint32_t n, uint32_t max_samples,
                              uint32_t interval) {
    free(e->g_idx); free(e->g_ring);
    e->g_idx = dup_u32(idx, n);
    e->g_n = n; e->g_max = max_samples; e->g_interval = interval ? interval : 1; e->g_count = 0;
    e->g_ring = (real *)xcalloc((size_t)max_samples * 2 * n, sizeof(real));
}

uint32_t PFX(read_gather_buffered)(const orc *e, double *out) {
    const size_t n = (size_t)e->g_count * 2 * e->g_n;
    if (out)
        for (size_t i = 0; i < n; ++i) out[i] = e->g_ring[i];
    return e->g_count;
}

static void gather_hook(orc *e) {
    if (!e->g_n) return;
    if ((e->numTS + 1) % e->g_interval != 0 || e->g_count >= e->g_max) return;
    real *row = e->g_ring + (size_t)e->g_count * 2 * e->g_n;
    for (uint32_t i = 0; i < e->g_n; ++i) {
        row[2 * i] = e->volt[e->g_idx[i]];
        row[2 * i + 1] = e->curr[e->g_idx[i]];
    }
    e->g_count++;
}

/* ------------------------------------------------------------------ */
/* steady-state detection  (src/webgpu-fdtd.mjs:1037-1095)             */
/* ------------------------------------------------------------------ */

void PFX(cfg_steady_state)(orc *e, const uint32_t *idx, uint32_t n, uint32_t period, double threshold) {
    free(e->ss_idx); free(e->ss_e1); free(e->ss_e2);
    e->ss_idx = dup_u32(idx, n);
    e->ss_n = n; e->ss_period = period; e->ss_threshold = threshold > 0 ? threshold : 1e-6;
    e->ss_current = 0; e->ss_recording = 0;
    e->ss_e1 = (real *)xcalloc(n, sizeof(real));
    e->ss_e2 = (real *)xcalloc(n, sizeof(real));
}

void PFX(ss_set)(orc *e, int recording, uint32_t current) { e->ss_recording = recording; e->ss_current = current; }

uint32_t PFX(ss_read)(const orc *e, double *e1, double *e2) {
    for (uint32_t n = 0; n < e->ss_n; ++n) {
        if (e1) e1[n] = e->ss_e1[n];
        if (e2) e2[n] = e->ss_e2[n];
    }
    return e->ss_current;
}

void PFX(ss_accumulate)(orc *e) {
    if (!e->ss_n || !e->ss_recording) return;
    for (uint32_t n = 0; n < e->ss_n; ++n) {
        const double v = e->volt[e->ss_idx[n]];
        const double en = v * v;
        if (e->ss_current < e->ss_period) e->ss_e1[n] = ST(e->ss_e1[n] + en);
        else e->ss_e2[n] = ST(e->ss_e2[n] + en);
    }
    e->ss_current++;
}

/* checkConvergence: both periods recorded and max |e2-e1|/|e1| < threshold */
int PFX(ss_converged)(const orc *e) {
    if (!e->ss_n && !e->ss_period) return 0;
    if (e->ss_current < 2 * e->ss_period) return 0;
    double mx = 0;
    for (uint32_t n = 0; n < e->ss_n; ++n) {
        const double e1 = e->ss_e1[n], e2 = e->ss_e2[n];
        if (fabs(e1) < 1e-30) continue;
        const double r = fabs(e2 - e1) / fabs(e1);
        if (r > mx) mx = r;
    }
    return mx < e->ss_threshold;
}

/* ------------------------------------------------------------------ */
/* periodic boundaries  (src/webgpu-engine.mjs:475-584, 1976-2155)      */
/* voltage: tangential E of face N-2 -> face 0; current: tangential H  */
/* of face 0 -> face N-2 over (face-1) x (face-1); Bloch phase rotates */
/* (field, imag) by the phase with the imaginary part kept per face.   */
/* ------------------------------------------------------------------ */

void PFX(cfg_pbc)(orc *e, uint32_t n, const uint32_t *axis, const double *phase) {
    for (uint32_t k = 0; k < e->pbc_n; ++k) { free(e->pbc_imag[k]); e->pbc_imag[k] = NULL; }
    e->pbc_n = n > 3 ? 3 : n;
    const uint32_t dims[3] = {e->nx, e->ny, e->nz};
    for (uint32_t k = 0; k < e->pbc_n; ++k) {
        e->pbc_axis[k] = axis[k];
        e->pbc_cos[k] = (float)cos(phase[k]);
        e->pbc_sin[k] = (float)sin(phase[k]);
        const size_t F = (size_t)dims[(axis[k] + 1) % 3] * dims[(axis[k] + 2) % 3];
        e->pbc_imag[k] = e->pbc_sin[k] != 0.0f ? (real *)xcalloc(4 * F, sizeof(real)) : NULL;
    }
}

static void pbc_apply(orc *e, int current) {
    const uint32_t dims[3] = {e->nx, e->ny, e->nz};
    real *fld = current ? e->curr : e->volt;
    for (uint32_t k = 0; k < e->pbc_n; ++k) {
        const uint32_t ax = e->pbc_axis[k], t0 = (ax + 1) % 3, t1 = (ax + 2) % 3;
        const uint32_t f0 = dims[t0], f1 = dims[t1];
        const uint32_t src = current ? 0 : dims[ax] - 2, dst = current ? dims[ax] - 2 : 0;
        const uint32_t b0 = current ? f0 - 1 : f0, b1 = current ? f1 - 1 : f1;
        const double c = e->pbc_cos[k], sn = e->pbc_sin[k];
        real *im = e->pbc_imag[k] ? e->pbc_imag[k] + (current ? 2 : 0) * (size_t)f0 * f1 : NULL;
        for (uint32_t i0 = 0; i0 < b0; ++i0)
            for (uint32_t i1 = 0; i1 < b1; ++i1) {
                uint32_t ps[3], pd[3];
                ps[ax] = src; pd[ax] = dst;
                ps[t0] = pd[t0] = i0;
                ps[t1] = pd[t1] = i1;
                const size_t fl = (size_t)i0 * f1 + i1;
                for (int q = 0; q < 2; ++q) {
                    const uint32_t comp = q ? t1 : t0;
                    const size_t si = IDX(e, comp, ps[0], ps[1], ps[2]), di = IDX(e, comp, pd[0], pd[1], pd[2]);
                    if (!im) {
                        fld[di] = ST(c * fld[si]);
                    } else {
                        real *ip = im + (size_t)q * f0 * f1 + fl;
                        const double re = fld[si], ii = *ip;
                        fld[di] = ST(c * re - sn * ii);
                        *ip = ST(sn * re + c * ii);
                    }
                }
            }
    }
}

void PFX(pbc_voltage)(orc *e) { pbc_apply(e, 0); }
void PFX(pbc_current)(orc *e) { pbc_apply(e, 1); }

/* ------------------------------------------------------------------ */
/* one complete step  (src/webgpu-fdtd.mjs:1094-1172 order)             */
/* ------------------------------------------------------------------ */

void PFX(step)(orc *e) {
    PFX(ss_accumulate)(e); /* pre-voltage, priority +2M */
    PFX(pml_step)(e, 0);
    PFX(update_volt_ade)(e);
    PFX(mur_pre)(e);
    PFX(pre_voltage_rlc)(e);
    PFX(update_voltages)(e);
    PFX(pml_step)(e, 1);
    PFX(apply_tfsf_voltage)(e);
    pbc_apply(e, 0); /* after TFSF-V, before Mur post (src/webgpu-engine.mjs:2461) */
    PFX(mur_post)(e);
    if (e->ade_apply) ade_apply(e, 0);
    PFX(apply_excitation)(e);
    PFX(mur_apply)(e);
    PFX(apply_rlc)(e);
    PFX(pml_step)(e, 2);
    PFX(update_curr_ade)(e);
    PFX(update_currents)(e);
    PFX(pml_step)(e, 3);
    PFX(apply_tfsf_current)(e);
    pbc_apply(e, 1); /* after TFSF-I (src/webgpu-engine.mjs:2477) */
    if (e->ade_apply) ade_apply(e, 1);
    /* engine hooks after all field updates (src/webgpu-engine.mjs:2483-2500) */
    if (e->nf_P) PFX(nf2ff_accumulate)(e, e->numTS + 1);
    gather_hook(e);
    e->numTS++;
}

void PFX(iterate)(orc *e, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) PFX(step)(e);
}
