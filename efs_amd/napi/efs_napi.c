/*
 * efs_napi.c -- Node-API binding of libefs_engine.so (include/efs_engine.h,
 * include/efs_batch.h).
 *
 * A thin, synchronous layer: every function takes the context handle (an
 * external value) plus numbers and TypedArrays, calls the matching efs_*
 * entry point and throws a JavaScript Error carrying efs_last_error() when
 * the call fails.  efs_amd/js/webgpu-engine.mjs builds the WebGPUEngine-
 * compatible class on top of it.  Built with plain gcc against node_api.h
 * (N-API, Node >= 12); napi_* symbols resolve from the node binary.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "efs_batch.h"
#include "efs_engine.h"

#define MAXARGS 16

typedef struct {
    napi_env env;
    napi_value argv[MAXARGS];
    size_t argc;
    int failed;
} call_t;

static int begin(napi_env env, napi_callback_info info, call_t *c) {
    c->env = env;
    c->argc = MAXARGS;
    c->failed = 0;
    if (napi_get_cb_info(env, info, &c->argc, c->argv, NULL, NULL) != napi_ok) {
        napi_throw_error(env, NULL, "efs: bad call");
        return -1;
    }
    for (size_t i = c->argc; i < MAXARGS; ++i) napi_get_undefined(env, &c->argv[i]);
    return 0;
}

static void throwf(call_t *c, const char *msg) {
    if (!c->failed) napi_throw_error(c->env, NULL, msg);
    c->failed = 1;
}

static efs_ctx *ctx_arg(call_t *c, size_t i) {
    void *p = NULL;
    if (napi_get_value_external(c->env, c->argv[i], &p) != napi_ok || !p) {
        throwf(c, "efs: invalid engine handle");
        return NULL;
    }
    return (efs_ctx *)p;
}

static uint32_t u32_arg(call_t *c, size_t i) {
    uint32_t v = 0;
    if (napi_get_value_uint32(c->env, c->argv[i], &v) != napi_ok) throwf(c, "efs: expected an unsigned integer");
    return v;
}

static double f64_arg(call_t *c, size_t i) {
    double v = 0;
    if (napi_get_value_double(c->env, c->argv[i], &v) != napi_ok) throwf(c, "efs: expected a number");
    return v;
}

/* TypedArray of the given element type; NULL/undefined allowed when opt. */
static void *arr_arg(call_t *c, size_t i, napi_typedarray_type want, size_t *len, int opt) {
    napi_valuetype t;
    *len = 0;
    napi_typeof(c->env, c->argv[i], &t);
    if (t == napi_undefined || t == napi_null) {
        if (!opt) throwf(c, "efs: missing typed array argument");
        return NULL;
    }
    bool is_ta = false;
    napi_is_typedarray(c->env, c->argv[i], &is_ta);
    if (!is_ta) {
        throwf(c, "efs: expected a typed array");
        return NULL;
    }
    napi_typedarray_type type;
    void *data = NULL;
    napi_value ab;
    size_t off = 0;
    if (napi_get_typedarray_info(c->env, c->argv[i], &type, len, &data, &ab, &off) != napi_ok) {
        throwf(c, "efs: unreadable typed array");
        return NULL;
    }
    if (type != want) {
        throwf(c, want == napi_float32_array   ? "efs: expected a Float32Array"
                  : want == napi_uint32_array  ? "efs: expected a Uint32Array"
                                               : "efs: expected a Float64Array");
        return NULL;
    }
    return data;
}

#define F32(i, n) ((float *)arr_arg(&c, (i), napi_float32_array, &(n), 0))
#define F32OPT(i, n) ((float *)arr_arg(&c, (i), napi_float32_array, &(n), 1))
#define U32(i, n) ((uint32_t *)arr_arg(&c, (i), napi_uint32_array, &(n), 0))
#define F64(i, n) ((double *)arr_arg(&c, (i), napi_float64_array, &(n), 0))

static napi_value undef(napi_env env) {
    napi_value u;
    napi_get_undefined(env, &u);
    return u;
}

static napi_value num(napi_env env, double v) {
    napi_value r;
    napi_create_double(env, v, &r);
    return r;
}

/* Map a status to a thrown Error with the engine's message. */
static napi_value status(call_t *c, efs_ctx *ctx, int rc) {
    if (c->failed) return NULL;
    if (rc != EFS_OK) {
        throwf(c, efs_last_error(ctx));
        return NULL;
    }
    return undef(c->env);
}

/* The context's sizes, to check the caller's arrays against what the
 * engine will read or write (a mismatch throws instead of overrunning the
 * typed array). */
static int sizes_of(call_t *c, efs_ctx *ctx, efs_sizes *s) {
    if (c->failed) return -1;
    if (efs_get_sizes(ctx, s) != EFS_OK) {
        throwf(c, "efs: invalid engine handle");
        return -1;
    }
    return 0;
}

static int check_len(call_t *c, const void *p, size_t len, size_t need, const char *msg) {
    if (c->failed) return -1;
    if (p && len != need) {
        char buf[256];
        snprintf(buf, sizeof buf, "efs: %s (got %zu, expected %zu)", msg, len, need);
        throwf(c, buf);
        return -1;
    }
    return 0;
}

#define BEGIN()                                  \
    call_t c;                                    \
    if (begin(env, info, &c)) return NULL;
#define CTX(i)                                   \
    efs_ctx *ctx = ctx_arg(&c, (i));             \
    if (!ctx) return NULL;
#define BAIL() \
    if (c.failed) return NULL;

static napi_value js_device_count(napi_env env, napi_callback_info info) {
    (void)info;
    return num(env, efs_device_count());
}

static napi_value js_version(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value s;
    napi_create_string_utf8(env, efs_version(), NAPI_AUTO_LENGTH, &s);
    return s;
}

/* create(device) -> handle, or null when no GPU is present. */
static napi_value js_create(napi_env env, napi_callback_info info) {
    BEGIN();
    uint32_t dev = c.argc > 0 ? u32_arg(&c, 0) : 0;
    BAIL();
    efs_ctx *ctx = NULL;
    int rc = efs_create(&ctx, (int)dev);
    if (rc == EFS_ERR_NO_DEVICE) {
        napi_value n;
        napi_get_null(env, &n);
        return n;
    }
    if (rc != EFS_OK) {
        throwf(&c, efs_last_error(NULL));
        return NULL;
    }
    napi_value h;
    napi_create_external(env, ctx, NULL, NULL, &h);
    return h;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    efs_destroy(ctx);
    return undef(env);
}

static napi_value js_last_error(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    napi_value s;
    napi_create_string_utf8(env, efs_last_error(ctx), NAPI_AUTO_LENGTH, &s);
    return s;
}

/* setStorage(h, fmt): 0 = FP32, 1 = FP16, 2 = FP64 (takes effect at init) */
static napi_value js_set_storage(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    uint32_t fmt = u32_arg(&c, 1);
    BAIL();
    return status(&c, ctx, efs_set_storage(ctx, (int)fmt));
}

/* fieldScales(h, out Float32Array(6)) */
static napi_value js_field_scales(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t n;
    float *out = F32(1, n);
    BAIL();
    if (n < 6) {
        throwf(&c, "efs: fieldScales needs a Float32Array of 6");
        return NULL;
    }
    return status(&c, ctx, efs_field_scales(ctx, out));
}

static napi_value js_init(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b, d, e;
    uint32_t nx = u32_arg(&c, 1), ny = u32_arg(&c, 2), nz = u32_arg(&c, 3);
    float *vv = F32(4, a), *vi = F32(5, b), *ii = F32(6, d), *iv = F32(7, e);
    BAIL();
    size_t need = 3 * (size_t)nx * ny * nz;
    if (a != need || b != need || d != need || e != need) {
        throwf(&c, "efs: coefficient arrays must hold 3*Nx*Ny*Nz floats");
        return NULL;
    }
    return status(&c, ctx, efs_init(ctx, nx, ny, nz, vv, vi, ii, iv));
}

/* cfgExcitation(h, signal, delay, amp, dir, pos, period) */
static napi_value js_cfg_excitation(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t ls, l1, l2, l3, l4;
    float *sig = F32(1, ls);
    uint32_t *delay = U32(2, l1);
    float *amp = F32(3, l2);
    uint32_t *dir = U32(4, l3), *pos = U32(5, l4);
    uint32_t period = u32_arg(&c, 6);
    BAIL();
    if (l1 != l2 || l2 != l3 || l3 != l4) {
        throwf(&c, "efs: excitation arrays must have equal length");
        return NULL;
    }
    return status(&c, ctx, efs_cfg_excitation(ctx, sig, (uint32_t)ls, delay, amp, dir, pos, (uint32_t)l2, period));
}

/* cfgPML(h, start, num, vv, vvfo, vvfn, ii, iifo, iifn) (concatenated) */
static napi_value js_cfg_pml(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t ls, ln, l[6];
    uint32_t *start = U32(1, ls), *nl = U32(2, ln);
    float *p[6];
    for (int k = 0; k < 6; ++k) p[k] = F32(3 + k, l[k]);
    BAIL();
    if (ls != ln || ls % 3) {
        throwf(&c, "efs: PML start/numLines must hold 3 values per region");
        return NULL;
    }
    size_t total = 0;
    for (size_t r = 0; r < ls / 3; ++r) total += 3 * (size_t)nl[3 * r] * nl[3 * r + 1] * nl[3 * r + 2];
    for (int k = 0; k < 6; ++k)
        if (ls && l[k] != total) {
            throwf(&c, "efs: PML coefficient arrays do not match the region sizes");
            return NULL;
        }
    return status(&c, ctx, efs_cfg_pml(ctx, (uint32_t)(ls / 3), start, nl, p[0], p[1], p[2], p[3], p[4], p[5]));
}

/* cfgMur(h, coeff, normalIdx, shiftedIdx) */
static napi_value js_cfg_mur(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b, d;
    float *co = F32(1, a);
    uint32_t *ni = U32(2, b), *si = U32(3, d);
    BAIL();
    if (a != b || b != d) {
        throwf(&c, "efs: Mur arrays must have equal length");
        return NULL;
    }
    return status(&c, ctx, efs_cfg_mur(ctx, (uint32_t)a, co, ni, si));
}

static napi_value js_ade_clear(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    return status(&c, ctx, efs_cfg_ade_clear(ctx));
}

/* adeAdd(h, hasLorentz, dir, pos, vInt, vExt, vLor, iInt, iExt, iLor) */
static napi_value js_ade_add(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    uint32_t lor = u32_arg(&c, 1), dir = u32_arg(&c, 2);
    size_t np_, l[6];
    uint32_t *pos = U32(3, np_);
    float *p[6];
    for (int k = 0; k < 6; ++k) p[k] = F32OPT(4 + k, l[k]);
    BAIL();
    for (int k = 0; k < 6; ++k)
        if (p[k] && l[k] != np_) {
            throwf(&c, "efs: ADE coefficient arrays must match pos_idx");
            return NULL;
        }
    return status(&c, ctx, efs_cfg_ade_add(ctx, lor, dir, (uint32_t)np_, pos, p[0], p[1], p[2], p[3], p[4], p[5]));
}

static napi_value js_ade_apply(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    uint32_t on = u32_arg(&c, 1);
    BAIL();
    return status(&c, ctx, efs_set_ade_apply(ctx, (int)on));
}

/* cfgTFSF(h, signal, period, vIdx, vDint, vDfrac, vAmp, iIdx, iDint, iDfrac, iAmp) */
static napi_value js_cfg_tfsf(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t ls, a[4], b[4];
    float *sig = F32(1, ls);
    uint32_t period = u32_arg(&c, 2);
    uint32_t *vi = U32(3, a[0]), *vd = U32(4, a[1]);
    float *vf = F32(5, a[2]), *va = F32(6, a[3]);
    uint32_t *ii = U32(7, b[0]), *id = U32(8, b[1]);
    float *if_ = F32(9, b[2]), *ia = F32(10, b[3]);
    BAIL();
    for (int k = 1; k < 4; ++k)
        if (a[k] != a[0] || b[k] != b[0]) {
            throwf(&c, "efs: TFSF point arrays must have equal length");
            return NULL;
        }
    return status(&c, ctx, efs_cfg_tfsf(ctx, sig, (uint32_t)ls, period, (uint32_t)a[0], vi, vd, vf, va,
                                        (uint32_t)b[0], ii, id, if_, ia));
}

/* cfgRLC(h, fieldIdx, direction, typeFlag, coeffs9) */
static napi_value js_cfg_rlc(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b, d, e;
    uint32_t *fi = U32(1, a), *dr = U32(2, b), *tf = U32(3, d);
    float *co = F32(4, e);
    BAIL();
    if (a != b || b != d || e != 9 * a) {
        throwf(&c, "efs: RLC arrays do not match (9 coefficients per element)");
        return NULL;
    }
    return status(&c, ctx, efs_cfg_rlc(ctx, (uint32_t)a, fi, dr, tf, co));
}

/* cfgPBC(h, axis Uint32Array, phase Float64Array) */
static napi_value js_cfg_pbc(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b;
    uint32_t *ax = U32(1, a);
    double *ph = F64(2, b);
    BAIL();
    if (a != b) {
        throwf(&c, "efs: PBC axis and phase arrays must have equal length");
        return NULL;
    }
    return status(&c, ctx, efs_cfg_pbc(ctx, (uint32_t)a, ax, ph));
}

/* cfgSteadyState(h, probeIdx, periodSamples) */
static napi_value js_cfg_ss(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t n;
    uint32_t *idx = U32(1, n);
    uint32_t period = u32_arg(&c, 2);
    BAIL();
    return status(&c, ctx, efs_cfg_steady_state(ctx, idx, (uint32_t)n, period));
}

/* ssSet(h, recording, currentSample) */
static napi_value js_ss_set(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    uint32_t rec = u32_arg(&c, 1), cur = u32_arg(&c, 2);
    BAIL();
    return status(&c, ctx, efs_ss_set(ctx, (int)rec, cur));
}

/* ssRead(h, e1 Float32Array | null, e2 | null) -> [currentSample, recording] */
static napi_value js_ss_read(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b;
    float *e1 = F32OPT(1, a), *e2 = F32OPT(2, b);
    BAIL();
    efs_sizes s;
    if (sizes_of(&c, ctx, &s) || check_len(&c, e1, a, s.steady_state_n, "ssRead needs one float per probe") ||
        check_len(&c, e2, b, s.steady_state_n, "ssRead needs one float per probe"))
        return NULL;
    uint32_t cur = 0;
    int rec = 0;
    if (!status(&c, ctx, efs_ss_read(ctx, e1, e2, &cur, &rec))) return NULL;
    napi_value arr;
    napi_create_array_with_length(env, 2, &arr);
    napi_set_element(env, arr, 0, num(env, cur));
    napi_set_element(env, arr, 1, num(env, rec));
    return arr;
}

static napi_value js_iterate(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    uint32_t n = u32_arg(&c, 1);
    BAIL();
    return status(&c, ctx, efs_iterate(ctx, n));
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    return status(&c, ctx, efs_sync(ctx));
}

/* tuneSweeps(h, reps) -> the chosen factor on the planner's workgroup target */
static napi_value js_tune_sweeps(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    uint32_t reps = u32_arg(&c, 1);
    BAIL();
    double scale = 1.0;
    if (!status(&c, ctx, efs_tune_sweeps(ctx, reps, &scale))) return NULL;
    return num(env, scale);
}

static napi_value js_num_ts(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    return num(env, efs_num_ts(ctx));
}

static napi_value js_set_num_ts(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    uint32_t v = u32_arg(&c, 1);
    BAIL();
    return status(&c, ctx, efs_set_num_ts(ctx, v));
}

/* getFields(h, voltOut, currOut): either may be null. */
static napi_value js_get_fields(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b;
    float *v = F32OPT(1, a), *cu = F32OPT(2, b);
    BAIL();
    efs_sizes s;
    if (sizes_of(&c, ctx, &s)) return NULL;
    const size_t n3 = 3 * (size_t)s.nx * s.ny * s.nz;
    if (check_len(&c, v, a, n3, "getFields: volt must hold 3*Nx*Ny*Nz floats") ||
        check_len(&c, cu, b, n3, "getFields: curr must hold 3*Nx*Ny*Nz floats"))
        return NULL;
    return status(&c, ctx, efs_get_fields(ctx, v, cu));
}

static napi_value js_upload_fields(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b;
    float *v = F32OPT(1, a), *cu = F32OPT(2, b);
    BAIL();
    efs_sizes s;
    if (sizes_of(&c, ctx, &s)) return NULL;
    const size_t n3 = 3 * (size_t)s.nx * s.ny * s.nz;
    if (check_len(&c, v, a, n3, "uploadFields: volt must hold 3*Nx*Ny*Nz floats") ||
        check_len(&c, cu, b, n3, "uploadFields: curr must hold 3*Nx*Ny*Nz floats"))
        return NULL;
    return status(&c, ctx, efs_upload_fields(ctx, v, cu));
}

static napi_value js_cfg_gather(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t n;
    uint32_t *idx = U32(1, n);
    BAIL();
    return status(&c, ctx, efs_cfg_gather(ctx, idx, (uint32_t)n));
}

static napi_value js_read_gather(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t n;
    float *out = F32(1, n);
    BAIL();
    efs_sizes s;
    if (sizes_of(&c, ctx, &s) || check_len(&c, out, n, 2 * (size_t)s.gather_n, "readGather: out must hold 2 floats per index"))
        return NULL;
    return status(&c, ctx, efs_read_gather(ctx, out));
}

static napi_value js_cfg_gather_buffered(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t n;
    uint32_t *idx = U32(1, n);
    uint32_t mx = u32_arg(&c, 2), iv = u32_arg(&c, 3);
    BAIL();
    return status(&c, ctx, efs_cfg_gather_buffered(ctx, idx, (uint32_t)n, mx, iv));
}

/* readGatherBuffered(h, out|null) -> number of samples */
static napi_value js_read_gather_buffered(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t n;
    float *out = F32OPT(1, n);
    BAIL();
    uint32_t ns = 0;
    efs_sizes s;
    if (sizes_of(&c, ctx, &s)) return NULL;
    if (out) {
        if (!status(&c, ctx, efs_read_gather_buffered(ctx, NULL, &ns))) return NULL;
        if (n < (size_t)ns * 2 * s.gather_buffered_n) {
            char buf[200];
            snprintf(buf, sizeof buf, "efs: readGatherBuffered: out holds %zu floats, %zu needed", n,
                     (size_t)ns * 2 * s.gather_buffered_n);
            throwf(&c, buf);
            return NULL;
        }
    }
    if (!status(&c, ctx, efs_read_gather_buffered(ctx, out, &ns))) return NULL;
    return num(env, ns);
}

/* cfgNF2FF(h, points, omega(F64), dT, edgeLens, elLen, maxTS, windowType, kBlock[, compensated = 1]) */
static napi_value js_cfg_nf2ff(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t np_, nf, ne, nl;
    uint32_t *pts = U32(1, np_);
    double *om = F64(2, nf);
    double dT = f64_arg(&c, 3);
    float *el = F32(4, ne);
    uint32_t *ell = U32(5, nl);
    uint32_t maxTS = u32_arg(&c, 6), win = u32_arg(&c, 7), kb = u32_arg(&c, 8);
    const uint32_t comp = c.argc > 9 ? u32_arg(&c, 9) : 1;
    BAIL();
    if (!status(&c, ctx, efs_set_nf2ff_compensated(ctx, comp ? 1 : 0))) return NULL;
    if (np_ % 3 || nl != 6) {
        throwf(&c, "efs: NF2FF points must be triplets and elLen must hold 6 lengths");
        return NULL;
    }
    size_t sum = 0;
    for (int k = 0; k < 6; ++k) sum += ell[k];
    if (sum != ne) {
        throwf(&c, "efs: NF2FF edge-length table does not match elLen");
        return NULL;
    }
    return status(&c, ctx, efs_cfg_nf2ff(ctx, pts, (uint32_t)(np_ / 3), om, (uint32_t)nf, dT, el, ell, maxTS, win, kb));
}

static napi_value js_read_nf2ff(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t a, b;
    float *e = F32(1, a), *h = F32(2, b);
    BAIL();
    efs_sizes s;
    if (sizes_of(&c, ctx, &s)) return NULL;
    const size_t need = (size_t)s.nf2ff_freqs * 6 * s.nf2ff_points;
    if (check_len(&c, e, a, need, "readNF2FF: accumE must hold F*6*P floats") ||
        check_len(&c, h, b, need, "readNF2FF: accumH must hold F*6*P floats"))
        return NULL;
    return status(&c, ctx, efs_read_nf2ff(ctx, e, h));
}

/* farfield(h, fIdx, meta, theta, phi, center(F64[3]), freq, radius, out
 *          [, mirrorDir, mirrorPos, mirrorType, imageNormals]) */
static napi_value js_farfield(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t lm, lt, lp, lc, lo;
    uint32_t f = u32_arg(&c, 1);
    float *meta = F32(2, lm), *th = F32(3, lt), *ph = F32(4, lp);
    double *ce = F64(5, lc);
    double freq = f64_arg(&c, 6), rad = f64_arg(&c, 7);
    float *out = F32(8, lo);
    const int mirror = c.argc > 9;
    const int mdir = mirror ? (int)f64_arg(&c, 9) : -1;
    const double mpos = mirror ? f64_arg(&c, 10) : 0.0;
    const int mtype = mirror ? (int)f64_arg(&c, 11) : 0;
    const int image_normals = mirror ? (int)f64_arg(&c, 12) : 0;
    BAIL();
    if (lc != 3 || lo != 5 * lt * lp) {
        throwf(&c, "efs: far-field center must hold 3 values and out 5 per angle");
        return NULL;
    }
    efs_sizes s;
    if (sizes_of(&c, ctx, &s) || check_len(&c, meta, lm, 8 * (size_t)s.nf2ff_points,
                                           "farfield: pointMeta must hold 8 floats per NF2FF point"))
        return NULL;
    if (!mirror)
        return status(&c, ctx, efs_farfield(ctx, f, meta, th, (uint32_t)lt, ph, (uint32_t)lp, ce, freq, rad, out));
    return status(&c, ctx, efs_farfield_mirror(ctx, f, meta, th, (uint32_t)lt, ph, (uint32_t)lp, ce, freq, rad,
                                               mdir, mpos, mtype, image_normals, out));
}

/* farfieldSurface(h, meta(F32, 8/point), accE(F64, 6/point), accH(F64),
 *                 theta, phi, center(F64[3]), freq, radius, out(F64, 5/angle)) */
static napi_value js_farfield_surface(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    size_t lm, le, lh, lt, lp, lc, lo;
    float *meta = F32(1, lm);
    double *ae = F64(2, le), *ah = F64(3, lh);
    float *th = F32(4, lt), *ph = F32(5, lp);
    double *ce = F64(6, lc);
    double freq = f64_arg(&c, 7), rad = f64_arg(&c, 8);
    double *out = F64(9, lo);
    BAIL();
    const size_t P = lm / 8;
    if (lm % 8 != 0 || le != 6 * P || lh != 6 * P || lc != 3 || lo != 5 * lt * lp) {
        throwf(&c, "efs: farfieldSurface needs 8 meta floats and 6 E and H values per point, a center of 3 and "
                   "5 outputs per angle");
        return NULL;
    }
    return status(&c, ctx, efs_farfield_surface(ctx, (uint32_t)P, meta, ae, ah, th, (uint32_t)lt, ph, (uint32_t)lp,
                                                ce, freq, rad, out));
}

static napi_value js_energy(napi_env env, napi_callback_info info) {
    BEGIN();
    CTX(0);
    double e = 0;
    if (!status(&c, ctx, efs_energy(ctx, &e))) return NULL;
    return num(env, e);
}


/* ------------------------------------------------------------ batched form
 * include/efs_batch.h: batchCreate(device, storage) -> handle | null,
 * batchAdd(h, nx, ny, nz, vv, vi, ii, iv) -> id, per-instance configure
 * calls, batchFinalize(h, interval, maxSamples), batchTune(h, steps) ->
 * [blocks, waves], batchIterate, reads. */
static efs_batch *batch_arg(call_t *c, size_t i) {
    void *p = NULL;
    if (napi_get_value_external(c->env, c->argv[i], &p) != napi_ok || !p) {
        throwf(c, "efs: invalid batch handle");
        return NULL;
    }
    return (efs_batch *)p;
}

static napi_value bstatus(call_t *c, efs_batch *b, int rc) {
    if (c->failed) return NULL;
    if (rc != EFS_OK) {
        throwf(c, efs_batch_last_error(b));
        return NULL;
    }
    return undef(c->env);
}

#define BATCH(i)                                 \
    efs_batch *b = batch_arg(&c, (i));           \
    if (!b) return NULL;

static napi_value js_batch_create(napi_env env, napi_callback_info info) {
    BEGIN();
    uint32_t dev = c.argc > 0 ? u32_arg(&c, 0) : 0;
    uint32_t storage = c.argc > 1 ? u32_arg(&c, 1) : 0;
    BAIL();
    efs_batch *b = NULL;
    int rc = efs_batch_create(&b, (int)dev);
    if (rc == EFS_ERR_NO_DEVICE) {
        napi_value n;
        napi_get_null(env, &n);
        return n;
    }
    if (rc != EFS_OK) {
        throwf(&c, efs_batch_last_error(NULL));
        return NULL;
    }
    rc = efs_batch_set_storage(b, (int)storage);
    if (rc != EFS_OK) {
        throwf(&c, efs_batch_last_error(b));
        efs_batch_destroy(b);
        return NULL;
    }
    napi_value h;
    napi_create_external(env, b, NULL, NULL, &h);
    return h;
}

static napi_value js_batch_destroy(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    efs_batch_destroy(b);
    return undef(env);
}

static napi_value js_batch_add(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    size_t a, bb, d, e;
    uint32_t nx = u32_arg(&c, 1), ny = u32_arg(&c, 2), nz = u32_arg(&c, 3);
    float *vv = F32(4, a), *vi = F32(5, bb), *ii = F32(6, d), *iv = F32(7, e);
    BAIL();
    size_t need = 3 * (size_t)nx * ny * nz;
    if (a != need || bb != need || d != need || e != need) {
        throwf(&c, "efs: coefficient arrays must hold 3*Nx*Ny*Nz floats");
        return NULL;
    }
    uint32_t id = 0;
    if (!bstatus(&c, b, efs_batch_add(b, nx, ny, nz, vv, vi, ii, iv, &id))) return NULL;
    return num(env, id);
}

/* batchCfgExcitation(h, id, signal, delay, amp, dir, pos, period) */
static napi_value js_batch_cfg_excitation(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    size_t ls, l1, l2, l3, l4;
    uint32_t id = u32_arg(&c, 1);
    float *sig = F32(2, ls);
    uint32_t *delay = U32(3, l1);
    float *amp = F32(4, l2);
    uint32_t *dir = U32(5, l3), *pos = U32(6, l4);
    uint32_t period = u32_arg(&c, 7);
    BAIL();
    if (l1 != l2 || l2 != l3 || l3 != l4) {
        throwf(&c, "efs: excitation arrays must have equal length");
        return NULL;
    }
    return bstatus(&c, b, efs_batch_cfg_excitation(b, id, sig, (uint32_t)ls, delay, amp, dir, pos, (uint32_t)l2,
                                                    period));
}

/* batchCfgMur(h, id, coeff, normalIdx, shiftedIdx) */
static napi_value js_batch_cfg_mur(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    size_t n1, n2, n3;
    uint32_t id = u32_arg(&c, 1);
    float *coeff = F32(2, n1);
    uint32_t *ni = U32(3, n2), *si = U32(4, n3);
    BAIL();
    if (n1 != n2 || n2 != n3) {
        throwf(&c, "efs: Mur arrays must have equal length");
        return NULL;
    }
    return bstatus(&c, b, efs_batch_cfg_mur(b, id, (uint32_t)n1, coeff, ni, si));
}

static napi_value js_batch_cfg_gather(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    size_t n;
    uint32_t id = u32_arg(&c, 1);
    uint32_t *idx = U32(2, n);
    BAIL();
    return bstatus(&c, b, efs_batch_cfg_gather(b, id, idx, (uint32_t)n));
}

static napi_value js_batch_finalize(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    uint32_t iv = u32_arg(&c, 1), mx = u32_arg(&c, 2);
    BAIL();
    return bstatus(&c, b, efs_batch_finalize(b, iv, mx));
}

static napi_value js_batch_tune(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    uint32_t steps = u32_arg(&c, 1);
    BAIL();
    uint32_t blocks = 0, waves = 0;
    if (!bstatus(&c, b, efs_batch_tune(b, steps, &blocks, &waves))) return NULL;
    napi_value arr;
    napi_create_array_with_length(env, 2, &arr);
    napi_set_element(env, arr, 0, num(env, blocks));
    napi_set_element(env, arr, 1, num(env, waves));
    return arr;
}

static napi_value js_batch_reset(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    return bstatus(&c, b, efs_batch_reset(b));
}

static napi_value js_batch_iterate(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    uint32_t n = u32_arg(&c, 1);
    BAIL();
    return bstatus(&c, b, efs_batch_iterate(b, n));
}

static napi_value js_batch_sync(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    return bstatus(&c, b, efs_batch_sync(b));
}

static napi_value js_batch_num_ts(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    uint32_t v = 0;
    if (!bstatus(&c, b, efs_batch_num_ts(b, &v))) return NULL;
    return num(env, v);
}

/* batchReadGather(h, id, out|null) -> number of samples */
static napi_value js_batch_read_gather(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    size_t n;
    uint32_t id = u32_arg(&c, 1);
    float *out = F32OPT(2, n);
    BAIL();
    uint32_t ns = 0, gn = 0;
    if (!bstatus(&c, b, efs_batch_sizes(b, id, NULL, &gn))) return NULL;
    if (out) {
        if (!bstatus(&c, b, efs_batch_read_gather(b, id, NULL, &ns))) return NULL;
        if (n < (size_t)ns * 2 * gn) {
            char buf[200];
            snprintf(buf, sizeof buf, "efs: batchReadGather: out holds %zu floats, %zu needed", n, (size_t)ns * 2 * gn);
            throwf(&c, buf);
            return NULL;
        }
    }
    if (!bstatus(&c, b, efs_batch_read_gather(b, id, out, &ns))) return NULL;
    return num(env, ns);
}

static napi_value js_batch_get_fields(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    size_t a, d;
    uint32_t id = u32_arg(&c, 1);
    float *v = F32(2, a), *cu = F32(3, d);
    BAIL();
    uint32_t dims[3] = {0, 0, 0};
    if (!bstatus(&c, b, efs_batch_sizes(b, id, dims, NULL))) return NULL;
    const size_t n3 = 3 * (size_t)dims[0] * dims[1] * dims[2];
    if (check_len(&c, v, a, n3, "batchGetFields: volt must hold 3*Nx*Ny*Nz floats") ||
        check_len(&c, cu, d, n3, "batchGetFields: curr must hold 3*Nx*Ny*Nz floats"))
        return NULL;
    return bstatus(&c, b, efs_batch_get_fields(b, id, v, cu));
}

/* batchInfo(h) -> [blocks, waves, graphSteps, cells] */
static napi_value js_batch_info(napi_env env, napi_callback_info info) {
    BEGIN();
    BATCH(0);
    uint32_t blocks = 0, waves = 0, gs = 0;
    uint64_t cells = 0;
    if (!bstatus(&c, b, efs_batch_info(b, &blocks, &waves, &gs, &cells))) return NULL;
    napi_value arr;
    napi_create_array_with_length(env, 4, &arr);
    napi_set_element(env, arr, 0, num(env, blocks));
    napi_set_element(env, arr, 1, num(env, waves));
    napi_set_element(env, arr, 2, num(env, gs));
    napi_set_element(env, arr, 3, num(env, (double)cells));
    return arr;
}

#define FN(name, fn) {name, 0, fn, 0, 0, 0, napi_default, 0}

static napi_value Init(napi_env env, napi_value exports) {
    napi_property_descriptor d[] = {
        FN("version", js_version),
        FN("deviceCount", js_device_count),
        FN("create", js_create),
        FN("destroy", js_destroy),
        FN("lastError", js_last_error),
        FN("init", js_init),
        FN("setStorage", js_set_storage),
        FN("fieldScales", js_field_scales),
        FN("cfgExcitation", js_cfg_excitation),
        FN("cfgPML", js_cfg_pml),
        FN("cfgMur", js_cfg_mur),
        FN("adeClear", js_ade_clear),
        FN("adeAdd", js_ade_add),
        FN("adeApply", js_ade_apply),
        FN("cfgTFSF", js_cfg_tfsf),
        FN("cfgRLC", js_cfg_rlc),
        FN("cfgPBC", js_cfg_pbc),
        FN("cfgSteadyState", js_cfg_ss),
        FN("ssSet", js_ss_set),
        FN("ssRead", js_ss_read),
        FN("iterate", js_iterate),
        FN("sync", js_sync),
        FN("tuneSweeps", js_tune_sweeps),
        FN("numTS", js_num_ts),
        FN("setNumTS", js_set_num_ts),
        FN("getFields", js_get_fields),
        FN("uploadFields", js_upload_fields),
        FN("cfgGather", js_cfg_gather),
        FN("readGather", js_read_gather),
        FN("cfgGatherBuffered", js_cfg_gather_buffered),
        FN("readGatherBuffered", js_read_gather_buffered),
        FN("cfgNF2FF", js_cfg_nf2ff),
        FN("readNF2FF", js_read_nf2ff),
        FN("farfield", js_farfield),
        FN("farfieldSurface", js_farfield_surface),
        FN("energy", js_energy),
        FN("batchCreate", js_batch_create),
        FN("batchDestroy", js_batch_destroy),
        FN("batchAdd", js_batch_add),
        FN("batchCfgExcitation", js_batch_cfg_excitation),
        FN("batchCfgMur", js_batch_cfg_mur),
        FN("batchCfgGather", js_batch_cfg_gather),
        FN("batchFinalize", js_batch_finalize),
        FN("batchTune", js_batch_tune),
        FN("batchReset", js_batch_reset),
        FN("batchIterate", js_batch_iterate),
        FN("batchSync", js_batch_sync),
        FN("batchNumTS", js_batch_num_ts),
        FN("batchReadGather", js_batch_read_gather),
        FN("batchGetFields", js_batch_get_fields),
        FN("batchInfo", js_batch_info),
    };
    napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
    return exports;
}

NAPI_MODULE(efs_engine, Init)
