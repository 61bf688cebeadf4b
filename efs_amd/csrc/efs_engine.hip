// efs_engine.hip -- MI355X FDTD engine: context, step driver and C ABI.
//
// Implements include/efs_engine.h.  One context = one HIP stream on one
// device; all buffers are device-resident in the reference's NIJK layout.
// The step order is the reference CPU engine's (src/webgpu-fdtd.mjs:
// 1094-1172) with the NF2FF/gather hooks of WebGPUEngine.iterate
// (src/webgpu-engine.mjs:2407-2504):
//
//   ADE(V) update, Mur pre, RLC pre           (read V; UPML cells via flux alias)
//   sweep_e  = UPML pre + core E + UPML post   (one pass)
//   TFSF(V), Mur post, [ADE(V) apply], excitation, Mur apply, RLC apply
//   ADE(I) update
//   sweep_h  = UPML pre + core H + UPML post   (one pass)
//   TFSF(I), [ADE(I) apply]
//   NF2FF sample (+ fold every K), buffered gather, numTS++
//
// Every index the kernels dereference is validated on the host at
// configure time; nothing reaches the GPU unchecked.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/efs_engine.h"
#include "hooks.hpp"
#include "nf2ff.hpp"
#include "reduce.hpp"
#include "sparse.hpp"
#include "sweep.hpp"
#include "sweep_kernels.hpp"

using namespace efs;

// Device scratch of the allocation-free far field (efs_farfield_noalloc):
// the FP64 partial integrals, 8 per (point chunk, angle).  One copy per GPU,
// shared by every context on it and held until the process ends.  In host
// code `scratch` only names the symbol: its device address comes from
// hipGetSymbolAddress, looked up once per device (efs_create, with the
// device current) and kept in g_scratch.
constexpr size_t kFarScratchDoubles = (size_t)4 << 17;  // 4 MiB
static __device__ __align__(256) double scratch[kFarScratchDoubles];

namespace {

std::string g_create_error;

constexpr int kMaxDevices = 64;
constexpr uint64_t kSideMaxCells = (uint64_t)4 << 20;  // EFS_SIDE_WORK=2: grids up to 4 Mi cells
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of this device's scratch (the device must be current).
hipError_t scratch_address(int device, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

struct Timed {
    int which;
    hipEvent_t a, b;
};

template <class T>
struct DevArr {
    T* p = nullptr;
    size_t n = 0;
};

}  // namespace

struct efs_ctx {
    uint32_t pt = 64;  // workgroup size of the per-step point kernels
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool inited = false;

    uint32_t nx = 0, ny = 0, nz = 0;
    size_t N = 0;      // cells per component (interface layout)
    // Device rows hold nzp >= nz cells (Nz rounded up to a multiple of 4, so
    // every grid runs the 4-cells-per-lane sweeps); padding cells have zero
    // coefficients and stay zero.  Np = Nx*Ny*nzp.
    uint32_t nzp = 0;
    size_t Np = 0;
    size_t S = 0;      // component stride of the device arrays (>= N, padded)
    bool vec4 = false; // 4 z cells per lane (Nz % 4 == 0)
    float* pool = nullptr;  // the six field/coefficient arrays, staggered
    // compressed coefficients: byte index per cell and component into a
    // table of <= 256 distinct (a, b) pairs per component (E: vv/vi, H: ii/iv)
    bool comp = false;
    uint8_t *cidx_e = nullptr, *cidx_h = nullptr;
    float2 *ctab_e = nullptr, *ctab_h = nullptr;
    // 16-bit form, when a component has more than 256 distinct pairs
    bool comp16 = false;
    uint16_t *cidx16_e = nullptr, *cidx16_h = nullptr;
    float2 *ctab16_e = nullptr, *ctab16_h = nullptr;
    uint32_t tabn_e = 0, tabn_h = 0;
    float *volt = nullptr, *curr = nullptr, *vv = nullptr, *vi = nullptr, *ii = nullptr, *iv = nullptr;
    uint32_t numTS = 0;
    SweepGeom geom{};   // 4-wide sweeps
    SweepGeom geom1{};  // 1-wide sweeps
    SweepBlock* sweep_tab4 = nullptr;  // their block tables
    SweepBlock* sweep_tab1 = nullptr;

    // excitation
    uint32_t exc_n = 0, exc_len = 0, exc_period = 0;
    float *exc_signal = nullptr, *exc_amp = nullptr;
    uint32_t *exc_delay = nullptr, *exc_gidx = nullptr;

    // UPML: regions with compact region-local arrays (6 coefficient arrays
    // and 2 flux arrays per region, 16-byte aligned region bases)
    PmlTable pml{};
    float *pml_coef = nullptr, *pml_vflux = nullptr, *pml_cflux = nullptr;
    bool pml_split = false;  // 4-wide sweeps with UPML (region z-bounds multiples of 4)
    bool pml_fused = true;   // one grid launch per sweep (else interior + region launches)
    bool pml_comp = false;   // UPML coefficients also as byte indices + tables (fused sweeps)
    // field storage: FP32, or FP16 with per-component power-of-two scales
    // (value = stored * fscale[k], k = volt 0..2, curr 0..2) kept on the
    // device and adjusted every range_interval steps by the range kernels
    int storage_req = EFS_STORAGE_F32;
    bool half = false;
    // FP64 storage (the precision reference run of the same engine): fields,
    // UPML flux, Mur state and the NF2FF buffer and accumulators in double,
    // updated in double by the 1-wide sweeps; coefficients stay the FP32
    // values given.  Arrays of R = double take twice the floats (rw()).
    bool dbl = false;
    float* fscale = nullptr;        // [6]
    unsigned int* f_amax = nullptr; // [6] running max |stored| (float bits)
    float* f_factor = nullptr;      // [6] last rescale factors
    float* f_smin = nullptr;        // [6] finest scale the range control may choose
    uint32_t range_interval = 32;
    bool fields_zero = true;        // no upload and no step since init: scales are free to choose
    uint64_t source_end = 0;        // first step index at which no excitation / TFSF sample is injected
    bool source_periodic = false;
    // regions as given (z-bounds) and, for regions widened to lane
    // boundaries, the padding cells (field device index, flux index)
    std::vector<uint32_t> pml_tz0, pml_tz1;
    uint32_t *pml_pad_g = nullptr, *pml_pad_f = nullptr;
    uint32_t pml_pad_n = 0;
    std::vector<uint64_t> h_pad;  // (device index << 32 | flux index) of the padding cells, sorted
    // padding cells that a hook writes after a sweep (V side: excitation,
    // Mur apply, TFSF-V, ADE write-back, RLC, PBC-V; I side: TFSF-I, ADE
    // write-back, PBC-I): their flux is set to the field again after the
    // hook phase (k_pml_pad_sync), rebuilt when a configuration changes
    bool pad_dirty = true;
    uint32_t *pad_hv_g = nullptr, *pad_hv_f = nullptr, *pad_hi_g = nullptr, *pad_hi_f = nullptr;
    uint32_t pad_hv_n = 0, pad_hi_n = 0;
    std::vector<uint32_t> h_tv_idx;    // TFSF voltage points (interface indices)
    std::vector<uint32_t> h_rlc_dev;   // RLC cells (device indices)
    uint8_t* pml_idx = nullptr;   // per region: E then H component indices (PmlRegion::pidx)
    float4* pml_tab_e = nullptr;  // [3][256] (vv, vvfo, vvfn, 0)
    float4* pml_tab_h = nullptr;  // [3][256] (ii, iifo, iifn, 0)

    // Mur
    uint32_t mur_n = 0;
    // entries kept: the last one of each normal cell, ordered by that cell
    std::vector<uint32_t> h_mur_normal, h_mur_shifted;  // interface indices
    std::vector<float> h_mur_coeff;
    float *mur_coeff = nullptr, *mur_saved = nullptr;
    uint32_t *mur_normal = nullptr, *mur_shifted = nullptr, *mur_normal_src = nullptr, *mur_shifted_src = nullptr;
    uint8_t* mur_early = nullptr;       // post done before the excitation (k_mur_post_some)
    uint32_t* mur_early_list = nullptr;
    uint32_t mur_ne = 0;
    // one-launch point work between the sweeps (k_post_e, k_post_h)
    bool fuse_hooks = true;
    bool post_e_ok = false;         // deferred lists built
    uint8_t *mur_defer = nullptr, *exc_defer = nullptr;
    uint2* mur_b = nullptr;         // deferred Mur entries: (normal cell, entry)
    uint32_t* exc_b = nullptr;
    uint32_t mur_nb = 0, exc_nb = 0;
    // group form of k_post_e (EFS_POST_GROUP, default 1): the early and
    // deferred entries done by its first grp_n workgroups, no k_post_e_b
    uint32_t grp_n = 0;
    uint32_t *grp_mur = nullptr, *grp_exc = nullptr, *grp_eoff = nullptr;
    bool mur_pre_ahead = false;     // the previous step's k_post_h ran this step's Mur pre pass
    // point work in extra workgroups of the sweep kernels (SideWork):
    // side_on from EFS_SIDE_WORK; side_e: the E-sweep side of
    // the previous step, pending while side_pend (it completes that step's
    // sample and gather)
    bool side_on = true;
    bool side_pend = false, side_pend_gather = false;
    SideWork side_e{};
    // E-sweep epilogue (sweep_epilogue, EFS_SWEEP_EPILOGUE): Mur and
    // excitation entries per block-table entry of the 4-wide launch
    std::vector<SweepBlock> h_tab4;
    double blocks_scale = 1.0;   // factor on the 4-wide launches' workgroup target (efs_tune_sweeps)
    uint64_t field_floats = 0;   // floats of the volt and curr arrays at the start of the pool
    uint64_t flux_floats = 0;    // floats of each UPML flux array
    bool ep_ok = false;
    uint32_t *ep_moff = nullptr, *ep_eoff = nullptr, *ep_exc = nullptr;
    uint4* ep_mur = nullptr;
    std::vector<uint32_t> h_exc_gidx;   // excitation cells (interface indices)

    // ADE (all orders/directions flattened)
    std::vector<uint32_t> h_ade_gidx;
    std::vector<uint8_t> h_ade_lor;
    std::vector<float> h_ade_c[6];  // v_int v_ext v_lor i_int i_ext i_lor
    bool ade_apply = false;
    bool ade_dirty = false;
    bool ade_i_live = false;  // some current-side ADE coefficient is non-zero
    std::vector<uint32_t> h_ti_idx;  // TFSF current points (interface indices)
    // point work of TFSF and ADE in the launches of other hooks
    // (plan_point_fusion, redone when a list changes)
    bool fuse_dirty = true;
    bool post_v_fuse = false;  // TFSF-V and the ADE write-back touch disjoint E cells
    bool post_i_fuse = false;  // no TFSF-I cell is read by the NF2FF sampling or the buffered gather
    bool ade_ahead = false;    // this step's E-side ADE update ran in the previous step's hooks launch
    bool ade_fuse = false;     // the E-side ADE update may run in the previous step's hooks launch
    std::vector<uint32_t> h_nf_pts;  // NF2FF points (ix, iy, iz), as configured
    std::vector<uint32_t> h_gb_idx;  // buffered gather cells (interface indices)
    bool cap_clash = false;          // an H-side hook writes inside the NF2FF capture rows
    bool cap_dirty = true;
    uint32_t ade_n = 0;
    uint32_t *ade_gidx = nullptr, *ade_src_v = nullptr, *ade_src_i = nullptr;
    uint8_t* ade_lor = nullptr;
    float* ade_c[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float *ade_va = nullptr, *ade_vl = nullptr, *ade_ia = nullptr, *ade_il = nullptr;

    bool alias_dirty = false;

    // TFSF
    uint32_t tfsf_len = 0, tfsf_period = 0, tfsf_nv = 0, tfsf_ni = 0;
    float* tfsf_signal = nullptr;
    uint32_t *tv_idx = nullptr, *tv_dint = nullptr, *ti_idx = nullptr, *ti_dint = nullptr;
    float *tv_dfrac = nullptr, *tv_amp = nullptr, *ti_dfrac = nullptr, *ti_amp = nullptr;

    // RLC
    uint32_t rlc_n = 0;
    uint32_t *rlc_gidx = nullptr, *rlc_type = nullptr;
    float *rlc_c = nullptr, *rlc_vdn = nullptr, *rlc_jn = nullptr, *rlc_vil = nullptr;

    // periodic boundaries: one entry per axis, in configuration order
    std::vector<PbcAxis> pbc_e, pbc_h;
    float* pbc_imag = nullptr;  // Bloch imaginary parts (4 face-sized buffers per phased axis)

    // steady-state detection
    uint32_t ss_n = 0, ss_period = 0, ss_current = 0;
    bool ss_recording = false;
    uint32_t* ss_gidx = nullptr;
    float* ss_e = nullptr;  // [2][ss_n]: period 1, period 2

    // probe gather
    uint32_t g_n = 0;
    uint32_t* g_idx = nullptr;
    float* g_out = nullptr;
    uint32_t gb_n = 0, gb_max = 0, gb_interval = 1, gb_count = 0;
    uint32_t* gb_idx = nullptr;
    float* gb_ring = nullptr;

    // NF2FF
    uint32_t nf_P = 0, nf_F = 0, nf_K = 0, nf_maxTS = 0, nf_window = 0;
    uint32_t nf_rows = 0;  // buffer / accumulator rows: E rows, then H rows from nf_geom.rowsE
    uint32_t nf_NC = 0;    // accumulator columns: 2F rounded up to kFoldCols
    uint32_t nf_slot = 0, nf_ts0 = 0;
    double nf_dT = 0, nf_wnorm = 1;
    Nf2ffGeom nf_geom{};
    uint32_t* nf_pts = nullptr;
    // NF2FF points are kept in grid order (sorted by linear cell index) so
    // neighbouring threads of the sampling kernel read neighbouring cells;
    // nf_perm[q] is the caller's index of internal point q.
    std::vector<uint32_t> nf_perm;
    // z planes of the surface captured by the H sweep (NfCapture); in use
    // when nf_capture() finds the step's configuration compatible
    NfCapture nf_cap{};
    float* nf_slab = nullptr;
    float *nf_el = nullptr, *nf_buf = nullptr, *nf_acc = nullptr;
    // compensated accumulation: rounding errors of acc (value = acc + cmp);
    // nullptr in the plain form
    float* nf_cmp = nullptr;
    bool nf_comp_req = true;
    double* nf_omega = nullptr;
    float* nf_tw = nullptr;  // [2][K][NC] twiddles

    // energy
    double* en_partial = nullptr;

    // this device's far-field scratch (efs_farfield_noalloc) and the NF2FF
    // point permutation on the device (internal point q -> caller's index)
    double* scratch = nullptr;
    hipError_t scratch_err = hipSuccess;
    uint32_t* nf_perm_d = nullptr;

    // profiling
    bool prof = false;
    std::vector<Timed> pending;
    std::vector<hipEvent_t> event_pool;
    double prof_ms[EFS_PROF_COUNT] = {0, 0, 0};
    uint64_t prof_n[EFS_PROF_COUNT] = {0, 0, 0};
};

namespace {

int fail(efs_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), EFS_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                              \
    } while (0)

#define CHECK_LAUNCH(ctx) HIP_TRY(ctx, hipGetLastError())

template <class T>
void dfree(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

template <class T>
int dalloc(efs_ctx* c, T*& p, size_t n) {
    dfree(p);
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        return fail(c, e == hipErrorOutOfMemory ? EFS_ERR_NOMEM : EFS_ERR_HIP, "hipMalloc(%zu bytes): %s",
                    n * sizeof(T), hipGetErrorString(e));
    }
    return EFS_OK;
}

template <class T>
int dupload(efs_ctx* c, T*& p, const T* src, size_t n) {
    int rc = dalloc(c, p, n);
    if (rc) return rc;
    if (n) HIP_TRY(c, hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    return EFS_OK;
}

template <class T>
int dzero(efs_ctx* c, T*& p, size_t n) {
    int rc = dalloc(c, p, n);
    if (rc) return rc;
    HIP_TRY(c, hipMemset(p, 0, (n ? n : 1) * sizeof(T)));
    return EFS_OK;
}

#define TRY(expr)              \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

inline uint32_t nblk(size_t n, int t = kPointThreads) { return (uint32_t)((n + t - 1) / t); }

int need_init(efs_ctx* c) {
    if (!c) return EFS_ERR_STATE;
    if (!c->inited) return fail(c, EFS_ERR_STATE, "Engine not initialized. Call init() first.");
    return EFS_OK;
}

// Interface field index (component*N + linear position) -> device index
// (component*S + linear position).
inline uint32_t tr(const efs_ctx* c, uint64_t gidx) {
    const uint64_t n = gidx / c->N, r = gidx - n * c->N;
    if (c->nzp == c->nz) return (uint32_t)(n * c->S + r);
    const uint64_t xy = r / c->nz;
    return (uint32_t)(n * c->S + xy * c->nzp + (r - xy * c->nz));
}

// Distinct (a, b) pairs of one component: byte index per cell and a
// 256-entry table.  Returns false when there are more than 256 pairs.
bool pair_table(const float* a, const float* b, size_t n, uint8_t* idx, float2* tab) {
    constexpr uint32_t H = 4096;  // open-addressing slots (power of two)
    std::vector<uint64_t> key(H);
    std::vector<int16_t> val(H, -1);
    uint32_t count = 0;
    uint64_t last_k = ~0ull;
    uint8_t last_v = 0;
    for (size_t i = 0; i < n; ++i) {
        uint32_t ua, ub;
        std::memcpy(&ua, a + i, 4);
        std::memcpy(&ub, b + i, 4);
        const uint64_t k = ((uint64_t)ua << 32) | ub;
        if (k == last_k) {
            idx[i] = last_v;
            continue;
        }
        uint32_t h = (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> 52) & (H - 1);
        while (val[h] >= 0 && key[h] != k) h = (h + 1) & (H - 1);
        if (val[h] < 0) {
            if (count == 256) return false;
            key[h] = k;
            val[h] = (int16_t)count;
            tab[count] = make_float2(a[i], b[i]);
            ++count;
        }
        last_k = k;
        last_v = (uint8_t)val[h];
        idx[i] = last_v;
    }
    for (uint32_t t = count; t < 256; ++t) tab[t] = make_float2(0.f, 0.f);
    return true;
}

// The same for at most 65,536 pairs: 16-bit indices, the distinct pairs in
// first-seen order in tab.
bool pair_table16(const float* a, const float* b, size_t n, uint16_t* idx, std::vector<float2>& tab) {
    constexpr uint32_t H = 1u << 17;  // open-addressing slots (load <= 1/2)
    std::vector<uint64_t> key(H);
    std::vector<int32_t> val(H, -1);
    tab.clear();
    uint64_t last_k = ~0ull;
    uint16_t last_v = 0;
    for (size_t i = 0; i < n; ++i) {
        uint32_t ua, ub;
        std::memcpy(&ua, a + i, 4);
        std::memcpy(&ub, b + i, 4);
        const uint64_t k = ((uint64_t)ua << 32) | ub;
        if (k == last_k) {
            idx[i] = last_v;
            continue;
        }
        uint32_t h = (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> 47) & (H - 1);
        while (val[h] >= 0 && key[h] != k) h = (h + 1) & (H - 1);
        if (val[h] < 0) {
            if (tab.size() == 65536) return false;
            key[h] = k;
            val[h] = (int32_t)tab.size();
            tab.push_back(make_float2(a[i], b[i]));
        }
        last_k = k;
        last_v = (uint16_t)val[h];
        idx[i] = last_v;
    }
    return true;
}

// Distinct (a, b, c) triples of one component, accumulated over several
// arrays (the UPML regions): byte index per cell into a 256-entry table of
// (a, b, c, 0).  add() returns false once there are more than 256 triples.
struct TripleTable {
    static constexpr uint32_t H = 4096;  // open-addressing slots (power of two)
    std::vector<uint64_t> k0 = std::vector<uint64_t>(H);
    std::vector<uint32_t> k1 = std::vector<uint32_t>(H);
    std::vector<int16_t> val = std::vector<int16_t>(H, -1);
    uint32_t count = 0;
    float4 tab[256] = {};

    bool add(const float* a, const float* b, const float* c, size_t n, uint8_t* idx) {
        uint64_t last0 = ~0ull;
        uint32_t last1 = ~0u;
        uint8_t last_v = 0;
        for (size_t i = 0; i < n; ++i) {
            uint32_t ua, ub, uc;
            std::memcpy(&ua, a + i, 4);
            std::memcpy(&ub, b + i, 4);
            std::memcpy(&uc, c + i, 4);
            const uint64_t q0 = ((uint64_t)ua << 32) | ub;
            if (q0 == last0 && uc == last1) {
                idx[i] = last_v;
                continue;
            }
            uint32_t h = (uint32_t)(((q0 ^ ((uint64_t)uc * 0xC2B2AE3D27D4EB4Full)) * 0x9E3779B97F4A7C15ull) >> 52) & (H - 1);
            while (val[h] >= 0 && (k0[h] != q0 || k1[h] != uc)) h = (h + 1) & (H - 1);
            if (val[h] < 0) {
                if (count == 256) return false;
                k0[h] = q0;
                k1[h] = uc;
                val[h] = (int16_t)count;
                tab[count] = make_float4(a[i], b[i], c[i], 0.f);
                ++count;
            }
            last0 = q0;
            last1 = uc;
            last_v = (uint8_t)val[h];
            idx[i] = last_v;
        }
        return true;
    }
};

std::vector<uint32_t> tr_all(const efs_ctx* c, const uint32_t* gidx, size_t n) {
    std::vector<uint32_t> out(n);
    for (size_t k = 0; k < n; ++k) out[k] = tr(c, gidx[k]);
    return out;
}

uint64_t env_u64(const char* name, uint64_t dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    return std::strtoull(v, nullptr, 10);
}

// ---- widened UPML regions: padding cells' flux follows the field ----
FieldView fv(const efs_ctx* c, int curr);

int pml_pad_sync(efs_ctx* c) {
    if (!c->pml_pad_n) return EFS_OK;
    hipLaunchKernelGGL(k_pml_pad_sync, dim3(nblk(c->pml_pad_n)), dim3(kPointThreads), 0, c->stream, fv(c, 0),
                       c->pml_vflux, c->pml_pad_g, c->pml_pad_f, c->pml_pad_n);
    hipLaunchKernelGGL(k_pml_pad_sync, dim3(nblk(c->pml_pad_n)), dim3(kPointThreads), 0, c->stream, fv(c, 1),
                       c->pml_cflux, c->pml_pad_g, c->pml_pad_f, c->pml_pad_n);
    CHECK_LAUNCH(c);
    return EFS_OK;
}

// The padding cells among the cells the hooks write after each sweep.
int rebuild_pad_hooks(efs_ctx* c) {
    c->pad_dirty = false;
    c->pad_hv_n = c->pad_hi_n = 0;
    if (!c->pml_pad_n) return EFS_OK;
    const bool pbc = !c->pbc_e.empty() || !c->pbc_h.empty();  // whole face planes: every padding cell
    std::vector<uint32_t> v, i;
    for (uint32_t g : c->h_exc_gidx) v.push_back(tr(c, g));
    for (uint32_t g : c->h_mur_normal) v.push_back(tr(c, g));
    for (uint32_t g : c->h_tv_idx) v.push_back(tr(c, g));
    v.insert(v.end(), c->h_rlc_dev.begin(), c->h_rlc_dev.end());
    if (c->ade_apply)
        for (uint32_t g : c->h_ade_gidx) {
            v.push_back(tr(c, g));
            i.push_back(tr(c, g));
        }
    for (uint32_t g : c->h_ti_idx) i.push_back(tr(c, g));
    auto pick = [&](std::vector<uint32_t>& dev, uint32_t*& dg, uint32_t*& df, uint32_t& n) -> int {
        std::vector<uint32_t> pg, pf;
        if (pbc) {
            for (uint64_t e : c->h_pad) {
                pg.push_back((uint32_t)(e >> 32));
                pf.push_back((uint32_t)e);
            }
        } else {
            std::sort(dev.begin(), dev.end());
            dev.erase(std::unique(dev.begin(), dev.end()), dev.end());
            for (uint32_t d : dev) {
                auto it = std::lower_bound(c->h_pad.begin(), c->h_pad.end(), (uint64_t)d << 32);
                if (it != c->h_pad.end() && (uint32_t)(*it >> 32) == d) {
                    pg.push_back(d);
                    pf.push_back((uint32_t)*it);
                }
            }
        }
        n = (uint32_t)pg.size();
        if (n) {
            TRY(dupload(c, dg, pg.data(), n));
            TRY(dupload(c, df, pf.data(), n));
        }
        return EFS_OK;
    };
    TRY(pick(v, c->pad_hv_g, c->pad_hv_f, c->pad_hv_n));
    TRY(pick(i, c->pad_hi_g, c->pad_hi_f, c->pad_hi_n));
    return EFS_OK;
}

// ---- alias resolution: field index -> UPML flux index when in a region ----
uint32_t alias_of(const efs_ctx* c, uint32_t gidx) {
    const uint32_t n = (uint32_t)(gidx / c->N);
    const size_t lin = gidx - (size_t)n * c->N;
    const uint32_t x = (uint32_t)(lin / ((size_t)c->ny * c->nz));
    const uint32_t rem = (uint32_t)(lin - (size_t)x * c->ny * c->nz);
    const uint32_t y = rem / c->nz, z = rem - y * c->nz;
    for (uint32_t k = 0; k < c->pml.n; ++k) {
        const PmlRegion& R = c->pml.r[k];
        // the region as given (a widened region's padding cells read the field)
        if (x >= R.x0 && x < R.x1 && y >= R.y0 && y < R.y1 && z >= c->pml_tz0[k] && z < c->pml_tz1[k]) {
            const uint64_t p = R.flux + (uint64_t)n * R.cs + ((uint64_t)(x - R.x0) * R.ny + (y - R.y0)) * R.nz +
                               (z - R.z0);
            return kAliasFlux | (uint32_t)p;
        }
    }
    return tr(c, gidx);
}

// Lists of k_post_e / k_post_e_b.  Deferred cells: the shifted cells of
// early entries (phase A reads them).  A Mur entry whose normal cell is
// deferred applies in phase B; an excitation entry on a deferred cell adds
// in phase B, one on a Mur normal cell is dropped (the apply overwrites
// it).  Without ADE write-back, "early" is exactly: shifted cell = some
// normal or excitation cell.  Group form: the early and the deferred Mur
// entries and the deferred excitation entries go to k_post_e's first
// workgroup instead (phase A skips them all).
int plan_post_e(efs_ctx* c, const std::vector<uint8_t>& early) {
    c->post_e_ok = false;
    if (!c->fuse_hooks) return EFS_OK;
    std::vector<uint32_t> sh_early, nor(c->h_mur_normal);
    for (uint32_t k = 0; k < c->mur_n; ++k)
        if (early[k]) sh_early.push_back(c->h_mur_shifted[k]);
    std::sort(sh_early.begin(), sh_early.end());
    std::sort(nor.begin(), nor.end());
    auto in = [](const std::vector<uint32_t>& v, uint32_t x) { return std::binary_search(v.begin(), v.end(), x); };
    std::vector<uint8_t> mdef(c->mur_n, 0), edef(std::max<uint32_t>(c->exc_n, 1), 0);
    std::vector<uint2> mb;
    std::vector<uint32_t> eb;
    for (uint32_t k = 0; k < c->mur_n; ++k)
        if (in(sh_early, c->h_mur_normal[k])) {
            mdef[k] = 1;
            mb.push_back(make_uint2(tr(c, c->h_mur_normal[k]), k));
        }
    for (uint32_t k = 0; k < c->exc_n; ++k) {
        const bool on_normal = in(nor, c->h_exc_gidx[k]);
        if (on_normal || in(sh_early, c->h_exc_gidx[k])) {
            edef[k] = 1;
            if (!on_normal) eb.push_back(k);
        }
    }
    c->grp_n = 0;
    if (env_u64("EFS_POST_GROUP", 1) != 0 && !(mb.empty() && eb.empty())) {
        // linked sets: an early entry with the Mur entry and the deferred
        // excitation entries that write its shifted cell (nodes: Mur
        // entries, then excitation entries)
        const uint32_t M = c->mur_n, E = c->exc_n;
        std::vector<uint32_t> parent(M + E);
        for (uint32_t a = 0; a < M + E; ++a) parent[a] = a;
        auto root = [&](uint32_t a) {
            while (parent[a] != a) a = parent[a] = parent[parent[a]];
            return a;
        };
        std::vector<std::pair<uint32_t, uint32_t>> by_normal(M), by_exc;
        for (uint32_t k = 0; k < M; ++k) by_normal[k] = {c->h_mur_normal[k], k};
        std::sort(by_normal.begin(), by_normal.end());
        for (uint32_t k : eb) by_exc.push_back({c->h_exc_gidx[k], k});
        std::sort(by_exc.begin(), by_exc.end());
        std::vector<uint8_t> in_grp(M + E, 0);
        for (uint32_t k = 0; k < M; ++k) {
            if (early[k] || mdef[k]) in_grp[k] = 1;
            if (!early[k]) continue;
            const uint32_t cell = c->h_mur_shifted[k];
            auto m = std::lower_bound(by_normal.begin(), by_normal.end(), std::make_pair(cell, 0u));
            if (m != by_normal.end() && m->first == cell) parent[root(m->second)] = root(k);
            for (auto e = std::lower_bound(by_exc.begin(), by_exc.end(), std::make_pair(cell, 0u));
                 e != by_exc.end() && e->first == cell; ++e)
                parent[root(M + e->second)] = root(k);
        }
        for (uint32_t k : eb) in_grp[M + k] = 1;
        std::vector<std::vector<uint32_t>> sets(M + E);
        for (uint32_t a = 0; a < M + E; ++a)
            if (in_grp[a]) sets[root(a)].push_back(a);
        // pack whole sets into groups of at most kPostThreads Mur entries
        std::vector<uint32_t> gm, ge, eoff{0};
        uint32_t fill = 0;  // Mur entries of the open group
        bool open = false, ok = true;
        auto close = [&] {
            gm.resize(gm.size() + (kPostThreads - fill), ~0u);
            eoff.push_back((uint32_t)ge.size());
        };
        for (const auto& st : sets) {
            if (st.empty()) continue;
            uint32_t nm = 0;
            for (uint32_t a : st) nm += a < M;
            if (nm > (uint32_t)kPostThreads) {
                ok = false;
                break;
            }
            if (!open || fill + nm > (uint32_t)kPostThreads) {
                if (open) close();
                open = true;
                fill = 0;
            }
            for (uint32_t a : st) {
                if (a < M) gm.push_back(a);
                else ge.push_back(a - M);
            }
            fill += nm;
        }
        if (ok && open) {
            close();
            c->grp_n = (uint32_t)(gm.size() / kPostThreads);
            for (uint32_t k = 0; k < M; ++k)
                if (in_grp[k]) mdef[k] = 1;
            TRY(dupload(c, c->grp_mur, gm.data(), gm.size()));
            if (!ge.empty()) TRY(dupload(c, c->grp_exc, ge.data(), ge.size()));
            TRY(dupload(c, c->grp_eoff, eoff.data(), eoff.size()));
            mb.clear();
            eb.clear();
        }
    }
    TRY(dupload(c, c->mur_defer, mdef.data(), mdef.size()));
    TRY(dupload(c, c->exc_defer, edef.data(), edef.size()));
    if (!mb.empty()) TRY(dupload(c, c->mur_b, mb.data(), mb.size()));
    if (!eb.empty()) TRY(dupload(c, c->exc_b, eb.data(), eb.size()));
    c->mur_nb = (uint32_t)mb.size();
    c->exc_nb = (uint32_t)eb.size();
    c->post_e_ok = true;
    return EFS_OK;
}

// Lists of sweep_epilogue: each Mur entry goes to the block-table entry
// that holds both its cells, each excitation entry to the one that holds
// its cell (one on a Mur normal cell is dropped: the apply overwrites it).
// Off (k_post_e runs) when some entry's cells lie in two blocks or a block
// gets more Mur entries than its threads hold.
int plan_epilogue(efs_ctx* c) {
    c->ep_ok = false;
    if (env_u64("EFS_SWEEP_EPILOGUE", 1) == 0 || !c->side_on || !c->fuse_hooks || !c->vec4 || c->dbl ||
        (c->pml.n && !(c->pml_split && c->pml_fused)) || (!c->mur_n && !c->exc_n))
        return EFS_OK;
    const uint32_t nx = c->nx, ny = c->ny, nz = c->nz, nzp = c->nzp, tl = c->geom.tile_lanes;
    const auto& tab = c->h_tab4;
    const uint32_t T = (uint32_t)(((uint64_t)ny * nzp / 4 + tl - 1) / tl);
    // [tile][x] -> the table entry of the block holding the tile's lanes
    // there; a block that takes part of its tile is listed in sub[tile][x]
    // with its first lane
    std::vector<int32_t> blk((size_t)T * nx, -1);
    std::vector<std::vector<std::pair<uint32_t, int32_t>>> sub((size_t)T * nx);
    for (size_t e = 0; e < tab.size(); ++e)
        for (uint32_t x = tab[e].x0; x < tab[e].x1; ++x) {
            const size_t i = (size_t)tab[e].tile * nx + x;
            if (block_lane_count(tab[e].flags) == 0) blk[i] = (int32_t)e;
            else sub[i].push_back({block_lane_first(tab[e].flags), (int32_t)e});
        }
    for (auto& s : sub) std::sort(s.begin(), s.end());
    auto block_of = [&](uint64_t gidx) -> int32_t {
        const uint64_t lin = gidx % c->N;
        const uint32_t x = (uint32_t)(lin / ((uint64_t)ny * nz));
        const uint32_t rem = (uint32_t)(lin - (uint64_t)x * ny * nz);
        const uint32_t y = rem / nz, z = rem - y * nz;
        const uint64_t lane = ((uint64_t)y * nzp + z) / 4;
        const size_t i = (size_t)(lane / tl) * nx + x;
        int32_t b = blk[i];
        for (const auto& s : sub[i])
            if (lane % tl >= s.first) b = s.second;
        return b;
    };
    std::vector<std::vector<uint32_t>> mur(tab.size()), exc(tab.size());
    for (uint32_t k = 0; k < c->mur_n; ++k) {
        const int32_t b = block_of(c->h_mur_normal[k]);
        if (b < 0 || b != block_of(c->h_mur_shifted[k])) return EFS_OK;
        mur[b].push_back(k);
    }
    std::vector<uint32_t> nor(c->h_mur_normal);
    std::sort(nor.begin(), nor.end());
    for (uint32_t k = 0; k < c->exc_n; ++k) {
        if (std::binary_search(nor.begin(), nor.end(), c->h_exc_gidx[k])) continue;
        const int32_t b = block_of(c->h_exc_gidx[k]);
        if (b < 0) return EFS_OK;
        exc[b].push_back(k);
    }
    std::vector<uint32_t> moff{0}, eoff{0}, elist;
    std::vector<uint4> mlist;
    for (size_t b = 0; b < tab.size(); ++b) {
        if (mur[b].size() > (size_t)kEpiPerThread * kSideThreads) return EFS_OK;
        for (uint32_t k : mur[b]) {
            uint32_t bits;
            std::memcpy(&bits, &c->h_mur_coeff[k], sizeof bits);
            // a shifted cell in the block's first two planes is read from
            // the LDS copy the march keeps of them
            const uint64_t sg = c->h_mur_shifted[k], lin = sg % c->N;
            const uint32_t comp = (uint32_t)(sg / c->N);
            const uint32_t x = (uint32_t)(lin / ((uint64_t)ny * nz));
            const uint32_t rem = (uint32_t)(lin - (uint64_t)x * ny * nz);
            const uint32_t y = rem / nz, z = rem - y * nz;
            const uint64_t lane = ((uint64_t)y * nzp + z) / 4;
            uint32_t src = tr(c, sg);
            if (x - tab[b].x0 < kEpLdsPlanes)
                src = kEpInLds | (((x - tab[b].x0) * 3 + comp) * kEpLdsCells +
                                  4 * (uint32_t)(lane - (uint64_t)tab[b].tile * tl - block_lane_first(tab[b].flags)) +
                                  z % 4);
            mlist.push_back(make_uint4(src, tr(c, c->h_mur_normal[k]), k, bits));
        }
        elist.insert(elist.end(), exc[b].begin(), exc[b].end());
        moff.push_back((uint32_t)mlist.size());
        eoff.push_back((uint32_t)elist.size());
    }
    TRY(dupload(c, c->ep_moff, moff.data(), moff.size()));
    TRY(dupload(c, c->ep_eoff, eoff.data(), eoff.size()));
    if (!mlist.empty()) TRY(dupload(c, c->ep_mur, mlist.data(), mlist.size()));
    if (!elist.empty()) TRY(dupload(c, c->ep_exc, elist.data(), elist.size()));
    c->ep_ok = true;
    return EFS_OK;
}

int rebuild_aliases(efs_ctx* c) {
    if (c->mur_n) {
        std::vector<uint32_t> ns(c->mur_n), ss(c->mur_n);
        for (uint32_t k = 0; k < c->mur_n; ++k) {
            ns[k] = alias_of(c, c->h_mur_normal[k]);
            ss[k] = alias_of(c, c->h_mur_shifted[k]);
        }
        TRY(dupload(c, c->mur_normal_src, ns.data(), ns.size()));
        TRY(dupload(c, c->mur_shifted_src, ss.data(), ss.size()));
        // entries whose shifted cell is written between post and apply
        std::vector<uint32_t> written(c->h_mur_normal);
        written.insert(written.end(), c->h_exc_gidx.begin(), c->h_exc_gidx.end());
        if (c->ade_apply) written.insert(written.end(), c->h_ade_gidx.begin(), c->h_ade_gidx.end());
        std::sort(written.begin(), written.end());
        std::vector<uint8_t> early(c->mur_n, 0);
        std::vector<uint32_t> list;
        for (uint32_t k = 0; k < c->mur_n; ++k)
            if (std::binary_search(written.begin(), written.end(), c->h_mur_shifted[k])) {
                early[k] = 1;
                list.push_back(k);
            }
        TRY(dupload(c, c->mur_early, early.data(), early.size()));
        TRY(dupload(c, c->mur_early_list, list.data(), list.size()));
        c->mur_ne = (uint32_t)list.size();
        TRY(plan_post_e(c, early));
    }
    TRY(plan_epilogue(c));
    if (c->ade_n) {
        std::vector<uint32_t> sv(c->ade_n);
        for (uint32_t k = 0; k < c->ade_n; ++k) sv[k] = alias_of(c, c->h_ade_gidx[k]);
        TRY(dupload(c, c->ade_src_v, sv.data(), sv.size()));
        TRY(dupload(c, c->ade_src_i, sv.data(), sv.size()));  // same offsets into the current flux
    }
    c->alias_dirty = false;
    return EFS_OK;
}

inline size_t rw(const efs_ctx* c);

int rebuild_ade(efs_ctx* c) {
    c->ade_n = (uint32_t)c->h_ade_gidx.size();
    if (c->ade_n) {
        const std::vector<uint32_t> g = tr_all(c, c->h_ade_gidx.data(), c->ade_n);
        TRY(dupload(c, c->ade_gidx, g.data(), c->ade_n));
        TRY(dupload(c, c->ade_lor, c->h_ade_lor.data(), c->ade_n));
        for (int k = 0; k < 6; ++k) TRY(dupload(c, c->ade_c[k], c->h_ade_c[k].data(), c->ade_n));
        // all-zero current-side coefficients keep the state at 0 (A = 0 A + 0 V)
        // and the write-back subtracts 0: both passes are skipped
        c->ade_i_live = false;
        for (int k = 3; k < 6; ++k)
            for (float v : c->h_ade_c[k]) c->ade_i_live |= v != 0.0f;
        // the state in the arithmetic type (double in the FP64 storage form)
        TRY(dzero(c, c->ade_va, c->ade_n * rw(c)));
        TRY(dzero(c, c->ade_vl, c->ade_n * rw(c)));
        TRY(dzero(c, c->ade_ia, c->ade_n * rw(c)));
        TRY(dzero(c, c->ade_il, c->ade_n * rw(c)));
    }
    c->ade_dirty = false;
    c->alias_dirty = true;
    return EFS_OK;
}

// ---- profiling ----
hipEvent_t take_event(efs_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct ProfScope {
    efs_ctx* c;
    int which;
    hipEvent_t a = nullptr;
    ProfScope(efs_ctx* ctx, int w) : c(ctx), which(w) {
        if (c->prof) {
            a = take_event(c);
            if (a) hipEventRecord(a, c->stream);
        }
    }
    ~ProfScope() {
        if (c->prof && a) {
            hipEvent_t b = take_event(c);
            if (b) {
                hipEventRecord(b, c->stream);
                c->pending.push_back({which, a, b});
            } else {
                c->event_pool.push_back(a);
            }
        }
    }
};

int drain_profile(efs_ctx* c) {
    if (c->pending.empty()) return EFS_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto& t : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            c->prof_ms[t.which] += ms;
            c->prof_n[t.which] += 1;
        }
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->pending.clear();
    return EFS_OK;
}

// ---- NF2FF fold ----
int nf2ff_fold(efs_ctx* c) {
    if (!c->nf_P || c->nf_slot == 0) return EFS_OK;
    const uint32_t kused = c->nf_slot;
    const uint32_t nt = 2 * kused * c->nf_NC;
    if (c->dbl) {  // FP64 storage: FP64 twiddles, buffer and accumulators
        hipLaunchKernelGGL(k_nf2ff_twiddle<double>, dim3(nblk(nt)), dim3(kPointThreads), 0, c->stream,
                           (double*)c->nf_tw, c->nf_K, kused, c->nf_F, c->nf_NC, c->nf_ts0, c->nf_omega, c->nf_dT,
                           c->nf_maxTS, c->nf_window, c->nf_wnorm);
        const dim3 grid(c->nf_rows / kFold64Rows, c->nf_NC / kFoldCols);
        hipLaunchKernelGGL(k_nf2ff_fold64, grid, dim3(kFold64Rows), (size_t)kused * kFoldCols * sizeof(double),
                           c->stream, (const double*)c->nf_buf, c->nf_rows, c->nf_geom.rowsE, c->nf_K, kused,
                           (const double*)c->nf_tw, c->nf_NC, (double*)c->nf_acc);
        CHECK_LAUNCH(c);
        c->nf_slot = 0;
        return EFS_OK;
    }
    hipLaunchKernelGGL(k_nf2ff_twiddle<float>, dim3(nblk(nt)), dim3(kPointThreads), 0, c->stream, c->nf_tw, c->nf_K,
                       kused, c->nf_F, c->nf_NC, c->nf_ts0, c->nf_omega, c->nf_dT, c->nf_maxTS, c->nf_window,
                       c->nf_wnorm);
    CHECK_LAUNCH(c);
    const size_t lds = (size_t)((kused + 1) & ~1u) * kFoldCols * sizeof(float);
    const dim3 grid(c->nf_rows / kFoldRows, c->nf_NC / kFoldCols);
    hipLaunchKernelGGL(k_nf2ff_fold, grid, dim3(256), lds, c->stream, c->nf_buf, c->nf_rows,
                       c->nf_geom.rowsE, c->nf_K, kused, c->nf_tw, c->nf_NC, c->nf_acc, c->nf_cmp);
    CHECK_LAUNCH(c);
    c->nf_slot = 0;
    return EFS_OK;
}

void free_all(efs_ctx* c) {
    dfree(c->pool);
    dfree(c->sweep_tab4); dfree(c->sweep_tab1);
    c->volt = c->curr = c->vv = c->vi = c->ii = c->iv = nullptr;
    dfree(c->cidx_e); dfree(c->cidx_h); dfree(c->ctab_e); dfree(c->ctab_h);
    dfree(c->cidx16_e); dfree(c->cidx16_h); dfree(c->ctab16_e); dfree(c->ctab16_h);
    c->comp = false;
    c->comp16 = false;
    dfree(c->exc_signal); dfree(c->exc_amp); dfree(c->exc_delay); dfree(c->exc_gidx);
    dfree(c->pml_coef); dfree(c->pml_vflux); dfree(c->pml_cflux);
    dfree(c->pml_idx); dfree(c->pml_tab_e); dfree(c->pml_tab_h);
    dfree(c->pml_pad_g); dfree(c->pml_pad_f);
    c->pml_pad_n = 0;
    dfree(c->pad_hv_g); dfree(c->pad_hv_f); dfree(c->pad_hi_g); dfree(c->pad_hi_f);
    c->pad_hv_n = c->pad_hi_n = 0;
    dfree(c->fscale); dfree(c->f_amax); dfree(c->f_factor); dfree(c->f_smin);
    c->half = false;
    dfree(c->mur_coeff); dfree(c->mur_saved); dfree(c->mur_normal); dfree(c->mur_shifted);
    dfree(c->mur_early); dfree(c->mur_early_list);
    dfree(c->mur_defer); dfree(c->exc_defer); dfree(c->mur_b); dfree(c->exc_b);
    dfree(c->grp_mur); dfree(c->grp_exc); dfree(c->grp_eoff);
    dfree(c->ep_moff); dfree(c->ep_mur); dfree(c->ep_eoff); dfree(c->ep_exc);
    c->ep_ok = false;
    c->post_e_ok = false;
    c->mur_nb = c->exc_nb = c->grp_n = 0;
    dfree(c->mur_normal_src); dfree(c->mur_shifted_src);
    dfree(c->ade_gidx); dfree(c->ade_src_v); dfree(c->ade_src_i); dfree(c->ade_lor);
    for (auto& p : c->ade_c) dfree(p);
    dfree(c->ade_va); dfree(c->ade_vl); dfree(c->ade_ia); dfree(c->ade_il);
    dfree(c->tfsf_signal); dfree(c->tv_idx); dfree(c->tv_dint); dfree(c->tv_dfrac); dfree(c->tv_amp);
    dfree(c->ti_idx); dfree(c->ti_dint); dfree(c->ti_dfrac); dfree(c->ti_amp);
    dfree(c->rlc_gidx); dfree(c->rlc_type); dfree(c->rlc_c); dfree(c->rlc_vdn); dfree(c->rlc_jn); dfree(c->rlc_vil);
    dfree(c->g_idx); dfree(c->g_out); dfree(c->gb_idx); dfree(c->gb_ring);
    dfree(c->pbc_imag); dfree(c->ss_gidx); dfree(c->ss_e);
    dfree(c->nf_pts); dfree(c->nf_el); dfree(c->nf_buf); dfree(c->nf_acc); dfree(c->nf_cmp); dfree(c->nf_omega); dfree(c->nf_perm_d); dfree(c->nf_tw);
    dfree(c->nf_slab);
    dfree(c->en_partial);
}

void reset_extensions(efs_ctx* c) {
    c->exc_n = 0;
    c->pml = PmlTable{};
    c->pml_tz0.clear();
    c->pml_tz1.clear();
    c->pml_pad_n = 0;
    c->h_pad.clear();
    c->pad_dirty = true;
    c->pad_hv_n = c->pad_hi_n = 0;
    c->h_tv_idx.clear();
    c->h_rlc_dev.clear();
    c->mur_n = 0;
    c->mur_ne = 0;
    c->h_mur_normal.clear();
    c->h_mur_shifted.clear();
    c->h_exc_gidx.clear();
    c->h_ade_gidx.clear();
    c->h_ade_lor.clear();
    for (auto& v : c->h_ade_c) v.clear();
    c->ade_n = 0;
    c->ade_dirty = false;
    c->alias_dirty = false;
    c->tfsf_nv = c->tfsf_ni = 0;
    c->h_ti_idx.clear();
    c->cap_dirty = true;
    c->fuse_dirty = true;
    c->rlc_n = 0;
    c->g_n = 0;
    c->gb_n = 0;
    c->gb_count = 0;
    c->nf_P = 0;
    c->nf_cap = NfCapture{};
    c->nf_slot = 0;
    c->pbc_e.clear();
    c->pbc_h.clear();
    c->source_end = 0;
    c->source_periodic = false;
    c->ss_n = 0;
    c->ss_current = 0;
    c->ss_recording = false;
}

bool box_overlap(const PmlRegion& a, const PmlRegion& b) {
    return a.x0 < b.x1 && b.x0 < a.x1 && a.y0 < b.y1 && b.y0 < a.y1 && a.z0 < b.z1 && b.z0 < a.z1;
}

// ---- one timestep ----
// Sweep geometries (4-wide and 1-wide lanes): block tables of equal-cost
// workgroups.  Cost model in bytes per cell: a core cell moves V r+w, H and
// its coefficients (39 B with byte-indexed coefficients, 60 B without); a
// UPML cell in the same launch adds flux r+w and three UPML coefficients
// (60 B, or 27 B byte-indexed); in the interior launch of the split form it
// costs only the loop.  Blocks that contain UPML cells are flagged (they
// stage the UPML coefficient tables).
// The launch gets ~2048 workgroups (8 per CU on 256 CUs) but no chunk
// shorter than 2 planes (a chunk re-reads one halo plane): a 256^3 grid
// gets chunks of 8 planes on average.  No chunk is longer than 48 planes
// (EFS_SWEEP_MAX_CHUNK) on average: the workgroups of one XCD march side by
// side over neighbouring tiles and read each other's edge rows from its
// L2, and over long chunks they drift apart and those reads go to HBM
// (800 x 800 x 600 in FP16: 2048 chunks of 182 planes 65.6k MCells/s,
// chunks of ~46 planes 73.8k).  Grids small enough for side work
// (a few hundred thousand cells) get 1-plane chunks: there the march is
// latency-bound, not HBM-bound, and a shorter march ends sooner.
int plan_sweeps(efs_ctx* c) {
    const uint32_t nx = c->nx, ny = c->ny, nz = c->nzp;  // device rows
    const uint64_t cols = (uint64_t)ny * nz;
    const bool upml_in_grid = !(c->vec4 && c->pml_split && !c->pml_fused);
    const double core_b = c->comp ? 39.0 : c->comp16 ? 42.0 : 60.0;
    const double upml_b = upml_in_grid ? core_b + (c->pml_comp ? 27.0 : 60.0) : 2.0;
    const uint64_t want = env_u64("EFS_SWEEP_BLOCKS", 2048);
    const uint64_t max_chunk = std::max<uint64_t>(1, env_u64("EFS_SWEEP_MAX_CHUNK", 48));
    const uint32_t min_planes = (uint32_t)std::max<uint64_t>(1, env_u64("EFS_SWEEP_MIN_PLANES", 1));
    // grids small enough for side work march 1-plane chunks (their blocks
    // are latency-bound; the x-face chunks keep 2 planes for the Mur pairs)
    const uint32_t min_chunk =
        (uint32_t)std::max<uint64_t>(1, env_u64("EFS_SWEEP_MIN_CHUNK", c->side_on ? 1 : 2));
    for (int v = 0; v < 2; ++v) {
        // 4-wide tiles of grids with side work hold whole rows of lanes when
        // that leaves at most 10 % of the threads idle, so that a Mur pair
        // across y (rows y and y+1) stays in one block (sweep_epilogue)
        uint32_t tl = kSweepThreads;
        if (v == 0 && c->side_on && nz % 4 == 0) {
            const uint32_t lpr = nz / 4, r = lpr <= (uint32_t)kSweepThreads ? kSweepThreads / lpr * lpr : 0u;
            if (r * 10 >= (uint32_t)kSweepThreads * 9) tl = r;
        }
        const uint64_t tcells = (uint64_t)tl * (v == 0 ? 4u : 1u);
        const uint32_t T = (uint32_t)((cols + tcells - 1) / tcells);
        // UPML cells of each region in each tile (per plane of the region)
        std::vector<std::vector<double>> cnt(c->pml.n, std::vector<double>(T, 0.0));
        for (uint32_t k = 0; k < c->pml.n; ++k) {
            const PmlRegion& R = c->pml.r[k];
            for (uint32_t y = R.y0; y < R.y1; ++y) {
                uint64_t a = (uint64_t)y * nz + R.z0;
                const uint64_t e = (uint64_t)y * nz + R.z1;
                while (a < e) {
                    const uint64_t t = a / tcells, te = std::min(e, (t + 1) * tcells);
                    cnt[k][t] += (double)(te - a);
                    a = te;
                }
            }
        }
        std::vector<double> cost((size_t)T * nx);  // [tile][x]
        std::vector<uint8_t> up((size_t)T * nx, 0);  // tile and plane hold UPML cells
        double total = 0.0;
        for (uint32_t t = 0; t < T; ++t) {
            const double cells = (double)(std::min(cols, (uint64_t)(t + 1) * tcells) - (uint64_t)t * tcells);
            for (uint32_t x = 0; x < nx; ++x) cost[(size_t)t * nx + x] = core_b * cells;
        }
        for (uint32_t k = 0; k < c->pml.n; ++k)
            for (uint32_t t = 0; t < T; ++t)
                if (cnt[k][t] > 0)
                    for (uint32_t x = c->pml.r[k].x0; x < c->pml.r[k].x1; ++x) {
                        cost[(size_t)t * nx + x] += (upml_b - core_b) * cnt[k][t];
                        up[(size_t)t * nx + x] = 1;
                    }
        for (double q : cost) total += q;
        // small grids have few column tiles: cut them into more, shorter
        // chunks (down to min_planes planes, default 1) so that the launch
        // still fills the GPU
        uint64_t want_v = std::max<uint64_t>(want, ((uint64_t)T * nx + max_chunk - 1) / max_chunk);
        if (v == 0) want_v = std::max<uint64_t>(1, (uint64_t)std::llround((double)want_v * c->blocks_scale));
        const double nblocks =
            (double)std::max<uint64_t>(1, std::min<uint64_t>(want_v, (uint64_t)T * std::max(1u, nx / min_planes)));
        const double target = total / nblocks;
        std::vector<SweepBlock> tab;
        std::vector<double> pre(nx + 1);
        auto chunks_of = [&](uint32_t t) {
            const double* ct = &cost[(size_t)t * nx];
            pre[0] = 0.0;
            for (uint32_t x = 0; x < nx; ++x) pre[x + 1] = pre[x] + ct[x];
            return (uint32_t)std::max(1.0, std::min((double)std::max(1u, nx / min_chunk),
                                                    std::floor(pre[nx] / target + 0.5)));
        };
        // the last tile of a plane holds fewer lanes: equal bytes would give
        // it longer chunks, but a march's time grows with its planes when
        // the grid is too small to load HBM (its block would end last), so
        // it gets at least the chunk count of the full tiles
        const bool partial = cols % tcells != 0;
        uint32_t n_full = 1;
        if (partial)
            for (uint32_t t = 0; t + 1 < T; ++t) n_full = std::max(n_full, chunks_of(t));
        for (uint32_t t = 0; t < T; ++t) {
            uint32_t n = chunks_of(t);
            if (partial && t + 1 == T) n = std::max(n, std::min(n_full, std::max(1u, nx / min_chunk)));
            // boundary j at the plane whose prefix cost is nearest j/n of the tile's total
            uint32_t x = 0, b0 = 0;
            for (uint32_t j = 1; j <= n; ++j) {
                uint32_t b = nx;
                if (j < n) {
                    const double goal = pre[nx] * j / n;
                    while (x < nx && pre[x + 1] <= goal) ++x;
                    b = (x < nx && goal - pre[x] > pre[x + 1] - goal) ? x + 1 : x;
                    // grids with side work keep planes 0, 1 and Nx-2, Nx-1
                    // in one chunk each (Mur pairs across x)
                    if (c->side_on && nx > 3) {
                        if (b == 1) b = 2;
                        if (b == nx - 1) b = nx - 2;
                    }
                }
                if (b > b0) {
                    uint32_t flags = 0;
                    for (uint32_t xx = b0; xx < b; ++xx)
                        if (up[(size_t)t * nx + xx]) flags = kBlockUpml;
                    tab.push_back(SweepBlock{t, b0, b, flags});
                    b0 = b;
                }
            }
        }
        // with side work the blocks of the x-face chunks end with the Mur of
        // a whole face plane (sweep_epilogue) and finish last: they are cut
        // into EFS_XFACE_PARTS blocks of whole rows of their tile (the rows
        // 0, 1 and Ny-2, Ny-1 of the y-face Mur pairs stay together)
        const uint32_t lpr = nz / 4;
        const uint32_t parts = (uint32_t)std::min<uint64_t>(16, env_u64("EFS_XFACE_PARTS", 2));
        if (v == 0 && c->side_on && nx > 3 && nz % 4 == 0 && tl % lpr == 0 && parts > 1) {
            const uint32_t lanes = (uint32_t)(cols / 4);
            std::vector<SweepBlock> out;
            for (const SweepBlock& b : tab) {
                const uint32_t tl_t = std::min(tl, lanes - b.tile * tl), r0 = b.tile * tl / lpr;
                const uint32_t rows = tl_t / lpr;
                if ((b.x0 != 0 && b.x1 != nx) || b.flags != (b.flags & kBlockUpml) || rows < 2) {
                    out.push_back(b);
                    continue;
                }
                // row boundaries r0 = s_0 < s_1 < ... < s_m = r0 + rows
                std::vector<uint32_t> cut{r0};
                for (uint32_t j = 1; j < parts; ++j) {
                    uint32_t rs = r0 + (uint32_t)((uint64_t)rows * j / parts);
                    if (rs == 1) rs = 2;
                    if (rs == ny - 1) rs = ny - 2;
                    if (rs > cut.back() && rs < r0 + rows) cut.push_back(rs);
                }
                cut.push_back(r0 + rows);
                const uint32_t last = tl_t - (cut[cut.size() - 2] - r0) * lpr;  // the partial row of a last tile
                for (size_t j = 0; j + 1 < cut.size(); ++j) {
                    const uint32_t l0 = (cut[j] - r0) * lpr;
                    const uint32_t n = j + 2 == cut.size() ? last : (cut[j + 1] - cut[j]) * lpr;
                    out.push_back(SweepBlock{b.tile, b.x0, b.x1, b.flags | (cut.size() > 2 ? block_lanes(l0, n) : 0u)});
                }
            }
            tab.swap(out);
        }
        std::stable_sort(tab.begin(), tab.end(), [](const SweepBlock& a, const SweepBlock& b) {
            return a.x0 + a.x1 != b.x0 + b.x1 ? a.x0 + a.x1 < b.x0 + b.x1 : a.tile < b.tile;
        });
        SweepGeom g{};
        g.nx = nx; g.ny = ny; g.nz = nz;
        g.cols = (uint32_t)cols;
        g.S = c->S;
        g.blocks = (uint32_t)tab.size();
        g.tile_lanes = tl;
        if (v == 0) c->h_tab4 = tab;
        SweepBlock*& dtab = v == 0 ? c->sweep_tab4 : c->sweep_tab1;
        TRY(dupload(c, dtab, tab.data(), tab.size()));
        g.tab = dtab;
        (v == 0 ? c->geom : c->geom1) = g;
    }
    return EFS_OK;
}

// A field array as the point kernels see it (storage type and scales).
FieldView fv(const efs_ctx* c, int curr) {
    return FieldView{curr ? (void*)c->curr : (void*)c->volt, c->fscale + 3 * curr, (uint64_t)c->S,
                     c->half ? 1u : 0u};
}

FieldView64 fv64(const efs_ctx* c, int curr) {
    return FieldView64{reinterpret_cast<double*>(curr ? c->curr : c->volt), (uint64_t)c->S};
}

// Floats per element of the arrays held in the arithmetic type (2 for FP64).
inline size_t rw(const efs_ctx* c) { return c->dbl ? 2 : 1; }

template <class FV>
typename FV::R* rp(float* p) { return reinterpret_cast<typename FV::R*>(p); }

// The E sweeps end with Mur post + apply and the excitation of their own
// cells (sweep_epilogue) in the steps of this configuration.
bool epilogue_on(const efs_ctx* c) {
    const bool has_pml = c->pml.n > 0;
    return c->ep_ok && c->vec4 && (!has_pml || (c->pml_split && c->pml_fused)) && !c->tfsf_nv &&
           c->pbc_e.empty() && !(c->ade_n && c->ade_apply);
}

// The point-kernel launches that the FP64 form shares, for view type FV.
template <class FV>
void launch_mur_pre(efs_ctx* c, FV v) {
    hipLaunchKernelGGL(k_mur_pre<FV>, dim3(nblk(c->mur_n, c->pt)), dim3(c->pt), 0, c->stream, v,
                       rp<FV>(c->pml_vflux), c->mur_coeff, c->mur_normal_src, c->mur_shifted_src,
                       rp<FV>(c->mur_saved), c->mur_n);
}

template <class FV>
void launch_mur_post_some(efs_ctx* c, FV v) {
    hipLaunchKernelGGL(k_mur_post_some<FV>, dim3(nblk(c->mur_ne, c->pt)), dim3(c->pt), 0, c->stream, v,
                       c->mur_coeff, c->mur_shifted, c->mur_early_list, rp<FV>(c->mur_saved), c->mur_ne);
}

template <class FV>
void launch_excitation(efs_ctx* c, FV v, uint32_t t) {
    hipLaunchKernelGGL(k_excitation<FV>, dim3(nblk(c->exc_n, c->pt)), dim3(c->pt), 0, c->stream, v, c->exc_signal,
                       c->exc_len, c->exc_period, c->exc_delay, c->exc_amp, c->exc_gidx, c->exc_n, t);
}

// ADE state update / write-back and TFSF injection on the voltage (curr = 0)
// or current (curr = 1) side.
template <class FV>
void launch_ade_update(efs_ctx* c, FV f, int curr) {
    hipLaunchKernelGGL(k_ade_update<FV>, dim3(nblk(c->ade_n, c->pt)), dim3(c->pt), 0, c->stream, f,
                       rp<FV>(curr ? c->pml_cflux : c->pml_vflux), curr ? c->ade_src_i : c->ade_src_v, c->ade_lor,
                       c->ade_c[3 * curr], c->ade_c[3 * curr + 1], c->ade_c[3 * curr + 2],
                       rp<FV>(curr ? c->ade_ia : c->ade_va), rp<FV>(curr ? c->ade_il : c->ade_vl), c->ade_n);
}

template <class FV>
void launch_ade_apply(efs_ctx* c, FV f, int curr) {
    hipLaunchKernelGGL(k_ade_apply<FV>, dim3(nblk(c->ade_n, c->pt)), dim3(c->pt), 0, c->stream, f, c->ade_gidx,
                       rp<FV>(curr ? c->ade_ia : c->ade_va), c->ade_n);
}

template <class FV>
void launch_tfsf(efs_ctx* c, FV f, int curr, uint32_t t) {
    hipLaunchKernelGGL(k_tfsf<FV>, dim3(nblk(curr ? c->tfsf_ni : c->tfsf_nv, c->pt)), dim3(c->pt), 0, c->stream, f,
                       c->tfsf_signal, c->tfsf_len, c->tfsf_period, curr ? c->ti_idx : c->tv_idx,
                       curr ? c->ti_dint : c->tv_dint, curr ? c->ti_dfrac : c->tv_dfrac, curr ? c->ti_amp : c->tv_amp,
                       curr ? c->tfsf_ni : c->tfsf_nv, t);
}

// Mur post + apply and the excitation: k_post_e, then k_post_e_b for the
// deferred entries.
template <class FV>
void launch_post_e(efs_ctx* c, FV v, uint32_t t) {
    PostE<typename FV::R> q{};
    q.signal = c->exc_signal;
    q.len = c->exc_len;
    q.period = c->exc_period;
    q.delay = c->exc_delay;
    q.gidx = c->exc_gidx;
    q.amp = c->exc_amp;
    q.exc_n = c->exc_n;
    q.numTS = t;
    q.exc_defer = c->exc_defer;
    q.exc_b = c->exc_b;
    q.exc_nb = c->exc_nb;
    q.coeff = c->mur_coeff;
    q.normal = c->mur_normal;
    q.shifted = c->mur_shifted;
    q.saved = rp<FV>(c->mur_saved);
    q.mur_n = c->mur_n;
    q.mur_defer = c->mur_defer;
    q.mur_b = c->mur_b;
    q.mur_nb = c->mur_nb;
    q.grp = c->grp_n;
    q.grp_mur = c->grp_mur;
    q.grp_exc = c->grp_exc;
    q.grp_eoff = c->grp_eoff;
    hipLaunchKernelGGL(k_post_e<FV>, dim3(q.grp + nblk((size_t)c->mur_n + c->exc_n, kPostThreads)),
                       dim3(kPostThreads), 0, c->stream, v, q);
    if (c->exc_nb + c->mur_nb)
        hipLaunchKernelGGL(k_post_e_b<FV>, dim3(nblk((size_t)c->exc_nb + c->mur_nb, kPostThreads)),
                           dim3(kPostThreads), 0, c->stream, v, q);
}

template <class FV>
void launch_mur_finish(efs_ctx* c, FV v) {
    hipLaunchKernelGGL(k_mur_finish<FV>, dim3(nblk(c->mur_n, c->pt)), dim3(c->pt), 0, c->stream, v, c->mur_coeff,
                       c->mur_normal, c->mur_shifted, c->mur_early, rp<FV>(c->mur_saved), c->mur_n);
}

// Whether a hook that changes H after the H sweep writes inside the
// captured rows (z planes zf-1 .. zf+2 of the capture rectangle): TFSF
// current points, or ADE write-back cells with live current coefficients
// (taken as live while the ADE lists await their rebuild).
bool hooks_clash(const efs_ctx* c) {
    const NfCapture& k = c->nf_cap;
    if (!k.on) return false;
    const uint32_t N = (uint32_t)c->N, nz = c->nz, ny = c->ny;
    auto hit = [&](uint32_t gi) {
        const uint32_t j = gi % N, z = j % nz, y = (j / nz) % ny, x = j / (nz * ny);
        if (x < k.x0 || x > k.x1 || y < k.y0 || y > k.y1) return false;
        for (uint32_t f = 0; f < k.nf; ++f)
            if (z + 1 >= k.z[f] && z <= k.z[f] + 2) return true;
        return false;
    };
    for (uint32_t g : c->h_ti_idx)
        if (hit(g)) return true;
    if (c->ade_apply && (c->ade_i_live || c->ade_dirty))
        for (uint32_t g : c->h_ade_gidx)
            if (hit(g)) return true;
    return false;
}

// The NF2FF capture for this step, or an inactive one: it needs the 4-wide
// grid sweeps, no UPML cell in the captured rows unless the sweeps are the
// fused form, no periodic H faces, and no other hook that changes H after
// the H sweep inside the captured rows (TFSF-I, ADE write-back).
NfCapture nf_capture(const efs_ctx* c) {
    const NfCapture& k = c->nf_cap;
    if (!k.on || !c->nf_P || !c->vec4 || (c->pml.n && !c->pml_split)) return NfCapture{};
    if (!c->pbc_h.empty() || (c->cap_dirty ? hooks_clash(c) : c->cap_clash)) return NfCapture{};
    // In the fused sweeps a lane computes its UPML cells' final values
    // before it captures them; only the two-launch form updates UPML cells
    // in a later launch, after the capture.
    for (uint32_t r = 0; r < c->pml.n && !c->pml_fused; ++r) {
        // the region as given: padding cells of a widened region update as
        // core cells, so the capture holds their values too
        const PmlRegion& R = c->pml.r[r];
        if (R.x0 > k.x1 || R.x1 <= k.x0 || R.y0 > k.y1 || R.y1 <= k.y0) continue;
        for (uint32_t f = 0; f < k.nf; ++f)
            if (c->pml_tz0[r] <= k.z[f] + 1 && c->pml_tz1[r] > k.z[f]) return NfCapture{};
    }
    return k;
}

// Choose the captured planes: the (at most two) z planes holding the most
// surface points, each with its zf and zf+1 in one 4-cell lane and strictly
// inside the grid; the rectangle covers their points plus one row in +x, +y.
int plan_nf_capture(efs_ctx* c, const std::vector<uint32_t>& pts, uint32_t P) {
    c->nf_cap = NfCapture{};
    if (env_u64("EFS_NF2FF_CAPTURE", 1) == 0 || c->nz < 4) return EFS_OK;
    std::vector<uint32_t> cnt(c->nz, 0);
    for (uint32_t p = 0; p < P; ++p) cnt[pts[3 * (size_t)p + 2]]++;
    std::vector<uint32_t> zs;
    for (uint32_t z = 0; z < c->nz; ++z)
        if (cnt[z] >= 256 && z % 4 != 3 && z + 2 < c->nz) zs.push_back(z);
    std::sort(zs.begin(), zs.end(), [&](uint32_t a, uint32_t b) { return cnt[a] != cnt[b] ? cnt[a] > cnt[b] : a < b; });
    if (zs.size() > 2) zs.resize(2);
    // the H sweep captures at most one plane per 4-cell lane: two planes in
    // one lane keep the busier one (the other one's points are gathered)
    if (zs.size() == 2 && zs[0] / 4 == zs[1] / 4) zs.resize(1);
    if (zs.empty()) return EFS_OK;
    uint32_t x0 = ~0u, x1 = 0, y0 = ~0u, y1 = 0;
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t z = pts[3 * (size_t)p + 2];
        if (std::find(zs.begin(), zs.end(), z) == zs.end()) continue;
        x0 = std::min(x0, pts[3 * (size_t)p]);
        x1 = std::max(x1, pts[3 * (size_t)p] + 1);
        y0 = std::min(y0, pts[3 * (size_t)p + 1]);
        y1 = std::max(y1, pts[3 * (size_t)p + 1] + 1);
    }
    if (x1 + 1 >= c->nx || y1 + 1 >= c->ny) return EFS_OK;  // rows must have the core update
    NfCapture k{};
    k.on = 1;
    k.nf = (uint32_t)zs.size();
    for (uint32_t f = 0; f < k.nf; ++f) k.z[f] = zs[f];
    k.x0 = x0; k.x1 = x1; k.y0 = y0; k.y1 = y1;
    k.nx_c = x1 - x0 + 1;
    k.ny_c = y1 - y0 + 1;
    TRY(dzero(c, c->nf_slab, (size_t)k.nf * k.nx_c * k.ny_c * kCapFloats));
    k.slab = c->nf_slab;
    c->nf_cap = k;
    c->cap_dirty = true;
    c->fuse_dirty = true;
    return EFS_OK;
}

// FP16 storage: largest |stored| per component, then rescale where needed.
void launch_range_control(efs_ctx* c, uint32_t t) {
    const uint32_t N = (uint32_t)c->Np;  // padding cells are zero
    const dim3 grid(std::max(1u, std::min(256u, (N / 4 + 255) / 256)), 6);
    hipLaunchKernelGGL(k_absmax16, grid, dim3(256), 0, c->stream, (const f16*)c->volt, (const f16*)c->curr,
                       (uint64_t)c->S, N, c->f_amax);
    const uint32_t active = c->source_periodic || (uint64_t)t + 1 < c->source_end;
    hipLaunchKernelGGL(k_range_plan, dim3(1), dim3(64), 0, c->stream, c->f_amax, c->fscale, c->f_smin, c->f_factor,
                       active);
    hipLaunchKernelGGL(k_range_apply, grid, dim3(256), 0, c->stream, (f16*)c->volt, (f16*)c->curr, (uint64_t)c->S,
                       N, c->f_factor);
}

// NF2FF sample of this step (when configured), probe gather into `ring`
// (null: none due) and, with `pre`, the next step's Mur pre pass, in one
// launch (k_post_h).
template <class FV>
void launch_post_h(efs_ctx* c, FV v, FV i, NfCapture cap, float* ring, bool pre, bool tfsf_i, bool ade_next,
                   uint32_t t) {
    PostH<FV> q{};
    if (c->nf_P) {
        q.pts = c->nf_pts;
        q.el = c->nf_el;
        q.g = c->nf_geom;
        q.row_out = rp<FV>(c->nf_buf) + (size_t)c->nf_slot * c->nf_rows;
        q.cap = cap;
        q.nb_sample = nblk(c->nf_P, c->pt);
    }
    if (ring) {
        q.gidx = c->gb_idx;
        q.gout = ring;
        q.gn = c->gb_n;
        q.nb_gather = nblk(c->gb_n, c->pt);
    }
    uint32_t nb_pre = 0;
    if (pre) {
        q.vflux = rp<FV>(c->pml_vflux);
        q.coeff = c->mur_coeff;
        q.normal_src = c->mur_normal_src;
        q.shifted_src = c->mur_shifted_src;
        q.saved = rp<FV>(c->mur_saved);
        q.mur_n = c->mur_n;
        nb_pre = nblk(c->mur_n, c->pt);
    }
    q.nb_pre = nb_pre;
    if (tfsf_i) {
        q.t_signal = c->tfsf_signal;
        q.t_len = c->tfsf_len;
        q.t_period = c->tfsf_period;
        q.t_numTS = t;
        q.t_idx = c->ti_idx;
        q.t_dint = c->ti_dint;
        q.t_dfrac = c->ti_dfrac;
        q.t_amp = c->ti_amp;
        q.t_n = c->tfsf_ni;
        q.nb_tfsf = nblk(c->tfsf_ni, c->pt);
    }
    uint32_t nb_ade = 0;
    if (ade_next) {
        q.a_flux = rp<FV>(c->pml_vflux);
        q.a_src = c->ade_src_v;
        q.a_lor = c->ade_lor;
        q.a_int = c->ade_c[0];
        q.a_ext = c->ade_c[1];
        q.a_cl = c->ade_c[2];
        q.a_A = rp<FV>(c->ade_va);
        q.a_L = rp<FV>(c->ade_vl);
        q.a_n = c->ade_n;
        nb_ade = nblk(c->ade_n, c->pt);
    }
    hipLaunchKernelGGL(k_post_h<FV>, dim3(q.nb_sample + q.nb_gather + nb_pre + q.nb_tfsf + nb_ade), dim3(c->pt), 0,
                       c->stream, v, i, q);
}

// TFSF on the E side and the ADE write-back in one launch (k_post_v).
template <class FV>
void launch_post_v(efs_ctx* c, FV v, uint32_t t) {
    PostV<FV> q{};
    q.signal = c->tfsf_signal;
    q.len = c->tfsf_len;
    q.period = c->tfsf_period;
    q.numTS = t;
    q.t_idx = c->tv_idx;
    q.t_dint = c->tv_dint;
    q.t_dfrac = c->tv_dfrac;
    q.t_amp = c->tv_amp;
    q.t_n = c->tfsf_nv;
    q.nb_tfsf = nblk(c->tfsf_nv, c->pt);
    q.a_gidx = c->ade_gidx;
    q.a_A = rp<FV>(c->ade_va);
    q.a_n = c->ade_n;
    hipLaunchKernelGGL(k_post_v<FV>, dim3(q.nb_tfsf + nblk(c->ade_n, c->pt)), dim3(c->pt), 0, c->stream, v, q);
}

// Which TFSF and ADE point work may share the launches of other hooks:
// TFSF-V with the ADE write-back when no E cell is in both lists (their
// adds then commute exactly); TFSF-I with the NF2FF sampling and the
// buffered gather when none of the H cells those read is a TFSF-I cell.
void plan_point_fusion(efs_ctx* c) {
    c->post_v_fuse = c->post_i_fuse = c->ade_fuse = false;
    c->fuse_dirty = false;
    if (!c->fuse_hooks || env_u64("EFS_FUSE_POINTS", 1) == 0) return;
    c->ade_fuse = true;
    // cell sets as bitmaps over the 3N interface indices (linear-time
    // checks: the first iterate call after a change pays for them)
    std::vector<bool> mark;
    auto fill = [&](const std::vector<uint32_t>& v) {
        mark.assign(3 * c->N, false);
        for (uint32_t g : v)
            if (g < 3 * c->N) mark[g] = true;
    };
    if (c->tfsf_nv && c->ade_n) {
        fill(c->h_tv_idx);
        bool both = false;
        for (size_t k = 0; k < c->h_ade_gidx.size() && !both; ++k)
            both = c->h_ade_gidx[k] < 3 * c->N && mark[c->h_ade_gidx[k]];
        c->post_v_fuse = !both;
    }
    if (c->tfsf_ni) {
        fill(c->h_ti_idx);
        auto hit = [&](uint64_t g) { return g < 3 * c->N && mark[g]; };
        bool clash = false;
        const uint64_t sx = (uint64_t)c->ny * c->nz, sy = c->nz;
        // the H values of a sampled point: component n at the point and one
        // cell further along axis n
        for (size_t p = 0; c->nf_P && !clash && p < c->h_nf_pts.size() / 3; ++p) {
            const uint64_t b = c->h_nf_pts[3 * p] * sx + c->h_nf_pts[3 * p + 1] * sy + c->h_nf_pts[3 * p + 2];
            const uint64_t step[3] = {sx, sy, 1};
            for (uint32_t n = 0; n < 3 && !clash; ++n)
                clash = hit(n * c->N + b) || hit(n * c->N + b + step[n]);
        }
        for (size_t k = 0; c->gb_n && !clash && k < c->h_gb_idx.size(); ++k) clash = hit(c->h_gb_idx[k]);
        c->post_i_fuse = !clash;
    }
}

// Side work of the H sweep of step t (SideWork): the E half of the NF2FF
// sample, the E values of a due gather into `ring` (null: none due) and,
// with `pre`, the Mur pre pass of step t+1.  Leaves in c->side_e the rest
// of the sample and gather, for the next E sweep or side_flush.
SideWork side_h(efs_ctx* c, uint32_t t, float* ring, bool pre) {
    SideWork q{};
    q.volt = fv(c, 0);
    q.curr = fv(c, 1);
    if (c->nf_P) {
        if (c->nf_slot == 0) c->nf_ts0 = t + 1;
        q.pts = c->nf_pts;
        q.el = c->nf_el;
        q.g = c->nf_geom;
        q.row_out = c->nf_buf + (size_t)c->nf_slot * c->nf_rows;
        q.cap = nf_capture(c);
        q.nb_sample = nblk(c->nf_P, kSideThreads);
    }
    if (ring) {
        q.gidx = c->gb_idx;
        q.gout = ring;
        q.gn = c->gb_n;
        q.nb_gather = nblk(c->gb_n, kSideThreads);
    }
    const uint32_t nb = q.nb_sample + q.nb_gather;
    c->side_e = q;
    c->side_e.nb = (nb + 7u) & ~7u;
    c->side_pend = nb > 0;
    c->side_pend_gather = ring != nullptr;
    uint32_t nb_pre = 0;
    if (pre) {
        q.vflux = c->pml_vflux;
        q.coeff = c->mur_coeff;
        q.normal_src = c->mur_normal_src;
        q.shifted_src = c->mur_shifted_src;
        q.saved = c->mur_saved;
        q.mur_n = c->mur_n;
        nb_pre = nblk(c->mur_n, kSideThreads);
    }
    q.nb = (nb + nb_pre + 7u) & ~7u;
    return q;
}

// The epilogue of the E sweep of step t (sweep_epilogue) into q.
void add_epilogue(const efs_ctx* c, SideWork& q, uint32_t t) {
    q.volt = fv(c, 0);
    q.curr = fv(c, 1);
    q.ep_on = 1;
    q.ep_moff = c->ep_moff;
    q.ep_mur = c->ep_mur;
    q.ep_eoff = c->ep_eoff;
    q.ep_exc = c->ep_exc;
    q.m_saved = c->mur_saved;
    q.x_signal = c->exc_signal;
    q.x_len = c->exc_len;
    q.x_period = c->exc_period;
    q.numTS = t;
    q.x_delay = c->exc_delay;
    q.x_gidx = c->exc_gidx;
    q.x_amp = c->exc_amp;
}

// The pending sample and gather are complete (their E-sweep side was
// launched): count them, and fold a full K-step block.
int side_done(efs_ctx* c) {
    if (!c->side_pend) return EFS_OK;
    c->side_pend = false;
    if (c->side_pend_gather) c->gb_count++;
    if (c->side_e.nb_sample) {
        c->nf_slot++;
        if (c->nf_slot == c->nf_K) TRY(nf2ff_fold(c));
    }
    return EFS_OK;
}

// The E-sweep side of the last step of an iterate call, on its own.
int side_flush(efs_ctx* c) {
    if (!c->side_pend) return EFS_OK;
    hipLaunchKernelGGL(k_side_flush, dim3(c->side_e.nb), dim3(kSideThreads), 0, c->stream, c->side_e);
    CHECK_LAUNCH(c);
    return side_done(c);
}

// One step.  next_follows: the same efs_iterate call runs another step
// right after this one (its Mur pre pass may then run in this step's
// k_post_h, or in the side work of its H sweep).
int launch_step(efs_ctx* c, bool next_follows) {
    const uint32_t t = c->numTS;
    ProfScope step_scope(c, EFS_PROF_STEP);
    const bool has_pml = c->pml.n > 0;
    // pre-voltage: steady-state energies first (priority +2M)
    if (c->ss_n && c->ss_recording) {
        hipLaunchKernelGGL(k_ss_accumulate, dim3(nblk(c->ss_n, c->pt)), dim3(c->pt), 0, c->stream, fv(c, 0),
                           c->ss_gidx, c->ss_n, c->ss_period, c->ss_current, c->ss_e, c->ss_e + c->ss_n);
        c->ss_current++;
    }
    if (c->ade_n && !c->ade_ahead) {
        if (c->dbl) launch_ade_update(c, fv64(c, 0), 0);
        else launch_ade_update(c, fv(c, 0), 0);
    }
    c->ade_ahead = false;
    if (c->mur_n && !c->mur_pre_ahead) {
        if (c->dbl) launch_mur_pre(c, fv64(c, 0));
        else launch_mur_pre(c, fv(c, 0));
    }
    c->mur_pre_ahead = false;
    if (c->rlc_n) {
        hipLaunchKernelGGL(k_rlc_pre, dim3(nblk(c->rlc_n, c->pt)), dim3(c->pt), 0, c->stream, c->rlc_type,
                           c->rlc_c, c->rlc_vdn, c->rlc_vil, c->rlc_n);
    }
    // Mur post + apply and the excitation at the end of the E-sweep blocks
    const bool use_ep = epilogue_on(c);
    {
        ProfScope s(c, EFS_PROF_SWEEP_E);
        if (c->vec4 && (!has_pml || c->pml_split)) {
            const Coef cf{c->vv, c->vi, c->cidx_e, c->comp16 ? c->ctab16_e : c->ctab_e, c->pml_idx, c->pml_tab_e,
                          c->cidx16_e, c->tabn_e};
            const bool x16 = c->comp16;
            auto kern = has_pml ? (x16 ? sweep_e4mx : c->comp ? sweep_e4mc : sweep_e4m)
                                : (x16 ? sweep_e4x : c->comp ? sweep_e4c : sweep_e4);
            if (has_pml && c->pml_fused)
                kern = c->pml_comp ? (x16 ? sweep_e4fxp : c->comp ? sweep_e4fcp : sweep_e4fp)
                                   : (x16 ? sweep_e4fx : c->comp ? sweep_e4fc : sweep_e4f);
            // the previous step's pending sample and gather ride along, and
            // the blocks end with the Mur and excitation of their cells
            SideWork se = c->side_pend ? c->side_e : SideWork{};
            if (use_ep) add_epilogue(c, se, t);
            if (se.nb || se.ep_on) kern = side_twin(kern);
            if (c->half) {
                auto k16 = has_pml ? (c->pml_comp ? (x16 ? sweep_e4fxp_h : c->comp ? sweep_e4fcp_h : sweep_e4fp_h)
                                                  : (x16 ? sweep_e4fx_h : c->comp ? sweep_e4fc_h : sweep_e4f_h))
                                   : (x16 ? sweep_e4x_h : c->comp ? sweep_e4c_h : sweep_e4_h);
                if (se.nb || se.ep_on) k16 = side_twin(k16);
                hipLaunchKernelGGL(k16, dim3(se.nb + c->geom.blocks), dim3(kSweepThreads), 0, c->stream,
                                   (f16*)c->volt, (const f16*)c->curr, cf, c->geom, c->pml, c->pml_coef,
                                   c->pml_vflux, c->fscale, NfCapture{}, se);
            } else {
                hipLaunchKernelGGL(kern, dim3(se.nb + c->geom.blocks), dim3(kSweepThreads), 0, c->stream, c->volt,
                                   c->curr, cf, c->geom, c->pml, c->pml_coef, c->pml_vflux, c->fscale, NfCapture{},
                                   se);
            }
            if (has_pml && !c->pml_fused)
                hipLaunchKernelGGL(x16 ? sweep_e4ux : c->comp ? sweep_e4uc : sweep_e4u, dim3(c->pml.blocks),
                                   dim3(kSweepThreads), 0,
                                   c->stream, c->volt, c->curr, cf, c->geom, c->pml, c->pml_coef, c->pml_vflux,
                                   c->fscale, NfCapture{}, SideWork{});
            TRY(side_done(c));
        } else if (c->dbl) {
            auto kern = has_pml ? sweep_e1<true, double> : sweep_e1<false, double>;
            hipLaunchKernelGGL(kern, dim3(c->geom1.blocks), dim3(kSweepThreads), 0, c->stream, (double*)c->volt,
                               (const double*)c->curr, c->vv, c->vi, c->geom1, c->pml, c->pml_coef,
                               (double*)c->pml_vflux);
        } else {
            auto kern = has_pml ? sweep_e1<true> : sweep_e1<false>;
            hipLaunchKernelGGL(kern, dim3(c->geom1.blocks), dim3(kSweepThreads), 0, c->stream, c->volt, c->curr,
                               c->vv, c->vi, c->geom1, c->pml, c->pml_coef, c->pml_vflux);
        }
    }
    // post-voltage / apply-to-voltages; TFSF and the ADE write-back share
    // a launch when they are the only work here and touch disjoint cells
    const bool fuse_v = c->post_v_fuse && c->tfsf_nv && c->ade_n && c->ade_apply && c->pbc_e.empty() &&
                        !c->mur_n && !c->exc_n && !c->rlc_n;
    if (fuse_v) {
        if (c->dbl) launch_post_v(c, fv64(c, 0), t);
        else launch_post_v(c, fv(c, 0), t);
    } else if (c->tfsf_nv) {
        if (c->dbl) launch_tfsf(c, fv64(c, 0), 0, t);
        else launch_tfsf(c, fv(c, 0), 0, t);
    }
    for (const PbcAxis& a : c->pbc_e)  // after TFSF-V, before Mur post (src/webgpu-engine.mjs:2461)
        hipLaunchKernelGGL(k_pbc, dim3(nblk((size_t)a.b0 * a.b1, c->pt)), dim3(c->pt), 0, c->stream, fv(c, 0),
                           c->pbc_imag, a, c->ny, c->nzp, (uint64_t)c->S);
    if (use_ep) {
        // done by the E sweep's blocks
    } else if (c->mur_n && c->post_e_ok && !(c->ade_n && c->ade_apply)) {
        // Mur post + excitation + Mur apply in one launch
        if (c->dbl) launch_post_e(c, fv64(c, 0), t);
        else launch_post_e(c, fv(c, 0), t);
    } else {
        if (c->mur_n) {
            if (c->mur_ne) {
                if (c->dbl) launch_mur_post_some(c, fv64(c, 0));
                else launch_mur_post_some(c, fv(c, 0));
            }
        }
        if (c->ade_n && c->ade_apply && !fuse_v) {
            if (c->dbl) launch_ade_apply(c, fv64(c, 0), 0);
            else launch_ade_apply(c, fv(c, 0), 0);
        }
        if (c->exc_n) {
            if (c->dbl) launch_excitation(c, fv64(c, 0), t);
            else launch_excitation(c, fv(c, 0), t);
        }
        if (c->mur_n) {
            if (c->dbl) launch_mur_finish(c, fv64(c, 0));
            else launch_mur_finish(c, fv(c, 0));
        }
    }
    if (c->rlc_n) {
        hipLaunchKernelGGL(k_rlc_apply, dim3(nblk(c->rlc_n, c->pt)), dim3(c->pt), 0, c->stream, fv(c, 0),
                           c->rlc_gidx, c->rlc_type, c->rlc_c, c->rlc_vdn, c->rlc_jn, c->rlc_vil, c->rlc_n);
    }
    if (c->pad_hv_n)  // padding cells of widened UPML regions that the hooks above wrote
        hipLaunchKernelGGL(k_pml_pad_sync, dim3(nblk(c->pad_hv_n, c->pt)), dim3(c->pt), 0, c->stream, fv(c, 0),
                           c->pml_vflux, c->pad_hv_g, c->pad_hv_f, c->pad_hv_n);
    // pre-current
    if (c->ade_n && c->ade_i_live) {
        if (c->dbl) launch_ade_update(c, fv64(c, 1), 1);
        else launch_ade_update(c, fv(c, 1), 1);
    }
    // engine hooks of this step: in the H sweep's side work, or after it
    const bool range_due = c->half && (t + 1) % c->range_interval == 0;
    const bool gather_due = c->gb_n && (t + 1) % c->gb_interval == 0 && c->gb_count < c->gb_max;
    const bool pre_ahead = c->fuse_hooks && next_follows && c->mur_n && !range_due;
    const bool use_side = c->side_on && c->fuse_hooks && c->vec4 && (!has_pml || c->pml_split) && !c->dbl &&
                          !range_due && (c->nf_P || gather_due || pre_ahead);
    {
        ProfScope s(c, EFS_PROF_SWEEP_H);
        if (c->vec4 && (!has_pml || c->pml_split)) {
            const Coef cf{c->ii, c->iv, c->cidx_h, c->comp16 ? c->ctab16_h : c->ctab_h, c->pml_idx, c->pml_tab_h,
                          c->cidx16_h, c->tabn_h};
            const bool x16 = c->comp16;
            auto kern = has_pml ? (x16 ? sweep_h4mx : c->comp ? sweep_h4mc : sweep_h4m)
                                : (x16 ? sweep_h4x : c->comp ? sweep_h4c : sweep_h4);
            if (has_pml && c->pml_fused)
                kern = c->pml_comp ? (x16 ? sweep_h4fxp : c->comp ? sweep_h4fcp : sweep_h4fp)
                                   : (x16 ? sweep_h4fx : c->comp ? sweep_h4fc : sweep_h4f);
            float* ring = gather_due ? c->gb_ring + (size_t)c->gb_count * 2 * c->gb_n : nullptr;
            const SideWork sh = use_side ? side_h(c, t, ring, pre_ahead) : SideWork{};
            if (sh.nb) kern = side_twin(kern);
            if (c->half) {
                auto k16 = has_pml ? (c->pml_comp ? (x16 ? sweep_h4fxp_h : c->comp ? sweep_h4fcp_h : sweep_h4fp_h)
                                                  : (x16 ? sweep_h4fx_h : c->comp ? sweep_h4fc_h : sweep_h4f_h))
                                   : (x16 ? sweep_h4x_h : c->comp ? sweep_h4c_h : sweep_h4_h);
                if (sh.nb) k16 = side_twin(k16);
                hipLaunchKernelGGL(k16, dim3(sh.nb + c->geom.blocks), dim3(kSweepThreads), 0, c->stream,
                                   (f16*)c->curr, (const f16*)c->volt, cf, c->geom, c->pml, c->pml_coef,
                                   c->pml_cflux, c->fscale, nf_capture(c), sh);
            } else {
                hipLaunchKernelGGL(kern, dim3(sh.nb + c->geom.blocks), dim3(kSweepThreads), 0, c->stream, c->curr,
                                   c->volt, cf, c->geom, c->pml, c->pml_coef, c->pml_cflux, c->fscale,
                                   nf_capture(c), sh);
            }
            if (has_pml && !c->pml_fused)
                hipLaunchKernelGGL(x16 ? sweep_h4ux : c->comp ? sweep_h4uc : sweep_h4u, dim3(c->pml.blocks),
                                   dim3(kSweepThreads), 0,
                                   c->stream, c->curr, c->volt, cf, c->geom, c->pml, c->pml_coef, c->pml_cflux,
                                   c->fscale, NfCapture{}, SideWork{});
        } else if (c->dbl) {
            auto kern = has_pml ? sweep_h1<true, double> : sweep_h1<false, double>;
            hipLaunchKernelGGL(kern, dim3(c->geom1.blocks), dim3(kSweepThreads), 0, c->stream, (double*)c->curr,
                               (const double*)c->volt, c->ii, c->iv, c->geom1, c->pml, c->pml_coef,
                               (double*)c->pml_cflux);
        } else {
            auto kern = has_pml ? sweep_h1<true> : sweep_h1<false>;
            hipLaunchKernelGGL(kern, dim3(c->geom1.blocks), dim3(kSweepThreads), 0, c->stream, c->curr, c->volt,
                               c->ii, c->iv, c->geom1, c->pml, c->pml_coef, c->pml_cflux);
        }
    }
    // post-current: TFSF here, or in the hooks launch below (fuse_i)
    const bool hooks_fusable = !use_side && c->fuse_hooks && !range_due;
    bool fuse_i = hooks_fusable && c->tfsf_ni && c->post_i_fuse && c->pbc_h.empty() &&
                  !(c->ade_n && c->ade_apply && c->ade_i_live) && !c->pad_hi_n;
    // the next step's E-side ADE update reads E only, which nothing after
    // this step's write-back changes before that step's E sweep
    bool ade_next = hooks_fusable && next_follows && c->ade_n && c->ade_fuse;
    const int parts = (c->nf_P ? 1 : 0) + (gather_due ? 1 : 0) + (pre_ahead ? 1 : 0) + (fuse_i ? 1 : 0) +
                      (ade_next ? 1 : 0);
    if (parts < 2) fuse_i = ade_next = false;
    if (c->tfsf_ni && !fuse_i) {
        if (c->dbl) launch_tfsf(c, fv64(c, 1), 1, t);
        else launch_tfsf(c, fv(c, 1), 1, t);
    }
    for (const PbcAxis& a : c->pbc_h)  // after TFSF-I (src/webgpu-engine.mjs:2477)
        hipLaunchKernelGGL(k_pbc, dim3(nblk((size_t)a.b0 * a.b1, c->pt)), dim3(c->pt), 0, c->stream, fv(c, 1),
                           c->pbc_imag, a, c->ny, c->nzp, (uint64_t)c->S);
    if (c->ade_n && c->ade_apply && c->ade_i_live) {
        if (c->dbl) launch_ade_apply(c, fv64(c, 1), 1);
        else launch_ade_apply(c, fv(c, 1), 1);
    }
    if (c->pad_hi_n)
        hipLaunchKernelGGL(k_pml_pad_sync, dim3(nblk(c->pad_hi_n, c->pt)), dim3(c->pt), 0, c->stream, fv(c, 1),
                           c->pml_cflux, c->pad_hi_g, c->pad_hi_f, c->pad_hi_n);
    if (use_side) {  // the hooks ran beside the H sweep
        c->mur_pre_ahead = pre_ahead;
        CHECK_LAUNCH(c);
        c->numTS++;
        return EFS_OK;
    }
    // NF2FF sampling, the probe gather, the next step's Mur pre pass, TFSF
    // of the H side and the next step's E-side ADE update in one launch;
    // not when the FP16 range control (which rescales the fields) comes
    // between them
    if (c->fuse_hooks && !range_due && parts >= 2) {
        if (c->nf_P && c->nf_slot == 0) c->nf_ts0 = t + 1;
        float* ring = gather_due ? c->gb_ring + (size_t)c->gb_count * 2 * c->gb_n : nullptr;
        if (c->dbl) launch_post_h(c, fv64(c, 0), fv64(c, 1), NfCapture{}, ring, pre_ahead, fuse_i, ade_next, t);
        else launch_post_h(c, fv(c, 0), fv(c, 1), nf_capture(c), ring, pre_ahead, fuse_i, ade_next, t);
        if (gather_due) c->gb_count++;
        c->mur_pre_ahead = pre_ahead;
        c->ade_ahead = ade_next;
        if (c->nf_P) {
            c->nf_slot++;
            if (c->nf_slot == c->nf_K) TRY(nf2ff_fold(c));
        }
        CHECK_LAUNCH(c);
        c->numTS++;
        return EFS_OK;
    }
    if (c->nf_P) {
        if (c->nf_slot == 0) c->nf_ts0 = t + 1;
        if (c->dbl)
            hipLaunchKernelGGL(k_nf2ff_sample<FieldView64>, dim3(nblk(c->nf_P, c->pt)), dim3(c->pt), 0, c->stream,
                               fv64(c, 0), fv64(c, 1), c->nf_pts, c->nf_el, c->nf_geom,
                               (double*)c->nf_buf + (size_t)c->nf_slot * c->nf_rows, NfCapture{});
        else
            hipLaunchKernelGGL(k_nf2ff_sample<FieldView>, dim3(nblk(c->nf_P, c->pt)), dim3(c->pt), 0, c->stream,
                               fv(c, 0), fv(c, 1), c->nf_pts, c->nf_el, c->nf_geom,
                               c->nf_buf + (size_t)c->nf_slot * c->nf_rows, nf_capture(c));
        c->nf_slot++;
        if (c->nf_slot == c->nf_K) TRY(nf2ff_fold(c));
    }
    if (range_due) launch_range_control(c, t);
    if (gather_due) {
        float* ring = c->gb_ring + (size_t)c->gb_count * 2 * c->gb_n;
        if (c->dbl)
            hipLaunchKernelGGL(k_gather<FieldView64>, dim3(nblk(c->gb_n, c->pt)), dim3(c->pt), 0, c->stream,
                               fv64(c, 0), fv64(c, 1), c->gb_idx, ring, c->gb_n);
        else
            hipLaunchKernelGGL(k_gather<FieldView>, dim3(nblk(c->gb_n, c->pt)), dim3(c->pt), 0, c->stream, fv(c, 0),
                               fv(c, 1), c->gb_idx, ring, c->gb_n);
        c->gb_count++;
    }
    CHECK_LAUNCH(c);
    c->numTS++;
    return EFS_OK;
}

}  // namespace

// ===================================================================== C ABI

extern "C" {

const char* efs_version(void) { return "efs-mi355x 0.1"; }

int efs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int efs_create(efs_ctx** out, int device) {
    if (!out) return EFS_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_create_error = "no HIP device available";
        return EFS_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        g_create_error = "device index out of range";
        return EFS_ERR_ARG;
    }
    efs_ctx* c = new efs_ctx();
    c->device = device;
    {
        // workgroup size of the per-step point kernels (64, 128 or 256):
        // 64-thread groups spread short point lists over more CUs (the
        // 531k-cell patch example steps 1 % faster than with 256)
        const uint64_t t = env_u64("EFS_POINT_THREADS", 64);
        c->pt = (t == 128 || t == 256) ? (uint32_t)t : 64u;
    }
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    // the no-alloc far field's scratch; on failure only that entry point
    // reports it (c->scratch stays null, scratch_err keeps the code)
    if (e == hipSuccess) c->scratch_err = scratch_address(device, &c->scratch);
    if (e != hipSuccess) {
        g_create_error = std::string("HIP setup failed: ") + hipGetErrorString(e);
        if (c->stream) hipStreamDestroy(c->stream);
        delete c;
        return EFS_ERR_HIP;
    }
    *out = c;
    return EFS_OK;
}

void efs_destroy(efs_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    free_all(c);
    for (auto& t : c->pending) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    for (auto e : c->event_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* efs_last_error(const efs_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int efs_init(efs_ctx* c, uint32_t nx, uint32_t ny, uint32_t nz, const float* vv, const float* vi,
             const float* ii, const float* iv) {
    if (!c) return EFS_ERR_STATE;
    if (!nx || !ny || !nz) return fail(c, EFS_ERR_ARG, "grid size must be >= 1 in every direction");
    const uint64_t N = (uint64_t)nx * ny * nz;
    if (3 * N >= (1ull << 31)) return fail(c, EFS_ERR_ARG, "grid too large: 3*Nx*Ny*Nz must be < 2^31");
    if (!vv || !vi || !ii || !iv) return fail(c, EFS_ERR_ARG, "coefficient arrays vv, vi, ii, iv are required");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    free_all(c);
    reset_extensions(c);
    c->inited = false;
    c->nx = nx; c->ny = ny; c->nz = nz; c->N = N;
    c->nzp = env_u64("EFS_VEC", 4) == 4 ? (nz + 3) / 4 * 4 : nz;
    c->Np = (uint64_t)nx * ny * c->nzp;
    const uint64_t Np = c->Np;
    // Device layout: component stride S = N rounded up to 256 floats plus a
    // pad, the six arrays in one pool with a stagger between them, so the
    // concurrent streams of a sweep are not spaced at powers of two.
    const uint64_t pad = env_u64("EFS_PAD", 1088);
    const uint64_t stagger = env_u64("EFS_STAGGER", 4160);
    c->S = ((Np + 255) / 256) * 256 + (pad / 4) * 4;
    if (3 * c->S >= (1ull << 31)) return fail(c, EFS_ERR_ARG, "grid too large for 32-bit device indices");
    const uint64_t arr = 3 * c->S + (stagger / 64) * 64;
    c->dbl = c->storage_req == EFS_STORAGE_F64;
    const uint64_t farr = c->dbl ? 2 * arr : arr;  // the two field arrays (double: twice the floats)
    TRY(dalloc(c, c->pool, 2 * farr + 4 * arr));
    HIP_TRY(c, hipMemset(c->pool, 0, (2 * farr + 4 * arr) * sizeof(float)));
    c->volt = c->pool;
    c->curr = c->pool + farr;
    c->field_floats = 2 * farr;
    c->vv = c->pool + 2 * farr;
    c->vi = c->vv + arr;
    c->ii = c->vv + 2 * arr;
    c->iv = c->vv + 3 * arr;
    // coefficients in the device row layout: padded copies when nzp > nz,
    // with the last real z plane of the H side set to ii = 1, iv = 0 (those
    // cells keep I, the reference's last-plane rule, src/webgpu-engine.mjs:88)
    std::vector<float> padded[4];
    const float* src[4] = {vv, vi, ii, iv};
    if (c->nzp != nz) {
        for (int a = 0; a < 4; ++a) {
            padded[a].assign(3 * Np, 0.0f);
            for (uint32_t n = 0; n < 3; ++n)
                for (uint64_t xy = 0; xy < (uint64_t)nx * ny; ++xy) {
                    float* d = padded[a].data() + n * Np + xy * c->nzp;
                    std::memcpy(d, src[a] + n * N + xy * nz, nz * sizeof(float));
                    if (a == 2) d[nz - 1] = 1.0f;
                    if (a == 3) d[nz - 1] = 0.0f;
                }
            src[a] = padded[a].data();
        }
    }
    float* dst[4] = {c->vv, c->vi, c->ii, c->iv};
    for (int a = 0; a < 4; ++a)
        for (int n = 0; n < 3; ++n)
            HIP_TRY(c, hipMemcpy(dst[a] + n * c->S, src[a] + n * Np, Np * sizeof(float), hipMemcpyHostToDevice));
    // FP64 storage runs the 1-wide sweeps (on the same padded rows)
    c->vec4 = !c->dbl && c->nzp % 4 == 0 && env_u64("EFS_VEC", 4) == 4;
    c->half = c->storage_req == EFS_STORAGE_F16;
    c->range_interval = (uint32_t)std::max<uint64_t>(1, env_u64("EFS_RANGE_INTERVAL", 32));
    c->fuse_hooks = env_u64("EFS_FUSE_HOOKS", 1) != 0;
    {
        // side work: 0 off, 1 on, 2 (default) on for grids of at most
        // kSideMaxCells cells, whose steps are short enough for the saved
        // launches to matter
        const uint64_t m = env_u64("EFS_SIDE_WORK", 2);
        c->side_on = m == 1 || (m == 2 && c->Np <= kSideMaxCells);
    }
    c->side_pend = false;
    if (c->half && !c->vec4) return fail(c, EFS_ERR_ARG, "FP16 field storage needs the 4-wide sweeps (EFS_VEC=4)");
    {
        const float one[6] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
        TRY(dupload(c, c->fscale, one, 6));
        TRY(dzero(c, c->f_amax, 6));
        TRY(dupload(c, c->f_factor, one, 6));
        TRY(dupload(c, c->f_smin, one, 6));
    }
    // coefficient compression (4-wide sweeps): both fields, all components.
    // EFS_COMPRESS: 1 (default) byte form where it fits, else 16-bit; 16 the
    // 16-bit form even where the byte form fits; 0 FP32 arrays only.
    const uint64_t cmode = env_u64("EFS_COMPRESS", 1);
    c->comp = false;
    if (c->vec4 && cmode != 0 && cmode != 16) {
        std::vector<uint8_t> ie(3 * c->S, 0), ih(3 * c->S, 0);
        std::vector<float2> te(3 * 256), th(3 * 256);
        bool ok = true;
        for (int n = 0; n < 3 && ok; ++n) {
            ok = pair_table(src[0] + n * Np, src[1] + n * Np, Np, ie.data() + n * c->S, te.data() + n * 256) &&
                 pair_table(src[2] + n * Np, src[3] + n * Np, Np, ih.data() + n * c->S, th.data() + n * 256);
        }
        if (ok) {
            TRY(dupload(c, c->cidx_e, ie.data(), ie.size()));
            TRY(dupload(c, c->cidx_h, ih.data(), ih.size()));
            TRY(dupload(c, c->ctab_e, te.data(), te.size()));
            TRY(dupload(c, c->ctab_h, th.data(), th.size()));
            c->comp = true;
        }
    }
    // more than 256 pairs somewhere: 16-bit indices (six tables built on six
    // threads), up to 65,536 pairs per component
    c->comp16 = false;
    if (c->vec4 && !c->comp && cmode != 0) {
        std::vector<uint16_t> ie(3 * c->S, 0), ih(3 * c->S, 0);
        std::vector<float2> tabs[6];
        bool ok[6];
        std::vector<std::thread> th;
        for (int k = 0; k < 6; ++k)
            th.emplace_back([&, k] {
                const int n = k % 3, side = k / 3;
                ok[k] = pair_table16(src[2 * side] + n * Np, src[2 * side + 1] + n * Np, Np,
                                     (side ? ih : ie).data() + n * c->S, tabs[k]);
            });
        for (auto& t : th) t.join();
        if (ok[0] && ok[1] && ok[2] && ok[3] && ok[4] && ok[5]) {
            for (int side = 0; side < 2; ++side) {
                size_t m = 1;
                for (int n = 0; n < 3; ++n) m = std::max(m, tabs[3 * side + n].size());
                const uint32_t tabn = (uint32_t)((m + 63) & ~(size_t)63);
                std::vector<float2> t(3 * (size_t)tabn, make_float2(0.f, 0.f));
                for (int n = 0; n < 3; ++n) std::copy(tabs[3 * side + n].begin(), tabs[3 * side + n].end(),
                                                      t.begin() + (size_t)n * tabn);
                TRY(dupload(c, side ? c->cidx16_h : c->cidx16_e, (side ? ih : ie).data(), ie.size()));
                TRY(dupload(c, side ? c->ctab16_h : c->ctab16_e, t.data(), t.size()));
                (side ? c->tabn_h : c->tabn_e) = tabn;
            }
            c->comp16 = true;
        }
    }
    TRY(dalloc(c, c->en_partial, kEnergyMaxBlocks + 1));
    TRY(plan_sweeps(c));
    c->numTS = 0;
    c->fields_zero = true;
    c->inited = true;
    return EFS_OK;
}

// FP16 storage from zero fields: start the scales where a source puts its
// largest value near 2^kRangeTarget -- volt from the voltage amplitudes,
// curr from the current amplitudes or, without those, 2^8 below volt
// (about 1/Z0).  These are also the floors of the range control.
int source_scales(efs_ctx* c, const float* signal, uint32_t len, const float* vamp, uint32_t nv, const float* iamp,
                  uint32_t ni) {
    if (!c->half || !c->fields_zero) return EFS_OK;
    float smax = 0.f, va = 0.f, ia = 0.f;
    for (uint32_t k = 0; k < len; ++k) smax = std::max(smax, std::fabs(signal[k]));
    for (uint32_t k = 0; k < nv; ++k) va = std::max(va, std::fabs(vamp[k]));
    for (uint32_t k = 0; k < ni; ++k) ia = std::max(ia, std::fabs(iamp[k]));
    const float mv = smax * va, mi = smax * ia;
    if (!(mv > 0.f && std::isfinite(mv)) && !(mi > 0.f && std::isfinite(mi))) return EFS_OK;
    float sc[6];
    HIP_TRY(c, hipMemcpy(sc, c->fscale, sizeof sc, hipMemcpyDeviceToHost));
    const float vs = mv > 0.f && std::isfinite(mv) ? std::ldexp(1.0f, std::ilogb(mv) - kRangeTarget)
                                                   : std::ldexp(1.0f, std::ilogb(mi) - kRangeTarget + 8);
    const float is = mi > 0.f && std::isfinite(mi) ? std::ldexp(1.0f, std::ilogb(mi) - kRangeTarget)
                                                   : std::ldexp(vs, -8);
    for (int k = 0; k < 3; ++k) {
        sc[k] = vs;
        sc[3 + k] = is;
    }
    HIP_TRY(c, hipMemcpy(c->fscale, sc, sizeof sc, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->f_smin, sc, sizeof sc, hipMemcpyHostToDevice));
    return EFS_OK;
}

int efs_cfg_excitation(efs_ctx* c, const float* signal, uint32_t len, const uint32_t* delay, const float* amp,
                       const uint32_t* dir, const uint32_t* pos, uint32_t n, uint32_t period) {
    TRY(need_init(c));
    c->pad_dirty = true;
    if (n == 0) {
        c->exc_n = 0;
        c->h_exc_gidx.clear();
        c->alias_dirty = true;
        return EFS_OK;
    }
    if (!signal || len == 0) return fail(c, EFS_ERR_ARG, "excitation signal must not be empty");
    if (!delay || !amp || !dir || !pos) return fail(c, EFS_ERR_ARG, "excitation arrays are required");
    std::vector<uint32_t> g(n);
    for (uint32_t k = 0; k < n; ++k) {
        if (dir[k] > 2) return fail(c, EFS_ERR_ARG, "excitation %u: dir %u out of range", k, dir[k]);
        if (pos[k] >= c->N) return fail(c, EFS_ERR_ARG, "excitation %u: pos %u out of range", k, pos[k]);
        g[k] = tr(c, (uint64_t)dir[k] * c->N + pos[k]);
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    TRY(dupload(c, c->exc_signal, signal, len));
    TRY(dupload(c, c->exc_delay, delay, n));
    TRY(dupload(c, c->exc_amp, amp, n));
    TRY(dupload(c, c->exc_gidx, g.data(), n));
    c->h_exc_gidx.resize(n);
    for (uint32_t k = 0; k < n; ++k) c->h_exc_gidx[k] = dir[k] * (uint32_t)c->N + pos[k];
    c->alias_dirty = true;  // Mur entries reading an excitation cell post early
    c->exc_n = n;
    c->exc_len = len;
    c->exc_period = period;
    for (uint32_t k = 0; k < n; ++k) c->source_end = std::max<uint64_t>(c->source_end, (uint64_t)delay[k] + len);
    if (period) c->source_periodic = true;
    TRY(source_scales(c, signal, len, amp, n, nullptr, 0));
    return EFS_OK;
}

int efs_cfg_pml(efs_ctx* c, uint32_t nreg, const uint32_t* start, const uint32_t* num, const float* vv,
                const float* vvfo, const float* vvfn, const float* ii, const float* iifo, const float* iifn) {
    TRY(need_init(c));
    c->pad_dirty = true;
    if (nreg > (uint32_t)kMaxPml) return fail(c, EFS_ERR_ARG, "at most %d UPML regions are supported", kMaxPml);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    PmlTable t{};
    std::vector<uint32_t> tz0(nreg), tz1(nreg);  // the regions' z-bounds as given
    for (uint32_t k = 0; k < nreg; ++k) {
        const uint32_t* s = start + 3 * k;
        const uint32_t* m = num + 3 * k;
        const uint32_t dims[3] = {c->nx, c->ny, c->nz};
        for (int d = 0; d < 3; ++d)
            if (m[d] == 0 || (uint64_t)s[d] + m[d] > dims[d])
                return fail(c, EFS_ERR_ARG, "UPML region %u exceeds the grid in direction %d", k, d);
        PmlRegion& R = t.r[k];
        R.x0 = s[0]; R.x1 = s[0] + m[0];
        R.y0 = s[1]; R.y1 = s[1] + m[1];
        R.z0 = s[2]; R.z1 = s[2] + m[2];
        tz0[k] = R.z0;
        tz1[k] = R.z1;
        for (uint32_t j = 0; j < k; ++j)
            if (box_overlap(t.r[j], R)) return fail(c, EFS_ERR_ARG, "UPML regions %u and %u overlap", j, k);
    }
    // The 4-wide sweeps need region z-bounds on 4-cell lane boundaries.
    // openEMS's PML_n slabs are n + 1 lines deep, so they usually are not;
    // such regions are widened to the enclosing lane boundaries (within the
    // padded rows) when the widened boxes still do not overlap.  A padding
    // cell carries the UPML coefficients (0, 0, 1): its update is then
    // V' = F' = a*F + b*curl, and with its flux kept equal to its field
    // (k_pml_pad_sync after configuration and uploads) that is the core
    // update to the bit.  Reads of the flux (alias indices) use the given
    // bounds, so padding cells read the field.
    bool widen = false;
    if (c->vec4 && env_u64("EFS_PML_WIDEN", 1) != 0) {
        PmlTable w = t;
        bool need = false;
        for (uint32_t k = 0; k < nreg; ++k) {
            PmlRegion& R = w.r[k];
            const uint32_t z0 = R.z0 / 4 * 4, z1 = std::min((R.z1 + 3) / 4 * 4, c->nzp);
            need = need || z0 != R.z0 || z1 != R.z1;
            R.z0 = z0;
            R.z1 = z1;
        }
        bool ok = need;
        for (uint32_t k = 0; k < nreg && ok; ++k)
            for (uint32_t j = 0; j < k && ok; ++j) ok = !box_overlap(w.r[j], w.r[k]);
        if (ok) {
            t = w;
            widen = true;
        }
    }
    uint64_t total = 0;
    bool aligned = true;
    for (uint32_t k = 0; k < nreg; ++k) {
        PmlRegion& R = t.r[k];
        R.ny = R.y1 - R.y0;
        R.nz = R.z1 - R.z0;
        R.cs = (R.x1 - R.x0) * R.ny * R.nz;
        total = (total + 3) / 4 * 4;  // 16-byte aligned region bases
        R.flux = total;
        R.coef = 6 * total;
        total += 3 * (uint64_t)R.cs;
        aligned = aligned && (R.z0 % 4 == 0) && (R.nz % 4 == 0);
    }
    if (total >= (1ull << 31)) return fail(c, EFS_ERR_ARG, "UPML volume too large");
    // UPML launch geometry: per region, 256-lane tiles over its (y, z/4)
    // plane and x chunks; the launch's workgroups are shared out in
    // proportion to region volume
    const uint64_t want = env_u64("EFS_PML_BLOCKS", 512);
    uint32_t blocks = 0;
    for (uint32_t k = 0; k < nreg; ++k) {
        PmlRegion& R = t.r[k];
        const uint32_t lanes = R.ny * ((R.nz + 3) / 4);
        R.tiles = (lanes + kSweepThreads - 1) / kSweepThreads;
        const uint32_t nxr = R.x1 - R.x0;
        uint32_t chunks = std::max(1u, std::min(nxr, (uint32_t)(want * R.cs / std::max<uint64_t>(total / 3, 1) / R.tiles)));
        R.chunk_len = (nxr + chunks - 1) / chunks;
        chunks = (nxr + R.chunk_len - 1) / R.chunk_len;
        R.blocks = R.tiles * chunks;
        R.blk0 = blocks;
        blocks += R.blocks;
    }
    t.n = nreg;
    t.blocks = blocks;
    // the six coefficient arrays in the engine's region layout (padding
    // cells of widened regions: identity coefficients 0, 0, 1)
    std::vector<float> coef;
    std::vector<uint32_t> pad_g, pad_f;  // padding cells: field (device index) and flux index
    if (nreg) {
        if (!vv || !vvfo || !vvfn || !ii || !iifo || !iifn)
            return fail(c, EFS_ERR_ARG, "UPML coefficient arrays are required");
        coef.assign(6 * total, 0.f);
        const float* src[6] = {vv, vvfo, vvfn, ii, iifo, iifn};
        const float ident[3] = {0.f, 0.f, 1.f};
        uint64_t off_in = 0;
        for (uint32_t k = 0; k < nreg; ++k) {
            const PmlRegion& R = t.r[k];
            const uint64_t T = 3 * (uint64_t)R.cs;
            const uint32_t nzg = tz1[k] - tz0[k];  // the caller's row length
            const uint64_t csg = (uint64_t)(R.x1 - R.x0) * R.ny * nzg;
            for (int a = 0; a < 6; ++a)
                for (uint32_t n = 0; n < 3; ++n)
                    for (uint32_t x = 0; x < R.x1 - R.x0; ++x)
                        for (uint32_t y = 0; y < R.ny; ++y) {
                            float* d = coef.data() + R.coef + a * T + n * (uint64_t)R.cs + ((uint64_t)x * R.ny + y) * R.nz;
                            const float* sr = src[a] + off_in + n * csg + ((uint64_t)x * R.ny + y) * nzg;
                            for (uint32_t z = R.z0; z < R.z1; ++z)
                                d[z - R.z0] = (z >= tz0[k] && z < tz1[k]) ? sr[z - tz0[k]] : ident[a % 3];
                        }
            off_in += 3 * csg;
            if (widen)
                for (uint32_t n = 0; n < 3; ++n)
                    for (uint32_t x = R.x0; x < R.x1; ++x)
                        for (uint32_t y = R.y0; y < R.y1; ++y)
                            for (uint32_t z = R.z0; z < std::min(R.z1, c->nz); ++z) {
                                if (z >= tz0[k] && z < tz1[k]) continue;
                                pad_g.push_back((uint32_t)(n * c->S + ((uint64_t)x * c->ny + y) * c->nzp + z));
                                pad_f.push_back((uint32_t)(R.flux + n * (uint64_t)R.cs +
                                                           ((uint64_t)(x - R.x0) * R.ny + (y - R.y0)) * R.nz + (z - R.z0)));
                            }
        }
        TRY(dupload(c, c->pml_coef, coef.data(), coef.size()));
        TRY(dzero(c, c->pml_vflux, total * rw(c)));
        TRY(dzero(c, c->pml_cflux, total * rw(c)));
        c->flux_floats = total * rw(c);
    } else {
        dfree(c->pml_coef); dfree(c->pml_vflux); dfree(c->pml_cflux);
        c->flux_floats = 0;
    }
    c->pml_tz0 = tz0;
    c->pml_tz1 = tz1;
    c->pml_pad_n = (uint32_t)pad_g.size();
    c->h_pad.resize(pad_g.size());
    for (size_t q = 0; q < pad_g.size(); ++q) c->h_pad[q] = ((uint64_t)pad_g[q] << 32) | pad_f[q];
    std::sort(c->h_pad.begin(), c->h_pad.end());
    if (c->pml_pad_n) {
        TRY(dupload(c, c->pml_pad_g, pad_g.data(), pad_g.size()));
        TRY(dupload(c, c->pml_pad_f, pad_f.data(), pad_f.size()));
    }
    // UPML coefficient compression: one table per field and component over
    // all regions, byte indices per region cell (region-local NIJK, like the
    // FP32 arrays); used by the fused sweeps when every table fits.
    dfree(c->pml_idx); dfree(c->pml_tab_e); dfree(c->pml_tab_h);
    c->pml_comp = false;
    if (nreg && aligned && env_u64("EFS_PML_COMPRESS", 1) != 0) {
        uint64_t ibytes = 0;
        for (uint32_t k = 0; k < nreg; ++k) {
            t.r[k].pidx = ibytes;
            ibytes += 6 * (uint64_t)t.r[k].cs;  // multiple of 4: region nz % 4 == 0
        }
        std::vector<uint8_t> idx(ibytes);
        TripleTable tabs[2][3];
        bool ok = true;
        for (uint32_t k = 0; k < nreg && ok; ++k) {
            const uint64_t cs = t.r[k].cs, T = 3 * cs;
            const float* base = coef.data() + t.r[k].coef;  // the engine-layout arrays built above
            for (int h = 0; h < 2 && ok; ++h)
                for (int n = 0; n < 3 && ok; ++n)
                    ok = tabs[h][n].add(base + (3 * h) * T + n * cs, base + (3 * h + 1) * T + n * cs,
                                        base + (3 * h + 2) * T + n * cs, cs,
                                        idx.data() + t.r[k].pidx + (3 * h + n) * cs);
        }
        if (ok) {
            std::vector<float4> te(3 * 256), th(3 * 256);
            for (int n = 0; n < 3; ++n)
                for (int q = 0; q < 256; ++q) {
                    te[256 * n + q] = tabs[0][n].tab[q];
                    th[256 * n + q] = tabs[1][n].tab[q];
                }
            TRY(dupload(c, c->pml_idx, idx.data(), idx.size()));
            TRY(dupload(c, c->pml_tab_e, te.data(), te.size()));
            TRY(dupload(c, c->pml_tab_h, th.data(), th.size()));
            c->pml_comp = true;
        }
    }
    c->pml = t;
    c->pml_split = aligned && env_u64("EFS_PML_SPLIT", 1) != 0;
    c->pml_fused = env_u64("EFS_PML_FUSED", 1) != 0;
    if (c->half && nreg && !(c->pml_split && c->pml_fused)) {
        c->pml = PmlTable{};
        return fail(c, EFS_ERR_ARG, "FP16 field storage needs UPML regions the 4-wide sweeps can take "
                                    "(z-bounds on 4-cell boundaries after widening)");
    }
    TRY(pml_pad_sync(c));
    TRY(plan_sweeps(c));
    c->alias_dirty = true;
    return EFS_OK;
}

int efs_cfg_mur(efs_ctx* c, uint32_t n, const float* coeff, const uint32_t* normal_idx,
                const uint32_t* shifted_idx) {
    TRY(need_init(c));
    c->pad_dirty = true;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (n == 0) {
        c->mur_n = 0;
        return EFS_OK;
    }
    if (!coeff || !normal_idx || !shifted_idx) return fail(c, EFS_ERR_ARG, "Mur arrays are required");
    for (uint32_t k = 0; k < n; ++k)
        if (normal_idx[k] >= 3 * c->N || shifted_idx[k] >= 3 * c->N)
            return fail(c, EFS_ERR_ARG, "Mur point %u: index out of range", k);
    // keep each normal cell's last entry (the reference's loop lets it win),
    // ordered by the cell so that neighbouring threads touch neighbouring cells
    std::vector<uint32_t> ord(n);
    for (uint32_t k = 0; k < n; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return normal_idx[a] < normal_idx[b]; });
    std::vector<uint32_t> hn, hs;
    std::vector<float> hc;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t k = ord[j];
        if (j + 1 < n && normal_idx[ord[j + 1]] == normal_idx[k]) continue;
        hn.push_back(normal_idx[k]);
        hs.push_back(shifted_idx[k]);
        hc.push_back(coeff[k]);
    }
    const uint32_t m = (uint32_t)hn.size();
    c->h_mur_normal = hn;
    c->h_mur_shifted = hs;
    c->h_mur_coeff = hc;
    TRY(dupload(c, c->mur_coeff, hc.data(), m));
    const std::vector<uint32_t> tn = tr_all(c, hn.data(), m), ts = tr_all(c, hs.data(), m);
    TRY(dupload(c, c->mur_normal, tn.data(), m));
    TRY(dupload(c, c->mur_shifted, ts.data(), m));
    TRY(dzero(c, c->mur_saved, m * rw(c)));
    c->mur_n = m;
    c->alias_dirty = true;
    return EFS_OK;
}

int efs_cfg_ade_clear(efs_ctx* c) {
    TRY(need_init(c));
    c->pad_dirty = true;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->h_ade_gidx.clear();
    c->h_ade_lor.clear();
    for (auto& v : c->h_ade_c) v.clear();
    c->ade_n = 0;
    c->ade_dirty = false;
    c->alias_dirty = true;
    c->cap_dirty = true;
    c->fuse_dirty = true;
    return EFS_OK;
}

// Extensions the FP64 storage form does not run (it covers the core sweeps,
// UPML, Mur, excitation, gathers, NF2FF and energy).
static int need_not_f64(efs_ctx* c, uint32_t n, const char* what) {
    if (c->dbl && n) return fail(c, EFS_ERR_ARG, "%s is not supported with FP64 field storage", what);
    return EFS_OK;
}

int efs_cfg_ade_add(efs_ctx* c, uint32_t has_lorentz, uint32_t dir, uint32_t ncells, const uint32_t* pos_idx,
                    const float* v_int, const float* v_ext, const float* v_lor, const float* i_int,
                    const float* i_ext, const float* i_lor) {
    TRY(need_init(c));
    c->pad_dirty = true;

    if (dir > 2) return fail(c, EFS_ERR_ARG, "ADE direction %u out of range", dir);
    if (ncells && !pos_idx) return fail(c, EFS_ERR_ARG, "ADE pos_idx is required");
    for (uint32_t k = 0; k < ncells; ++k)
        if (pos_idx[k] >= c->N) return fail(c, EFS_ERR_ARG, "ADE cell %u: pos_idx out of range", k);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const float* src[6] = {v_int, v_ext, v_lor, i_int, i_ext, i_lor};
    for (uint32_t k = 0; k < ncells; ++k) {
        c->h_ade_gidx.push_back(dir * (uint32_t)c->N + pos_idx[k]);
        c->h_ade_lor.push_back(has_lorentz ? 1 : 0);
        for (int a = 0; a < 6; ++a) c->h_ade_c[a].push_back(src[a] ? src[a][k] : 0.f);
    }
    c->ade_dirty = true;
    c->cap_dirty = true;
    c->fuse_dirty = true;
    return EFS_OK;
}

int efs_set_ade_apply(efs_ctx* c, int apply) {
    TRY(need_init(c));
    c->pad_dirty = true;
    c->ade_apply = apply != 0;
    c->alias_dirty = true;
    c->cap_dirty = true;
    c->fuse_dirty = true;
    return EFS_OK;
}

int efs_cfg_tfsf(efs_ctx* c, const float* signal, uint32_t len, uint32_t period, uint32_t nv,
                 const uint32_t* v_idx, const uint32_t* v_dint, const float* v_dfrac, const float* v_amp,
                 uint32_t ni, const uint32_t* i_idx, const uint32_t* i_dint, const float* i_dfrac,
                 const float* i_amp) {
    TRY(need_init(c));
    c->pad_dirty = true;

    if ((nv || ni) && (!signal || len == 0)) return fail(c, EFS_ERR_ARG, "TFSF signal must not be empty");
    for (uint32_t k = 0; k < nv; ++k)
        if (v_idx[k] >= 3 * c->N) return fail(c, EFS_ERR_ARG, "TFSF voltage point %u out of range", k);
    for (uint32_t k = 0; k < ni; ++k)
        if (i_idx[k] >= 3 * c->N) return fail(c, EFS_ERR_ARG, "TFSF current point %u out of range", k);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (nv || ni) TRY(dupload(c, c->tfsf_signal, signal, len));
    if (nv) {
        const std::vector<uint32_t> t = tr_all(c, v_idx, nv);
        TRY(dupload(c, c->tv_idx, t.data(), nv));
        TRY(dupload(c, c->tv_dint, v_dint, nv));
        TRY(dupload(c, c->tv_dfrac, v_dfrac, nv));
        TRY(dupload(c, c->tv_amp, v_amp, nv));
    }
    if (ni) {
        const std::vector<uint32_t> t = tr_all(c, i_idx, ni);
        TRY(dupload(c, c->ti_idx, t.data(), ni));
        TRY(dupload(c, c->ti_dint, i_dint, ni));
        TRY(dupload(c, c->ti_dfrac, i_dfrac, ni));
        TRY(dupload(c, c->ti_amp, i_amp, ni));
    }
    c->tfsf_len = len;
    c->tfsf_period = period;
    for (uint32_t k = 0; k < nv; ++k) c->source_end = std::max<uint64_t>(c->source_end, (uint64_t)v_dint[k] + len);
    for (uint32_t k = 0; k < ni; ++k) c->source_end = std::max<uint64_t>(c->source_end, (uint64_t)i_dint[k] + len);
    if (period && (nv || ni)) c->source_periodic = true;
    c->tfsf_nv = nv;
    c->tfsf_ni = ni;
    c->h_ti_idx.assign(i_idx, i_idx + ni);
    c->h_tv_idx.assign(v_idx, v_idx + nv);
    c->cap_dirty = true;
    c->fuse_dirty = true;
    if (nv || ni) TRY(source_scales(c, signal, len, v_amp, nv, i_amp, ni));
    return EFS_OK;
}

int efs_cfg_rlc(efs_ctx* c, uint32_t n, const uint32_t* field_idx, const uint32_t* direction,
                const uint32_t* type_flag, const float* coeffs) {
    TRY(need_init(c));
    c->pad_dirty = true;
    TRY(need_not_f64(c, n, "RLC"));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (n == 0) {
        c->rlc_n = 0;
        c->h_rlc_dev.clear();
        return EFS_OK;
    }
    std::vector<uint32_t> g(n);
    for (uint32_t k = 0; k < n; ++k) {
        if (field_idx[k] >= c->N || direction[k] > 2 || type_flag[k] > 1)
            return fail(c, EFS_ERR_ARG, "RLC element %u: field_idx/direction/type_flag out of range", k);
        g[k] = tr(c, (uint64_t)direction[k] * c->N + field_idx[k]);
    }
    TRY(dupload(c, c->rlc_gidx, g.data(), n));
    c->h_rlc_dev = g;
    TRY(dupload(c, c->rlc_type, type_flag, n));
    TRY(dupload(c, c->rlc_c, coeffs, 9 * (size_t)n));
    TRY(dzero(c, c->rlc_vdn, 3 * (size_t)n));
    TRY(dzero(c, c->rlc_jn, 3 * (size_t)n));
    TRY(dzero(c, c->rlc_vil, n));
    c->rlc_n = n;
    return EFS_OK;
}

int efs_cfg_pbc(efs_ctx* c, uint32_t n, const uint32_t* axis, const double* phase) {
    TRY(need_init(c));
    c->pad_dirty = true;
    TRY(need_not_f64(c, n, "PBC"));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (n > 3) return fail(c, EFS_ERR_ARG, "at most 3 periodic axes");
    if (n && (!axis || !phase)) return fail(c, EFS_ERR_ARG, "PBC axes and phases are required");
    const uint32_t dims[3] = {c->nx, c->ny, c->nz};
    std::vector<PbcAxis> ve, vh;
    uint32_t imag = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (axis[k] > 2) return fail(c, EFS_ERR_ARG, "PBC axis %u out of range", axis[k]);
        const uint32_t ax = axis[k];
        if (dims[ax] < 2) return fail(c, EFS_ERR_ARG, "PBC axis %u needs at least 2 grid lines", ax);
        PbcAxis a{};
        a.axis = ax;
        a.t0 = (ax + 1) % 3;
        a.t1 = (ax + 2) % 3;
        const uint32_t f0 = dims[a.t0], f1 = dims[a.t1];
        a.f1 = f1;
        a.plane = f0 * f1;
        a.cs = (float)std::cos(phase[k]);  // the engine's f32 uniform (src/webgpu-engine.mjs:2091-2092)
        a.sn = (float)std::sin(phase[k]);
        a.has_imag = a.sn != 0.0f;
        a.imag_off = imag;
        if (a.has_imag) imag += 4 * a.plane;
        PbcAxis e = a, h = a;
        e.src = dims[ax] - 2; e.dst = 0; e.b0 = f0; e.b1 = f1;                       // E: plane N-2 -> plane 0
        h.src = 0; h.dst = dims[ax] - 2; h.b0 = f0 - 1; h.b1 = f1 - 1;               // H: plane 0 -> plane N-2
        h.imag_off += 2 * a.plane;
        ve.push_back(e);
        if (h.b0 && h.b1) vh.push_back(h);
    }
    TRY(dzero(c, c->pbc_imag, imag));
    c->pbc_e = ve;
    c->pbc_h = vh;
    return EFS_OK;
}

int efs_cfg_steady_state(efs_ctx* c, const uint32_t* probe_idx, uint32_t n, uint32_t period_samples) {
    TRY(need_init(c));
    TRY(need_not_f64(c, n, "steady-state detection"));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (n && !probe_idx) return fail(c, EFS_ERR_ARG, "steady-state probe indices are required");
    for (uint32_t k = 0; k < n; ++k)
        if (probe_idx[k] >= 3 * c->N) return fail(c, EFS_ERR_ARG, "steady-state probe %u: index out of range", k);
    const std::vector<uint32_t> g = tr_all(c, probe_idx, n);
    TRY(dupload(c, c->ss_gidx, g.data(), n));
    TRY(dzero(c, c->ss_e, 2 * (size_t)n));
    c->ss_n = n;
    c->ss_period = period_samples;
    c->ss_current = 0;
    c->ss_recording = false;
    return EFS_OK;
}

int efs_ss_set(efs_ctx* c, int recording, uint32_t current_sample) {
    TRY(need_init(c));
    c->ss_recording = recording != 0;
    c->ss_current = current_sample;
    return EFS_OK;
}

int efs_ss_read(efs_ctx* c, float* e1, float* e2, uint32_t* current_sample, int* recording) {
    TRY(need_init(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->ss_n && e1) HIP_TRY(c, hipMemcpy(e1, c->ss_e, c->ss_n * sizeof(float), hipMemcpyDeviceToHost));
    if (c->ss_n && e2)
        HIP_TRY(c, hipMemcpy(e2, c->ss_e + c->ss_n, c->ss_n * sizeof(float), hipMemcpyDeviceToHost));
    if (current_sample) *current_sample = c->ss_current;
    if (recording) *recording = c->ss_recording ? 1 : 0;
    return EFS_OK;
}

int efs_iterate(efs_ctx* c, uint32_t n) {
    TRY(need_init(c));
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->ade_dirty) TRY(rebuild_ade(c));
    if (c->alias_dirty) TRY(rebuild_aliases(c));
    if (c->pad_dirty) TRY(rebuild_pad_hooks(c));
    if (c->cap_dirty) {
        c->cap_clash = hooks_clash(c);
        c->cap_dirty = false;
    }
    if (n) c->fields_zero = false;
    c->mur_pre_ahead = false;
    c->ade_ahead = false;
    if (c->fuse_dirty) plan_point_fusion(c);
    for (uint32_t s = 0; s < n; ++s) TRY(launch_step(c, s + 1 < n));
    return side_flush(c);  // no sample or gather stays half done between calls
}

// The E and H sweeps of one step and nothing else (no point work, no
// NF2FF capture), on the 4-wide fused path: what efs_tune_sweeps times.
int launch_sweep_pair(efs_ctx* c) {
    const bool has_pml = c->pml.n > 0, x16 = c->comp16;
    const Coef ce{c->vv, c->vi, c->cidx_e, c->comp16 ? c->ctab16_e : c->ctab_e, c->pml_idx, c->pml_tab_e,
                  c->cidx16_e, c->tabn_e};
    const Coef ch{c->ii, c->iv, c->cidx_h, c->comp16 ? c->ctab16_h : c->ctab_h, c->pml_idx, c->pml_tab_h,
                  c->cidx16_h, c->tabn_h};
    const dim3 grid(c->geom.blocks), blk(kSweepThreads);
    if (c->half) {
        auto ke = has_pml ? (c->pml_comp ? (x16 ? sweep_e4fxp_h : c->comp ? sweep_e4fcp_h : sweep_e4fp_h)
                                         : (x16 ? sweep_e4fx_h : c->comp ? sweep_e4fc_h : sweep_e4f_h))
                          : (x16 ? sweep_e4x_h : c->comp ? sweep_e4c_h : sweep_e4_h);
        auto kh = has_pml ? (c->pml_comp ? (x16 ? sweep_h4fxp_h : c->comp ? sweep_h4fcp_h : sweep_h4fp_h)
                                         : (x16 ? sweep_h4fx_h : c->comp ? sweep_h4fc_h : sweep_h4f_h))
                          : (x16 ? sweep_h4x_h : c->comp ? sweep_h4c_h : sweep_h4_h);
        hipLaunchKernelGGL(ke, grid, blk, 0, c->stream, (f16*)c->volt, (const f16*)c->curr, ce, c->geom, c->pml,
                           c->pml_coef, c->pml_vflux, c->fscale, NfCapture{}, SideWork{});
        hipLaunchKernelGGL(kh, grid, blk, 0, c->stream, (f16*)c->curr, (const f16*)c->volt, ch, c->geom, c->pml,
                           c->pml_coef, c->pml_cflux, c->fscale, NfCapture{}, SideWork{});
    } else {
        auto ke = has_pml ? (c->pml_comp ? (x16 ? sweep_e4fxp : c->comp ? sweep_e4fcp : sweep_e4fp)
                                         : (x16 ? sweep_e4fx : c->comp ? sweep_e4fc : sweep_e4f))
                          : (x16 ? sweep_e4x : c->comp ? sweep_e4c : sweep_e4);
        auto kh = has_pml ? (c->pml_comp ? (x16 ? sweep_h4fxp : c->comp ? sweep_h4fcp : sweep_h4fp)
                                         : (x16 ? sweep_h4fx : c->comp ? sweep_h4fc : sweep_h4f))
                          : (x16 ? sweep_h4x : c->comp ? sweep_h4c : sweep_h4);
        hipLaunchKernelGGL(ke, grid, blk, 0, c->stream, c->volt, c->curr, ce, c->geom, c->pml, c->pml_coef,
                           c->pml_vflux, c->fscale, NfCapture{}, SideWork{});
        hipLaunchKernelGGL(kh, grid, blk, 0, c->stream, c->curr, c->volt, ch, c->geom, c->pml, c->pml_coef,
                           c->pml_cflux, c->fscale, NfCapture{}, SideWork{});
    }
    CHECK_LAUNCH(c);
    return EFS_OK;
}

// Sweep-shape tuner.  Candidates: the planner's workgroup target times
// 1/2, 1, 2, 3 and 4; each is timed over `reps` E + H sweep pairs (after one
// untimed pair) with HIP events on the engine's stream, and the fastest is
// kept (the planner's own target unless another is at least 1 % faster).
// The sweeps run on the live arrays, so the fields and UPML flux are copied
// aside first and back after: the simulation state is unchanged.  Grids with
// side work (small grids, latency-bound) and the other paths keep the
// planner's target.
int tune_sweeps(efs_ctx* c, uint32_t reps, double* scale_out) {
    const bool has_pml = c->pml.n > 0;
    const bool eligible = c->vec4 && !c->dbl && !c->side_on && (!has_pml || (c->pml_split && c->pml_fused));
    if (!eligible || reps == 0) {
        if (scale_out) *scale_out = c->blocks_scale;
        return EFS_OK;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float *fsave = nullptr, *vsave = nullptr, *csave = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    const double scale0 = c->blocks_scale;
    bool swept = false;  // the live arrays have been stepped (a failure must put them back)
    auto release = [&]() {
        dfree(fsave); dfree(vsave); dfree(csave);
        for (hipEvent_t& e : ev)
            if (e) hipEventDestroy(e);
    };
    auto run = [&]() -> int {
        TRY(dalloc(c, fsave, c->field_floats));
        HIP_TRY(c, hipMemcpyAsync(fsave, c->pool, c->field_floats * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        if (c->flux_floats) {
            TRY(dalloc(c, vsave, c->flux_floats));
            TRY(dalloc(c, csave, c->flux_floats));
            HIP_TRY(c, hipMemcpyAsync(vsave, c->pml_vflux, c->flux_floats * sizeof(float), hipMemcpyDeviceToDevice,
                                      c->stream));
            HIP_TRY(c, hipMemcpyAsync(csave, c->pml_cflux, c->flux_floats * sizeof(float), hipMemcpyDeviceToDevice,
                                      c->stream));
        }
        for (hipEvent_t& e : ev) HIP_TRY(c, hipEventCreate(&e));
        const double cand[] = {1.0, 0.5, 2.0, 3.0, 4.0};
        double best = 1.0, best_ms = 0.0;
        for (double s : cand) {
            c->blocks_scale = s;
            TRY(plan_sweeps(c));
            swept = true;
            TRY(launch_sweep_pair(c));
            HIP_TRY(c, hipEventRecord(ev[0], c->stream));
            for (uint32_t r = 0; r < reps; ++r) TRY(launch_sweep_pair(c));
            HIP_TRY(c, hipEventRecord(ev[1], c->stream));
            HIP_TRY(c, hipEventSynchronize(ev[1]));
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
            if (s == 1.0 || ms < 0.99 * best_ms) {
                best = s;
                best_ms = ms;
            }
        }
        c->blocks_scale = best;
        TRY(plan_sweeps(c));
        c->alias_dirty = true;  // lists built on the block table are rebuilt before the next step
        HIP_TRY(c, hipMemcpyAsync(c->pool, fsave, c->field_floats * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        if (c->flux_floats) {
            HIP_TRY(c, hipMemcpyAsync(c->pml_vflux, vsave, c->flux_floats * sizeof(float), hipMemcpyDeviceToDevice,
                                      c->stream));
            HIP_TRY(c, hipMemcpyAsync(c->pml_cflux, csave, c->flux_floats * sizeof(float), hipMemcpyDeviceToDevice,
                                      c->stream));
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        swept = false;
        if (scale_out) *scale_out = best;
        return EFS_OK;
    };
    const int rc = run();
    if (rc != EFS_OK && swept) {
        // back to the planner's shape and the saved state (best effort: the
        // error that got here is the one reported)
        c->blocks_scale = scale0;
        plan_sweeps(c);
        hipStreamSynchronize(c->stream);
        if (fsave) hipMemcpy(c->pool, fsave, c->field_floats * sizeof(float), hipMemcpyDeviceToDevice);
        if (vsave) hipMemcpy(c->pml_vflux, vsave, c->flux_floats * sizeof(float), hipMemcpyDeviceToDevice);
        if (csave) hipMemcpy(c->pml_cflux, csave, c->flux_floats * sizeof(float), hipMemcpyDeviceToDevice);
    }
    release();
    return rc;
}

int efs_tune_sweeps(efs_ctx* c, uint32_t reps, double* scale_out) {
    TRY(need_init(c));
    return tune_sweeps(c, reps, scale_out);
}

int efs_sync(efs_ctx* c) {
    if (!c) return EFS_ERR_STATE;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return EFS_OK;
}

uint32_t efs_num_ts(const efs_ctx* c) { return c ? c->numTS : 0; }

int efs_set_num_ts(efs_ctx* c, uint32_t ts) {
    TRY(need_init(c));
    c->numTS = ts;
    return EFS_OK;
}

int efs_get_fields(efs_ctx* c, float* volt, float* curr) {
    TRY(need_init(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t rows = (size_t)c->nx * c->ny;
    if (c->dbl) {  // narrow on the host
        std::vector<double> tmp(c->Np);
        for (int a = 0; a < 2; ++a) {
            float* out = a ? curr : volt;
            if (!out) continue;
            const double* src = (const double*)(a ? c->curr : c->volt);
            for (int n = 0; n < 3; ++n) {
                HIP_TRY(c, hipMemcpy(tmp.data(), src + n * c->S, c->Np * sizeof(double), hipMemcpyDeviceToHost));
                for (size_t r = 0; r < rows; ++r)
                    for (uint32_t z = 0; z < c->nz; ++z) out[n * c->N + r * c->nz + z] = (float)tmp[r * c->nzp + z];
            }
        }
        return EFS_OK;
    }
    if (c->half) {  // widen and scale on the host
        float sc[6];
        HIP_TRY(c, hipMemcpy(sc, c->fscale, sizeof sc, hipMemcpyDeviceToHost));
        std::vector<f16> tmp(c->Np);
        for (int a = 0; a < 2; ++a) {
            float* out = a ? curr : volt;
            if (!out) continue;
            const f16* src = (const f16*)(a ? c->curr : c->volt);
            for (int n = 0; n < 3; ++n) {
                HIP_TRY(c, hipMemcpy(tmp.data(), src + n * c->S, c->Np * sizeof(f16), hipMemcpyDeviceToHost));
                const float s = sc[3 * a + n];
                for (size_t r = 0; r < rows; ++r)
                    for (uint32_t z = 0; z < c->nz; ++z)
                        out[n * c->N + r * c->nz + z] = (float)tmp[r * c->nzp + z] * s;
            }
        }
        return EFS_OK;
    }
    for (int a = 0; a < 2; ++a) {
        float* out = a ? curr : volt;
        if (!out) continue;
        const float* src = a ? c->curr : c->volt;
        for (int n = 0; n < 3; ++n)
            HIP_TRY(c, hipMemcpy2D(out + n * c->N, c->nz * sizeof(float), src + n * c->S, c->nzp * sizeof(float),
                                   c->nz * sizeof(float), rows, hipMemcpyDeviceToHost));
    }
    return EFS_OK;
}

int efs_get_fields_f64(efs_ctx* c, double* volt, double* curr) {
    TRY(need_init(c));
    if (!c->dbl) {  // FP32 / FP16 storage: the FP32 values, widened
        std::vector<float> v(volt ? 3 * c->N : 0), i(curr ? 3 * c->N : 0);
        TRY(efs_get_fields(c, volt ? v.data() : nullptr, curr ? i.data() : nullptr));
        std::copy(v.begin(), v.end(), volt);
        std::copy(i.begin(), i.end(), curr);
        return EFS_OK;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t rows = (size_t)c->nx * c->ny;
    for (int a = 0; a < 2; ++a) {
        double* out = a ? curr : volt;
        if (!out) continue;
        const double* src = (const double*)(a ? c->curr : c->volt);
        for (int n = 0; n < 3; ++n)
            HIP_TRY(c, hipMemcpy2D(out + n * c->N, c->nz * sizeof(double), src + n * c->S, c->nzp * sizeof(double),
                                   c->nz * sizeof(double), rows, hipMemcpyDeviceToHost));
    }
    return EFS_OK;
}

int efs_upload_fields(efs_ctx* c, const float* volt, const float* curr) {
    TRY(need_init(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t rows = (size_t)c->nx * c->ny;
    if (c->dbl) {  // widen on the host (padding cells stay zero)
        std::vector<double> tmp(c->Np, 0.0);
        for (int a = 0; a < 2; ++a) {
            const float* in = a ? curr : volt;
            if (!in) continue;
            double* dst = (double*)(a ? c->curr : c->volt);
            for (int n = 0; n < 3; ++n) {
                for (size_t r = 0; r < rows; ++r)
                    for (uint32_t z = 0; z < c->nz; ++z) tmp[r * c->nzp + z] = in[n * c->N + r * c->nz + z];
                HIP_TRY(c, hipMemcpy(dst + n * c->S, tmp.data(), c->Np * sizeof(double), hipMemcpyHostToDevice));
            }
        }
        c->fields_zero = false;
        return pml_pad_sync(c);
    }
    if (c->half) {  // per component: scale 2^e with the largest value near 2^10, then narrow
        float sc[6];
        HIP_TRY(c, hipMemcpy(sc, c->fscale, sizeof sc, hipMemcpyDeviceToHost));
        std::vector<f16> tmp(c->Np, (f16)0.0f);
        for (int a = 0; a < 2; ++a) {
            const float* in = a ? curr : volt;
            if (!in) continue;
            f16* dst = (f16*)(a ? c->curr : c->volt);
            for (int n = 0; n < 3; ++n) {
                const float* src = in + n * c->N;
                float m = 0.f;
                for (size_t j = 0; j < c->N; ++j) m = std::max(m, std::fabs(src[j]));
                if (m > 0.f && std::isfinite(m)) sc[3 * a + n] = std::ldexp(1.0f, std::ilogb(m) - kRangeTarget);
                const float inv = 1.0f / sc[3 * a + n];
                for (size_t r = 0; r < rows; ++r)
                    for (uint32_t z = 0; z < c->nz; ++z) tmp[r * c->nzp + z] = (f16)(src[r * c->nz + z] * inv);
                HIP_TRY(c, hipMemcpy(dst + n * c->S, tmp.data(), c->Np * sizeof(f16), hipMemcpyHostToDevice));
            }
        }
        HIP_TRY(c, hipMemcpy(c->fscale, sc, sizeof sc, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->f_smin, sc, sizeof sc, hipMemcpyHostToDevice));
        c->fields_zero = false;
        return pml_pad_sync(c);
    }
    c->fields_zero = false;
    for (int a = 0; a < 2; ++a) {
        const float* in = a ? curr : volt;
        if (!in) continue;
        float* dst = a ? c->curr : c->volt;
        for (int n = 0; n < 3; ++n)
            HIP_TRY(c, hipMemcpy2D(dst + n * c->S, c->nzp * sizeof(float), in + n * c->N, c->nz * sizeof(float),
                                   c->nz * sizeof(float), rows, hipMemcpyHostToDevice));
    }
    return pml_pad_sync(c);
}

int efs_cfg_gather(efs_ctx* c, const uint32_t* idx, uint32_t n) {
    TRY(need_init(c));
    if (n == 0) return EFS_OK;  // the reference ignores an empty list
    for (uint32_t k = 0; k < n; ++k)
        if (idx[k] >= 3 * c->N) return fail(c, EFS_ERR_ARG, "gather index %u out of range", k);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const std::vector<uint32_t> t = tr_all(c, idx, n);
    TRY(dupload(c, c->g_idx, t.data(), n));
    TRY(dalloc(c, c->g_out, 2 * (size_t)n));
    c->g_n = n;
    return EFS_OK;
}

int efs_read_gather(efs_ctx* c, float* out) {
    TRY(need_init(c));
    if (!c->g_n) return fail(c, EFS_ERR_STATE, "probe gather not configured");
    if (c->dbl)
        hipLaunchKernelGGL(k_gather<FieldView64>, dim3(nblk(c->g_n)), dim3(kPointThreads), 0, c->stream, fv64(c, 0),
                           fv64(c, 1), c->g_idx, c->g_out, c->g_n);
    else
        hipLaunchKernelGGL(k_gather<FieldView>, dim3(nblk(c->g_n)), dim3(kPointThreads), 0, c->stream, fv(c, 0),
                           fv(c, 1), c->g_idx, c->g_out, c->g_n);
    CHECK_LAUNCH(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, c->g_out, 2 * (size_t)c->g_n * sizeof(float), hipMemcpyDeviceToHost));
    return EFS_OK;
}

int efs_cfg_gather_buffered(efs_ctx* c, const uint32_t* idx, uint32_t n, uint32_t max_samples, uint32_t interval) {
    TRY(need_init(c));
    if (n == 0) return EFS_OK;
    if (interval == 0) return fail(c, EFS_ERR_ARG, "sampleInterval must be >= 1");
    for (uint32_t k = 0; k < n; ++k)
        if (idx[k] >= 3 * c->N) return fail(c, EFS_ERR_ARG, "gather index %u out of range", k);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const std::vector<uint32_t> t = tr_all(c, idx, n);
    TRY(dupload(c, c->gb_idx, t.data(), n));
    c->h_gb_idx.assign(idx, idx + n);
    c->fuse_dirty = true;
    TRY(dzero(c, c->gb_ring, (size_t)max_samples * 2 * n));
    c->gb_n = n;
    c->gb_max = max_samples;
    c->gb_interval = interval;
    c->gb_count = 0;
    return EFS_OK;
}

int efs_read_gather_buffered(efs_ctx* c, float* out, uint32_t* num_samples) {
    TRY(need_init(c));
    if (num_samples) *num_samples = c->gb_n ? c->gb_count : 0;
    if (!out || !c->gb_n || !c->gb_count) return EFS_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, c->gb_ring, (size_t)c->gb_count * 2 * c->gb_n * sizeof(float),
                         hipMemcpyDeviceToHost));
    return EFS_OK;
}

int efs_cfg_nf2ff(efs_ctx* c, const uint32_t* points, uint32_t P, const double* omega, uint32_t F, double dT,
                  const float* edge_lens, const uint32_t* el_len, uint32_t max_ts, uint32_t window_type,
                  uint32_t k_block) {
    TRY(need_init(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (P == 0) {
        c->nf_P = 0;
        c->nf_cap = NfCapture{};
        return EFS_OK;
    }
    if (!points || !omega || F == 0 || !edge_lens || !el_len)
        return fail(c, EFS_ERR_ARG, "NF2FF points, omega and edge lengths are required");
    const uint32_t dims[3] = {c->nx, c->ny, c->nz};
    for (int a = 0; a < 3; ++a) {
        if (el_len[a] + 1 < dims[a]) return fail(c, EFS_ERR_ARG, "primal edge lengths %d too short", a);
        if (el_len[3 + a] < dims[a]) return fail(c, EFS_ERR_ARG, "dual edge lengths %d too short", a);
    }
    for (uint32_t p = 0; p < P; ++p)
        for (int a = 0; a < 3; ++a)
            if (points[3 * p + a] >= dims[a]) return fail(c, EFS_ERR_ARG, "NF2FF point %u outside the grid", p);
    const uint32_t K = k_block ? k_block : 64;
    if (K > 256) return fail(c, EFS_ERR_ARG, "NF2FF k_block must be <= 256");
    uint64_t el_total = 0;
    Nf2ffGeom g{};
    g.nx = c->nx; g.ny = c->ny; g.nz = c->nz; g.P = P;
    g.nzs = c->nzp;
    g.S = c->S;
    const uint64_t rowsE = (3 * (uint64_t)P + kFoldRows - 1) / kFoldRows * kFoldRows;
    if (2 * rowsE >= (1ull << 31)) return fail(c, EFS_ERR_ARG, "too many NF2FF points");
    g.rowsE = (uint32_t)rowsE;
    const uint32_t rows = 2 * g.rowsE;
    const uint32_t NC = (2 * F + kFoldCols - 1) / kFoldCols * kFoldCols;
    for (int a = 0; a < 6; ++a) {
        g.el_off[a] = (uint32_t)el_total;
        el_total += el_len[a];
    }
    std::vector<uint32_t> perm(P), sorted(3 * (size_t)P);
    for (uint32_t p = 0; p < P; ++p) perm[p] = p;
    auto lin = [&](uint32_t p) {
        return ((uint64_t)points[3 * p] * c->ny + points[3 * p + 1]) * c->nz + points[3 * p + 2];
    };
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return lin(a) < lin(b); });
    for (uint32_t q = 0; q < P; ++q)
        for (int a = 0; a < 3; ++a) sorted[3 * (size_t)q + a] = points[3 * (size_t)perm[q] + a];
    TRY(dupload(c, c->nf_pts, sorted.data(), sorted.size()));
    c->h_nf_pts = sorted;
    c->fuse_dirty = true;
    TRY(dupload(c, c->nf_perm_d, perm.data(), perm.size()));
    c->nf_perm = std::move(perm);
    TRY(plan_nf_capture(c, sorted, P));
    TRY(dupload(c, c->nf_el, edge_lens, el_total));
    TRY(dupload(c, c->nf_omega, omega, F));
    TRY(dzero(c, c->nf_buf, (size_t)K * rows * rw(c)));  // padding rows stay zero
    TRY(dzero(c, c->nf_acc, (size_t)rows * NC * rw(c)));
    dfree(c->nf_cmp);  // FP64 accumulators need no compensation
    if (c->nf_comp_req && !c->dbl) TRY(dzero(c, c->nf_cmp, (size_t)rows * NC));
    TRY(dalloc(c, c->nf_tw, (size_t)2 * K * NC * rw(c)));
    c->nf_geom = g;
    c->nf_P = P; c->nf_F = F; c->nf_K = K;
    c->nf_rows = rows; c->nf_NC = NC;
    c->nf_dT = dT; c->nf_maxTS = max_ts; c->nf_window = window_type;
    c->nf_wnorm = 1.0;
    if (window_type == 1 && max_ts > 1) {
        double s = 0;
        for (uint32_t n = 0; n < max_ts; ++n) s += 0.5 * (1 - std::cos(2 * M_PI * n / (max_ts - 1)));
        c->nf_wnorm = 1.0 / s;
    }
    c->nf_slot = 0;
    return EFS_OK;
}

// The accumulators in the caller's layout, as T (float or double): device
// [row][2f + re/im] (grid-ordered points) -> caller [f][6p + 2k + re/im];
// the compensated form adds its error array (in double).
}  // extern "C"

template <typename T>
static int read_nf2ff_t(efs_ctx* c, T* accE, T* accH) {
    TRY(need_init(c));
    if (!c->nf_P) return fail(c, EFS_ERR_STATE, "NF2FF accumulation not configured");
    TRY(nf2ff_fold(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t P = c->nf_P, NC = c->nf_NC, RE = c->nf_geom.rowsE;
    std::vector<double> acc((size_t)c->nf_rows * NC);
    if (c->dbl) {
        HIP_TRY(c, hipMemcpy(acc.data(), c->nf_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> a32(acc.size());
        HIP_TRY(c, hipMemcpy(a32.data(), c->nf_acc, a32.size() * sizeof(float), hipMemcpyDeviceToHost));
        std::copy(a32.begin(), a32.end(), acc.begin());
        if (c->nf_cmp) {  // compensated form: value = acc + cmp
            HIP_TRY(c, hipMemcpy(a32.data(), c->nf_cmp, a32.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < acc.size(); ++i) acc[i] += (double)a32[i];
        }
    }
    for (size_t q = 0; q < P; ++q) {
        const size_t p = c->nf_perm[q];
        for (int k = 0; k < 3; ++k) {
            const double* e = acc.data() + (k * P + q) * NC;
            const double* h = acc.data() + (RE + k * P + q) * NC;
            for (uint32_t f = 0; f < c->nf_F; ++f) {
                const size_t o = (size_t)f * 6 * P + 6 * p + 2 * k;
                if (accE) {
                    accE[o] = (T)e[2 * f];
                    accE[o + 1] = (T)e[2 * f + 1];
                }
                if (accH) {
                    accH[o] = (T)h[2 * f];
                    accH[o + 1] = (T)h[2 * f + 1];
                }
            }
        }
    }
    return EFS_OK;
}

extern "C" {

int efs_read_nf2ff(efs_ctx* c, float* accE, float* accH) { return read_nf2ff_t(c, accE, accH); }

int efs_read_nf2ff_f64(efs_ctx* c, double* accE, double* accH) { return read_nf2ff_t(c, accE, accH); }

}  // extern "C"

// Far-field constants for one call (frequency, radius, centre, mirror).
static FarParams far_params(const efs_ctx* c, uint32_t nT, uint32_t nP, const double* center, double frequency,
                            double radius, int mdir, double mpos, int mtype, int image_normals) {
    const double C0 = 299792458.0, Z0 = 376.73031346177066;
    const double k = 2 * M_PI * frequency / C0;
    const double fmag = k / (4 * M_PI * radius), fph = -k * radius;
    FarParams q{};
    q.P = c->nf_P; q.nT = nT; q.nP = nP;
    q.k = (float)k; q.Z0 = (float)Z0;
    q.cx = (float)center[0]; q.cy = (float)center[1]; q.cz = (float)center[2];
    q.fac_re = (float)(fmag * -std::sin(fph));
    q.fac_im = (float)(fmag * std::cos(fph));
    q.mdir = mdir;
    q.mpos = (float)mpos;
    q.msign = image_normals ? -1.0f : 1.0f;
    for (int k3 = 0; k3 < 3; ++k3) q.ef[k3] = q.hf[k3] = 1.0f;
    if (mdir >= 0) {
        const int a1 = (mdir + 1) % 3, a2 = (mdir + 2) % 3;
        if (mtype == 0) {  // PEC: tangential E and normal H reversed
            q.hf[mdir] = -1.0f; q.ef[a1] = -1.0f; q.ef[a2] = -1.0f;
        } else {           // PMC: normal E and tangential H reversed
            q.ef[mdir] = -1.0f; q.hf[a1] = -1.0f; q.hf[a2] = -1.0f;
        }
    }
    return q;
}

// Point chunks of the far-field launch: whole LDS tiles, about 2048
// workgroups even for a few angles, at most max_chunks chunks.
struct FarPlan {
    uint32_t ablk, chunk, nchunks;
};
static FarPlan far_plan_n(uint32_t P, size_t na, bool mirror, uint32_t max_chunks) {
    FarPlan p{};
    p.ablk = nblk(na, 256);
    const uint32_t Pv = mirror ? 2 * P : P;  // the image adds a second pass over the points
    const uint32_t tiles = (Pv + kFarTile - 1) / kFarTile;
    const uint32_t want = std::max(1u, std::min({tiles, (2048u + p.ablk - 1) / p.ablk, max_chunks}));
    p.chunk = ((tiles + want - 1) / want) * kFarTile;
    p.nchunks = (Pv + p.chunk - 1) / p.chunk;
    return p;
}

static FarPlan far_plan(const efs_ctx* c, size_t na, bool mirror, uint32_t max_chunks) {
    return far_plan_n(c->nf_P, na, mirror, max_chunks);
}

// The far-field kernels on the context's stream: per-chunk FP64 partials
// into `part` (nchunks * angles * 8 doubles), then their ordered sum and the
// 5 output floats per angle into `out`.  meta is in internal point order,
// or in the caller's order when perm is given.
static void launch_farfield(efs_ctx* c, uint32_t f, const float* meta, const uint32_t* perm, const float* theta,
                            const float* phi, const FarParams& q, const FarPlan& p, double* part, float* out) {
    if (c->dbl)
        hipLaunchKernelGGL(k_farfield<double>, dim3(p.ablk, p.nchunks), dim3(256), 0, c->stream,
                           (const double*)c->nf_acc + 2 * (size_t)f, nullptr, c->nf_NC, c->nf_geom.rowsE, meta, perm,
                           theta, phi, q, p.chunk, part);
    else
        hipLaunchKernelGGL(k_farfield<float>, dim3(p.ablk, p.nchunks), dim3(256), 0, c->stream,
                           c->nf_acc + 2 * (size_t)f, c->nf_cmp ? c->nf_cmp + 2 * (size_t)f : nullptr, c->nf_NC,
                           c->nf_geom.rowsE, meta, perm, theta, phi, q, p.chunk, part);
    hipLaunchKernelGGL(k_farfield_final<float>, dim3(p.ablk), dim3(256), 0, c->stream, part, p.nchunks, q, out);
}

static int farfield_checks(efs_ctx* c, uint32_t f, const void* meta, const void* theta, uint32_t nT, const void* phi,
                           uint32_t nP, const double* center, double radius, const void* out) {
    TRY(need_init(c));
    if (!c->nf_P) return fail(c, EFS_ERR_STATE, "NF2FF accumulation not configured");
    if (f >= c->nf_F) return fail(c, EFS_ERR_ARG, "frequency index out of range");
    if (!meta || !theta || !phi || !center || !out || !nT || !nP)
        return fail(c, EFS_ERR_ARG, "far-field arrays are required");
    if (!(radius > 0)) return fail(c, EFS_ERR_ARG, "radius must be > 0");
    if ((size_t)nT * nP > 0xffffffffull / 256) return fail(c, EFS_ERR_ARG, "too many far-field angles");
    return EFS_OK;
}

static int farfield_impl(efs_ctx* c, uint32_t f, const float* meta, const float* theta, uint32_t nT, const float* phi,
                         uint32_t nP, const double* center, double frequency, double radius, int mdir, double mpos,
                         int mtype, int image_normals, float* out) {
    TRY(farfield_checks(c, f, meta, theta, nT, phi, nP, center, radius, out));
    if (mdir > 2 || (mdir >= 0 && mtype != 0 && mtype != 1))
        return fail(c, EFS_ERR_ARG, "mirror direction must be -1..2 and type 0 (PEC) or 1 (PMC)");
    TRY(nf2ff_fold(c));
    float *d_meta = nullptr, *d_th = nullptr, *d_ph = nullptr, *d_out = nullptr;
    double* d_part = nullptr;
    int rc = EFS_OK;
    const size_t na = (size_t)nT * nP;
    const FarPlan plan = far_plan(c, na, mdir >= 0, 0xffffffffu);
    std::vector<float> meta_q(8 * (size_t)c->nf_P);  // caller order -> internal point order
    for (size_t q = 0; q < c->nf_P; ++q)
        std::memcpy(&meta_q[8 * q], meta + 8 * (size_t)c->nf_perm[q], 8 * sizeof(float));
    do {
        if ((rc = dupload(c, d_meta, meta_q.data(), meta_q.size()))) break;
        if ((rc = dupload(c, d_th, theta, nT))) break;
        if ((rc = dupload(c, d_ph, phi, nP))) break;
        if ((rc = dalloc(c, d_out, 5 * na))) break;
        if ((rc = dalloc(c, d_part, (size_t)plan.nchunks * na * 8))) break;
        const FarParams q = far_params(c, nT, nP, center, frequency, radius, mdir, mpos, mtype, image_normals);
        launch_farfield(c, f, d_meta, nullptr, d_th, d_ph, q, plan, d_part, d_out);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(out, d_out, 5 * na * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, EFS_ERR_HIP, "far field: %s", hipGetErrorString(e));
    } while (0);
    hipStreamSynchronize(c->stream);
    dfree(d_meta); dfree(d_th); dfree(d_ph); dfree(d_out); dfree(d_part);
    return rc;
}

extern "C" {

int efs_farfield(efs_ctx* c, uint32_t f, const float* meta, const float* theta, uint32_t nT, const float* phi,
                 uint32_t nP, const double* center, double frequency, double radius, float* out) {
    return farfield_impl(c, f, meta, theta, nT, phi, nP, center, frequency, radius, -1, 0.0, 0, 0, out);
}

int efs_farfield_mirror(efs_ctx* c, uint32_t f, const float* meta, const float* theta, uint32_t nT,
                        const float* phi, uint32_t nP, const double* center, double frequency, double radius,
                        int mirror_dir, double mirror_pos, int mirror_type, int image_normals, float* out) {
    return farfield_impl(c, f, meta, theta, nT, phi, nP, center, frequency, radius, mirror_dir, mirror_pos,
                         mirror_type, image_normals, out);
}

int efs_farfield_noalloc(efs_ctx* c, uint32_t f, const float* meta_dev, const float* theta_dev, uint32_t nT,
                         const float* phi_dev, uint32_t nP, const double* center, double frequency, double radius,
                         float* out_dev) {
    TRY(farfield_checks(c, f, meta_dev, theta_dev, nT, phi_dev, nP, center, radius, out_dev));
    if (!c->scratch)
        return fail(c, EFS_ERR_HIP, "far-field scratch unavailable: %s", hipGetErrorString(c->scratch_err));
    if (!c->nf_perm_d) return fail(c, EFS_ERR_STATE, "NF2FF accumulation not configured");
    const size_t na = (size_t)nT * nP;
    const size_t cap_chunks = kFarScratchDoubles / (8 * na);  // chunks of partials that fit the scratch
    if (cap_chunks < 1)
        return fail(c, EFS_ERR_ARG, "efs_farfield_noalloc: %zu angles exceed the scratch (at most %zu angles)", na,
                    kFarScratchDoubles / 8);
    TRY(nf2ff_fold(c));
    const FarPlan plan = far_plan(c, na, false, (uint32_t)std::min<size_t>(cap_chunks, 0xffffffffu));
    const FarParams q = far_params(c, nT, nP, center, frequency, radius, -1, 0.0, 0, 0);
    // every partial the final kernel reads is written by this call's first
    // kernel, so nothing needs clearing between calls
    launch_farfield(c, f, meta_dev, c->nf_perm_d, theta_dev, phi_dev, q, plan, c->scratch, out_dev);
    CHECK_LAUNCH(c);
    return EFS_OK;
}

int efs_farfield_surface(efs_ctx* c, uint32_t P, const float* meta, const double* acc_e, const double* acc_h,
                         const float* theta, uint32_t nT, const float* phi, uint32_t nP, const double* center,
                         double frequency, double radius, double* out) {
    if (!c) return EFS_ERR_STATE;
    if (!P || !meta || !acc_e || !acc_h || !theta || !phi || !center || !out || !nT || !nP)
        return fail(c, EFS_ERR_ARG, "far-field arrays are required");
    if (!(radius > 0)) return fail(c, EFS_ERR_ARG, "radius must be > 0");
    if (!(frequency > 0)) return fail(c, EFS_ERR_ARG, "frequency must be > 0");
    if ((size_t)nT * nP > 0xffffffffull / 256) return fail(c, EFS_ERR_ARG, "too many far-field angles");
    if ((uint64_t)P * 6 >= (1ull << 31)) return fail(c, EFS_ERR_ARG, "too many surface points");
    for (uint32_t p = 0; p < P; ++p) {
        const float nd = meta[8 * (size_t)p + 3];
        if (!(nd == 0.f || nd == 1.f || nd == 2.f))
            return fail(c, EFS_ERR_ARG, "point %u: normal direction must be 0, 1 or 2", p);
    }
    HIP_TRY(c, hipSetDevice(c->device));
    // the kernel's row layout: E rows c*P + p, then H rows 3P + c*P + p,
    // (re, im) per row
    std::vector<double> acc(12 * (size_t)P);
    for (uint32_t p = 0; p < P; ++p)
        for (uint32_t k = 0; k < 3; ++k)
            for (uint32_t r = 0; r < 2; ++r) {
                acc[2 * ((size_t)k * P + p) + r] = acc_e[6 * (size_t)p + 2 * k + r];
                acc[2 * (3 * (size_t)P + (size_t)k * P + p) + r] = acc_h[6 * (size_t)p + 2 * k + r];
            }
    const size_t na = (size_t)nT * nP;
    const FarPlan plan = far_plan_n(P, na, false, 0xffffffffu);
    const double C0 = 299792458.0, Z0 = 376.73031346177066;
    const double k = 2 * M_PI * frequency / C0;
    const double fmag = k / (4 * M_PI * radius), fph = -k * radius;
    FarParams q{};
    q.P = P; q.nT = nT; q.nP = nP;
    q.k = (float)k; q.Z0 = (float)Z0;
    q.cx = (float)center[0]; q.cy = (float)center[1]; q.cz = (float)center[2];
    q.fac_re = (float)(fmag * -std::sin(fph));
    q.fac_im = (float)(fmag * std::cos(fph));
    q.mdir = -1;
    q.msign = 1.0f;
    for (int k3 = 0; k3 < 3; ++k3) q.ef[k3] = q.hf[k3] = 1.0f;
    double *d_acc = nullptr, *d_part = nullptr, *d_out = nullptr;
    float *d_meta = nullptr, *d_th = nullptr, *d_ph = nullptr;
    int rc = EFS_OK;
    do {
        if ((rc = dupload(c, d_acc, acc.data(), acc.size()))) break;
        if ((rc = dupload(c, d_meta, meta, 8 * (size_t)P))) break;
        if ((rc = dupload(c, d_th, theta, nT))) break;
        if ((rc = dupload(c, d_ph, phi, nP))) break;
        if ((rc = dalloc(c, d_out, 5 * na))) break;
        if ((rc = dalloc(c, d_part, (size_t)plan.nchunks * na * 8))) break;
        hipLaunchKernelGGL(k_farfield<double>, dim3(plan.ablk, plan.nchunks), dim3(256), 0, c->stream,
                           (const double*)d_acc, (const double*)nullptr, 2u, 3 * P, (const float*)d_meta,
                           (const uint32_t*)nullptr, (const float*)d_th, (const float*)d_ph, q, plan.chunk, d_part);
        hipLaunchKernelGGL(k_farfield_final<double>, dim3(plan.ablk), dim3(256), 0, c->stream,
                           (const double*)d_part, plan.nchunks, q, d_out);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(out, d_out, 5 * na * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, EFS_ERR_HIP, "surface far field: %s", hipGetErrorString(e));
    } while (0);
    hipStreamSynchronize(c->stream);
    dfree(d_acc); dfree(d_meta); dfree(d_th); dfree(d_ph); dfree(d_out); dfree(d_part);
    return rc;
}

int efs_set_nf2ff_compensated(efs_ctx* c, int on) {
    if (!c) return EFS_ERR_STATE;
    if (on != 0 && on != 1) return fail(c, EFS_ERR_ARG, "compensated must be 0 or 1");
    c->nf_comp_req = on != 0;
    return EFS_OK;
}

int efs_energy(efs_ctx* c, double* out) {
    TRY(need_init(c));
    if (!out) return fail(c, EFS_ERR_ARG, "out is required");
    if (c->nx < 2 || c->ny < 2 || c->nz < 2) {
        *out = 0.0;
        return EFS_OK;
    }
    const uint32_t rows = (c->nx - 1) * (c->ny - 1);
    const uint32_t blocks = std::max(1u, std::min((uint32_t)kEnergyMaxBlocks, (rows + 3) / 4));
    if (c->dbl)
        hipLaunchKernelGGL(k_energy_partial<double>, dim3(blocks), dim3(kEnergyThreads), 0, c->stream,
                           (const double*)c->volt, (const double*)c->curr, c->nx, c->ny, c->nz, c->nzp,
                           (uint64_t)c->S, c->fscale, c->en_partial);
    else if (c->half)
        hipLaunchKernelGGL(k_energy_partial<f16>, dim3(blocks), dim3(kEnergyThreads), 0, c->stream,
                           (const f16*)c->volt, (const f16*)c->curr, c->nx, c->ny, c->nz, c->nzp, (uint64_t)c->S,
                           c->fscale, c->en_partial);
    else
        hipLaunchKernelGGL(k_energy_partial<float>, dim3(blocks), dim3(kEnergyThreads), 0, c->stream, c->volt,
                           c->curr, c->nx, c->ny, c->nz, c->nzp, (uint64_t)c->S, c->fscale, c->en_partial);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kEnergyThreads), 0, c->stream, c->en_partial, blocks,
                       c->en_partial + kEnergyMaxBlocks);
    CHECK_LAUNCH(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, c->en_partial + kEnergyMaxBlocks, sizeof(double), hipMemcpyDeviceToHost));
    return EFS_OK;
}

int efs_get_sizes(const efs_ctx* c, efs_sizes* out) {
    if (!c || !out) return EFS_ERR_ARG;
    *out = efs_sizes{};
    out->nx = c->nx; out->ny = c->ny; out->nz = c->nz;
    out->gather_n = c->g_n;
    out->gather_buffered_n = c->gb_n;
    out->gather_buffered_max = c->gb_max;
    out->nf2ff_points = c->nf_P;
    out->nf2ff_freqs = c->nf_F;
    out->steady_state_n = c->ss_n;
    return EFS_OK;
}

int efs_path_flags(const efs_ctx* c, uint32_t* flags) {
    if (!c || !flags) return EFS_ERR_ARG;
    const bool upml4 = c->vec4 && c->pml.n && c->pml_split;
    *flags = (c->vec4 ? EFS_PATH_VEC4 : 0u) | (c->comp ? EFS_PATH_COEF_IDX : 0u) | (upml4 ? EFS_PATH_UPML_4 : 0u) |
             (upml4 && c->pml_fused ? EFS_PATH_FUSED : 0u) |
             (upml4 && c->pml_fused && c->pml_comp ? EFS_PATH_UPML_IDX : 0u) | (c->half ? EFS_PATH_F16 : 0u) |
             (nf_capture(c).on ? EFS_PATH_NF2FF_CAPTURE : 0u) | (c->comp16 ? EFS_PATH_COEF_IDX16 : 0u) |
             (c->dbl ? EFS_PATH_F64 : 0u) |
             (c->side_on && c->fuse_hooks && c->vec4 && !c->dbl && (!c->pml.n || c->pml_split) ? EFS_PATH_SIDE_WORK
                                                                                                : 0u) |
             (c->ep_ok ? EFS_PATH_SWEEP_EPILOGUE : 0u) |
             (c->post_v_fuse ? EFS_PATH_FUSE_TFSF_ADE : 0u) | (c->post_i_fuse ? EFS_PATH_FUSE_TFSF_HOOKS : 0u);
    return EFS_OK;
}

int efs_set_storage(efs_ctx* c, int storage) {
    if (!c) return EFS_ERR_STATE;
    if (storage != EFS_STORAGE_F32 && storage != EFS_STORAGE_F16 && storage != EFS_STORAGE_F64)
        return fail(c, EFS_ERR_ARG, "unknown storage %d", storage);
    c->storage_req = storage;
    return EFS_OK;
}

int efs_field_scales(efs_ctx* c, float* scales) {
    TRY(need_init(c));
    if (!scales) return fail(c, EFS_ERR_ARG, "scales is required");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(scales, c->fscale, 6 * sizeof(float), hipMemcpyDeviceToHost));
    return EFS_OK;
}

int efs_profile_enable(efs_ctx* c, int on) {
    if (!c) return EFS_ERR_STATE;
    TRY(drain_profile(c));
    c->prof = on != 0;
    for (int k = 0; k < EFS_PROF_COUNT; ++k) {
        c->prof_ms[k] = 0;
        c->prof_n[k] = 0;
    }
    return EFS_OK;
}

int efs_profile_read(efs_ctx* c, int which, double* total_ms, uint64_t* count) {
    if (!c) return EFS_ERR_STATE;
    if (which < 0 || which >= EFS_PROF_COUNT) return fail(c, EFS_ERR_ARG, "unknown profile slot");
    TRY(drain_profile(c));
    if (total_ms) *total_ms = c->prof_ms[which];
    if (count) *count = c->prof_n[which];
    return EFS_OK;
}

}  // extern "C"
