/*
 * efs_engine.h -- C ABI of the MI355X FDTD time-stepping engine.
 *
 * Drop-in boundary for the engine interface of the reference's WebGPU engine
 * (class WebGPUEngine, src/webgpu-engine.mjs:994-3498).  Each entry point
 * names the method it replaces.  Conventions:
 *   - every array is a plain host pointer plus an explicit count; the engine
 *     copies what it needs at configure time and never keeps the caller's
 *     pointer (src/webgpu-engine.mjs:3231-3247);
 *   - every call returns 0 (EFS_OK) or an EFS_ERR_* code; the message of the
 *     last failure is returned by efs_last_error();
 *   - field and coefficient arrays use the reference's NIJK layout:
 *     addr(n,x,y,z) = n*Nx*Ny*Nz + x*Ny*Nz + y*Nz + z, 3 components;
 *   - "linear position" (pos, field_idx of RLC/ADE) is x*Ny*Nz + y*Nz + z
 *     without the component offset, as in the reference;
 *   - work is enqueued on the context's HIP stream; only the read/energy
 *     calls and efs_sync() wait for it.
 */
#ifndef EFS_ENGINE_H
#define EFS_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct efs_ctx efs_ctx;

enum {
    EFS_OK = 0,
    EFS_ERR_NO_DEVICE = 1, /* no GPU: initGPU() resolves false */
    EFS_ERR_STATE = 2,     /* call out of order, e.g. before efs_init */
    EFS_ERR_ARG = 3,       /* invalid size, index or region */
    EFS_ERR_HIP = 4,       /* HIP runtime failure */
    EFS_ERR_NOMEM = 5
};

/* Library version string, e.g. "efs-mi355x 0.1". */
const char *efs_version(void);
/* Number of visible HIP devices (0 when there is none). */
int efs_device_count(void);

/* initGPU() (src/webgpu-engine.mjs:1107-1135): create a context on HIP
 * device `device`.  Returns EFS_ERR_NO_DEVICE when none is present. */
int efs_create(efs_ctx **out, int device);
/* destroy() (src/webgpu-engine.mjs:3163-3210). */
void efs_destroy(efs_ctx *ctx);
/* Message of the last failing call on ctx ("" if none).  With ctx == NULL,
 * the message of the last failing efs_create. */
const char *efs_last_error(const efs_ctx *ctx);

/* init(gridSize, {vv,vi,ii,iv}) (src/webgpu-engine.mjs:1143-1185): each
 * coefficient array holds 3*nx*ny*nz floats.  Fields start at zero and
 * numTS at 0; every extension is cleared. */
int efs_init(efs_ctx *ctx, uint32_t nx, uint32_t ny, uint32_t nz, const float *vv, const float *vi,
             const float *ii, const float *iv);

/* configureExcitation (src/webgpu-engine.mjs:1199-1227).  signal[len];
 * per point n: delay, amp, dir (0..2), pos (linear position).  period 0 =
 * not periodic.  A signal position past the end reads signal[0] (openEMS
 * Apply2Voltages; src/webgpu-fdtd.mjs:169-183). */
int efs_cfg_excitation(efs_ctx *ctx, const float *signal, uint32_t len, const uint32_t *delay,
                       const float *amp, const uint32_t *dir, const uint32_t *pos, uint32_t n,
                       uint32_t period);

/* configurePML(regions) (src/webgpu-engine.mjs:1235-1323).  Replaces all
 * regions.  start[3*r], num[3*r] per region; the six coefficient arrays are
 * the regions' 3*nx*ny*nz-float NIJK arrays concatenated in region order.
 * Regions must lie inside the grid and must not overlap.  nregions 0 clears
 * the UPML. */
int efs_cfg_pml(efs_ctx *ctx, uint32_t nregions, const uint32_t *start, const uint32_t *num,
                const float *vv, const float *vvfo, const float *vvfn, const float *ii,
                const float *iifo, const float *iifn);

/* configureMur (src/webgpu-engine.mjs:1822-1918), single-list form: n points
 * with coeff, normal_idx and shifted_idx (full field indices incl. the
 * component offset).  The dual nyP/nyPP form is the concatenation of both
 * lists (done by the host wrappers).  n 0 clears. */
int efs_cfg_mur(efs_ctx *ctx, uint32_t n, const float *coeff, const uint32_t *normal_idx,
                const uint32_t *shifted_idx);

/* configureLorentzADE (src/webgpu-engine.mjs:1358-1475) is expressed as a
 * clear followed by one efs_cfg_ade_add per (order, direction).  NULL
 * coefficient pointers mean zeros, as in the reference. */
int efs_cfg_ade_clear(efs_ctx *ctx);
int efs_cfg_ade_add(efs_ctx *ctx, uint32_t has_lorentz, uint32_t dir, uint32_t ncells,
                    const uint32_t *pos_idx, const float *v_int, const float *v_ext, const float *v_lor,
                    const float *i_int, const float *i_ext, const float *i_lor);
/* 0 (default): reference semantics, the ADE state is never written back to
 * the fields (src/webgpu-fdtd.mjs:405-454).  1: openEMS write-back
 * (field -= ADE state after the E and after the H half step). */
int efs_set_ade_apply(efs_ctx *ctx, int apply);

/* configureTFSF (src/webgpu-engine.mjs:1527-1645): voltage points (nv) and
 * current points (ni), each {field_idx (full index), delay_int, delay_frac,
 * amp}. */
int efs_cfg_tfsf(efs_ctx *ctx, const float *signal, uint32_t len, uint32_t period, uint32_t nv,
                 const uint32_t *v_field_idx, const uint32_t *v_delay_int, const float *v_delay_frac,
                 const float *v_amp, uint32_t ni, const uint32_t *i_field_idx,
                 const uint32_t *i_delay_int, const float *i_delay_frac, const float *i_amp);

/* configureRLC (src/webgpu-engine.mjs:1722-1794): n elements; coeffs holds 9
 * floats per element in the order i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2.
 * type_flag 0 = parallel, 1 = series.  Pre-voltage and apply phases follow
 * the CPU engine's order (src/webgpu-fdtd.mjs:1135,1155). */
int efs_cfg_rlc(efs_ctx *ctx, uint32_t n, const uint32_t *field_idx, const uint32_t *direction,
                const uint32_t *type_flag, const float *coeffs);

/* configurePBC({axes: [{axis, phase}]}) (src/webgpu-engine.mjs:1976-2155,
 * copy kernel 475-584): n_axes (<= 3) periodic axes with Bloch phases in
 * radians.  After the E update (following TFSF) the tangential E of plane
 * N-2 is copied to plane 0; after the H update the tangential H of plane 0
 * to plane N-2 over (N0-1) x (N1-1) face cells.  A phase whose f32 sine is
 * nonzero rotates (field, imaginary part), the imaginary part kept per face
 * cell.  Axes apply in the given order; n_axes 0 clears. */
int efs_cfg_pbc(efs_ctx *ctx, uint32_t n_axes, const uint32_t *axis, const double *phase);

/* configureSteadyState({probe_idx, periodSamples, threshold})
 * (src/webgpu-engine.mjs:2161-2217; CPU form src/webgpu-fdtd.mjs:1037-1095):
 * probe_idx are full field indices.  While recording, every step first adds
 * volt^2 of each probe to period 1 (current sample < period_samples) or
 * period 2, and advances the current sample.  Recording starts off; the
 * convergence threshold is applied by the host (checkConvergence). */
int efs_cfg_steady_state(efs_ctx *ctx, const uint32_t *probe_idx, uint32_t n, uint32_t period_samples);
/* the ssRecording / ssCurrentSample fields */
int efs_ss_set(efs_ctx *ctx, int recording, uint32_t current_sample);
/* the period energies (n floats each; either may be NULL) and the fields */
int efs_ss_read(efs_ctx *ctx, float *period1, float *period2, uint32_t *current_sample, int *recording);

/* iterate(numSteps) (src/webgpu-engine.mjs:2407-2504): enqueue n complete
 * timesteps; does not wait. */
int efs_iterate(efs_ctx *ctx, uint32_t n);
/* device.queue.onSubmittedWorkDone(): wait for all enqueued work. */
int efs_sync(efs_ctx *ctx);
/* numTS field. */
uint32_t efs_num_ts(const efs_ctx *ctx);
int efs_set_num_ts(efs_ctx *ctx, uint32_t ts);

/* getFields() / uploadFields(v, c) (src/webgpu-engine.mjs:2533-2571):
 * 3*nx*ny*nz floats each; either pointer may be NULL. */
int efs_get_fields(efs_ctx *ctx, float *volt, float *curr);
int efs_upload_fields(efs_ctx *ctx, const float *volt, const float *curr);
/* The fields as FP64 (no counterpart in the reference): the stored values of
 * the FP64 storage form, the FP32 values widened otherwise.  Same layout and
 * NULL rules as efs_get_fields. */
int efs_get_fields_f64(efs_ctx *ctx, double *volt, double *curr);

/* configureProbeGather / readProbeGather (src/webgpu-engine.mjs:2577-2613,
 * 2697-2715): out receives 2*n floats [volt0, curr0, volt1, ...]. */
int efs_cfg_gather(efs_ctx *ctx, const uint32_t *indices, uint32_t n);
int efs_read_gather(efs_ctx *ctx, float *out);

/* configureProbeGatherBuffered / readProbeGatherBuffered
 * (src/webgpu-engine.mjs:2615-2694): one sample every `interval` steps,
 * taken after step s when (s+1) % interval == 0, at most max_samples.
 * out receives num_samples*2*n floats (pass NULL to query the count). */
int efs_cfg_gather_buffered(efs_ctx *ctx, const uint32_t *indices, uint32_t n, uint32_t max_samples,
                            uint32_t interval);
int efs_read_gather_buffered(efs_ctx *ctx, float *out, uint32_t *num_samples);

/* configureNF2FFAccumulation (src/webgpu-engine.mjs:2729-2827), generalised
 * from one omega to F frequencies.  points: 3*P grid indices (ix,iy,iz);
 * omega: F angular frequencies (FP64); edge_lens: the six edge-length
 * arrays (primal x, y, z, dual x, y, z) concatenated, el_len[6] their
 * lengths (each at least the grid size along its axis); window_type 1 =
 * Hann over max_ts steps.  Surface E/H are buffered for k_block steps
 * (0 = default 64) and folded into the F bins as one complex GEMM with
 * twiddles generated in FP64. */
int efs_cfg_nf2ff(efs_ctx *ctx, const uint32_t *points, uint32_t P, const double *omega, uint32_t F,
                  double dT, const float *edge_lens, const uint32_t *el_len, uint32_t max_ts,
                  uint32_t window_type, uint32_t k_block);
/* readNF2FFAccumulation (src/webgpu-engine.mjs:2844-2855): accE/accH each
 * receive F*6*P floats, frequency-major, each frequency in the reference
 * layout [6*p + 2*c + {0: re, 1: im}]. */
int efs_read_nf2ff(efs_ctx *ctx, float *accE, float *accH);
/* The same accumulators as FP64 (no counterpart in the reference): the FP64
 * form's values, or the FP32 accumulators plus their compensation terms
 * summed in double. */
int efs_read_nf2ff_f64(efs_ctx *ctx, double *accE, double *accH);

/* computeNF2FFfarField (src/webgpu-engine.mjs:2871-2967) for frequency bin
 * freq_idx: point_meta 8 floats per point (pos x,y,z [m], normalDir,
 * normSign, area, pad, pad); out receives 5 floats per (theta, phi) pair,
 * theta-major: P_rad, Et_re, Et_im, Ep_re, Ep_im. */
int efs_farfield(efs_ctx *ctx, uint32_t freq_idx, const float *point_meta, const float *theta,
                 uint32_t n_theta, const float *phi, uint32_t n_phi, const double *center,
                 double frequency, double radius, float *out);

/* The mirror option of the reference's CPU far field (src/nf2ff.mjs:377-398):
 * the surface is added again reflected in the plane x[mirror_dir] =
 * mirror_pos with the image fields of a PEC (mirror_type 0: tangential E and
 * normal H reversed) or PMC (1: normal E and tangential H reversed) plane.
 * image_normals = 0 (the reference's rule) keeps the original face normals
 * in the reflected copy, as src/nf2ff.mjs:377-398 does; image_normals = 1
 * reverses the normal sign of the reflected copy of each face perpendicular
 * to the plane, which faces the other way (image theory).  mirror_dir = -1
 * is efs_farfield. */
int efs_farfield_mirror(efs_ctx *ctx, uint32_t freq_idx, const float *point_meta, const float *theta,
                        uint32_t n_theta, const float *phi, uint32_t n_phi, const double *center, double frequency,
                        double radius, int mirror_dir, double mirror_pos, int mirror_type, int image_normals,
                        float *out);

/* The far field of efs_farfield with nothing allocated, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  point_meta_dev (8 floats per NF2FF point, the caller's point
 * order), theta_dev, phi_dev and out_dev (5 floats per angle, theta-major)
 * are DEVICE pointers; center is a host array of 3.  The FP64 partial sums
 * go to a static device scratch of 4 MiB per GPU (allocated with the code
 * object, held until the process ends), so n_theta * n_phi may be at most
 * 65,536 (larger requests fail with EFS_ERR_ARG).  The work is enqueued on
 * the context's stream and the call returns without waiting (efs_sync
 * waits).  Every context on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (other streams, other threads) race on it.  No mirror
 * option. */
int efs_farfield_noalloc(efs_ctx *ctx, uint32_t freq_idx, const float *point_meta_dev, const float *theta_dev,
                         uint32_t n_theta, const float *phi_dev, uint32_t n_phi, const double *center,
                         double frequency, double radius, float *out_dev);

/* computeNF2FF (src/nf2ff.mjs:102-436, the reference's CPU far field over
 * surface fields it reads back from openEMS dumps) on the GPU: the
 * radiation integrals of P caller-supplied surface points.  point_meta: 8
 * floats per point as for efs_farfield (Cartesian position [m], normal axis
 * 0..2, normal sign, area, 2 unused); acc_e / acc_h: the complex E and H at
 * each point, 6 doubles per point [6p + 2c + {0: re, 1: im}] (Cartesian
 * components).  A point of a cylindrical mesh (src/nf2ff.mjs:165-285) is
 * given as the Cartesian points its rotated normal decomposes into (the
 * integrals are linear in area * normal); the host wrappers do that.  out
 * receives 5 doubles per (theta, phi) pair, theta-major: P_rad, Et_re,
 * Et_im, Ep_re, Ep_im.  Needs a context (efs_create) but not efs_init;
 * synchronous. */
int efs_farfield_surface(efs_ctx *ctx, uint32_t P, const float *point_meta, const double *acc_e,
                         const double *acc_h, const float *theta, uint32_t n_theta, const float *phi,
                         uint32_t n_phi, const double *center, double frequency, double radius, double *out);

/* NF2FF accumulation form (no counterpart in the reference, which adds each
 * step's term into FP32 accumulators with a plain +=,
 * src/webgpu-engine.mjs:785-798): 1 (default) = compensated, every DFT
 * fold adds its block into the accumulators with an error-free TwoSum and
 * keeps the rounding errors in a second array, so the read value (acc +
 * error) does not drift with the number of folds; 0 = plain FP32 +=.
 * Takes effect at the next efs_cfg_nf2ff. */
int efs_set_nf2ff_compensated(efs_ctx *ctx, int on);

/* computeEnergy() (src/webgpu-engine.mjs:2974-3158): sum of volt^2+curr^2
 * over the (nx-1)(ny-1)(nz-1) interior, all components; FP32 squares,
 * FP64 reduction.  Synchronous (returns the current value, not the
 * previous call's); the lagging form is built on top by the host wrappers. */
int efs_energy(efs_ctx *ctx, double *out);

/* Sizes of what the context holds, for bindings that check the caller's
 * buffers before a read or upload (no counterpart in the reference, whose
 * readbacks are sized by WebGPU): grid, plain and buffered gather lists,
 * NF2FF points and frequencies, steady-state probes.  Zero where nothing is
 * configured. */
typedef struct {
    uint32_t nx, ny, nz;
    uint32_t gather_n;                   /* efs_cfg_gather indices */
    uint32_t gather_buffered_n;          /* efs_cfg_gather_buffered indices */
    uint32_t gather_buffered_max;        /* its max_samples */
    uint32_t nf2ff_points, nf2ff_freqs;  /* P, F */
    uint32_t steady_state_n;             /* steady-state probes */
} efs_sizes;
int efs_get_sizes(const efs_ctx *ctx, efs_sizes *out);

/* Which sweep forms the context uses (diagnostic; no counterpart in the
 * reference): bit set = form in use. */
enum {
    EFS_PATH_VEC4 = 1,      /* 4 cells per lane (else 1) */
    EFS_PATH_COEF_IDX = 2,  /* core coefficients as byte indices + tables */
    EFS_PATH_UPML_4 = 4,    /* UPML handled by the 4-wide sweeps */
    EFS_PATH_FUSED = 8,     /* one launch per sweep for core and UPML cells */
    EFS_PATH_UPML_IDX = 16, /* UPML coefficients as byte indices + tables */
    EFS_PATH_F16 = 32,      /* FP16 field storage */
    EFS_PATH_NF2FF_CAPTURE = 64, /* NF2FF z-face points read slabs the H sweep wrote */
    EFS_PATH_COEF_IDX16 = 128,   /* core coefficients as 16-bit indices + cached tables */
    EFS_PATH_F64 = 256,          /* FP64 field storage and arithmetic (1-wide sweeps) */
    EFS_PATH_SIDE_WORK = 512,    /* NF2FF sampling, probe gather and Mur pre pass run in extra
                                    workgroups of the sweep kernels (EFS_SIDE_WORK) */
    EFS_PATH_SWEEP_EPILOGUE = 1024, /* Mur post/apply and the excitation run at the end of the E-sweep
                                       blocks (EFS_SWEEP_EPILOGUE; set once an iterate call has
                                       planned it) */
    EFS_PATH_FUSE_TFSF_ADE = 2048,  /* TFSF of the E side and the ADE write-back share a launch
                                       (disjoint cells; EFS_FUSE_POINTS; set once an iterate call
                                       has planned it) */
    EFS_PATH_FUSE_TFSF_HOOKS = 4096 /* TFSF of the H side may run in the launch of the NF2FF
                                       sampling and gather (cells they do not read), beside the next
                                       step's E-side ADE update */
};
int efs_path_flags(const efs_ctx *ctx, uint32_t *flags);

/* Sweep-shape tuner (no counterpart in the reference, whose compute passes
 * have fixed workgroup sizes and dispatch shapes, src/webgpu-engine.mjs:
 * 7-122, 2407-2504).  Times the E + H sweeps of the configured grid under
 * the block planner's workgroup target times 1/2, 1, 2, 3 and 4 (reps
 * timed sweep pairs each) and keeps the fastest for the following steps;
 * *scale (may be null) receives the chosen factor.  The fields and UPML
 * flux are copied aside for the call (device memory of their size, freed
 * before it returns) and restored afterwards, so results do not change.  Grids small
 * enough for side work, FP64 storage and the split UPML form keep the
 * planner's target (factor 1).  Call after init and configurePML. */
int efs_tune_sweeps(efs_ctx *ctx, uint32_t reps, double *scale);

/* Field storage (no counterpart in the reference, whose fields are FP32):
 * EFS_STORAGE_F32 (default); EFS_STORAGE_F16 -- volt and curr held as FP16
 * with a power-of-two scale per component, updated in FP32 registers, the
 * scales following the field range on the device (checked every 32 steps);
 * FP16 needs UPML z-bounds that are multiples of 4.  EFS_STORAGE_F64 -- the
 * precision reference run of the same engine: fields, UPML flux, Mur state
 * and the NF2FF buffer, fold and accumulators in FP64 (the coefficients are
 * the FP32 values given, widened); it runs the core sweeps, UPML, Mur,
 * excitation, gathers, NF2FF, far field and energy, and refuses ADE, TFSF,
 * RLC, PBC and steady-state configuration.  Takes effect at the next
 * efs_init.  Every read returns FP32 values. */
enum { EFS_STORAGE_F32 = 0, EFS_STORAGE_F16 = 1, EFS_STORAGE_F64 = 2 };
int efs_set_storage(efs_ctx *ctx, int storage);
/* Current scales: volt components 0..2, curr 0..2 (value = stored * scale;
 * all 1 with FP32 storage). */
int efs_field_scales(efs_ctx *ctx, float *scales);

/* Kernel timing with HIP events on the engine's stream. */
enum {
    EFS_PROF_SWEEP_E = 0, /* fused E sweep (core + UPML) */
    EFS_PROF_SWEEP_H = 1, /* fused H sweep (core + UPML) */
    EFS_PROF_STEP = 2,    /* one complete timestep, all kernels */
    EFS_PROF_COUNT = 3
};
int efs_profile_enable(efs_ctx *ctx, int on);
/* Total milliseconds and number of timed launches since enabling. */
int efs_profile_read(efs_ctx *ctx, int which, double *total_ms, uint64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* EFS_ENGINE_H */
