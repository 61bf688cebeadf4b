/*
 * efs_batch.h -- C ABI of the batched form: many independent FDTD grids
 * (instances, each with its own size, graded-mesh coefficients, Mur lists,
 * excitation and probes) advanced together on one GPU.
 *
 * The reference has no batched engine; callers that sweep a design
 * parameter run one WebGPUEngine per design (scripts/lib/sim-engine.mjs:
 * 537-813 drives one at a time).  Per instance, every step does exactly
 * what one engine context does for the same arrays (efs_engine.h:
 * efs_init + efs_cfg_mur + efs_cfg_excitation + efs_cfg_gather_buffered,
 * then efs_iterate): the same update arithmetic, the same phase order
 * (Mur pre, E sweep, Mur post, excitation, Mur apply, H sweep, gather).
 * What changes is the launch shape: one launch per half-step covers every
 * instance through a packed table of (instance, tile, x-range) blocks, and
 * a block of steps is replayed from one captured HIP graph.
 *
 * Conventions follow efs_engine.h (status codes, EFS_ERR_*, host copies of
 * every array, NIJK indices relative to each instance's own grid).
 * Life cycle: efs_batch_create -> efs_batch_add (per instance) ->
 * efs_batch_cfg_* (per instance) -> efs_batch_finalize -> [efs_batch_tune]
 * -> efs_batch_iterate / reads -> efs_batch_destroy.
 */
#ifndef EFS_BATCH_H
#define EFS_BATCH_H

#include <stdint.h>

#include "efs_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct efs_batch efs_batch;

/* A batch on HIP device `device` (EFS_ERR_NO_DEVICE without one). */
int efs_batch_create(efs_batch **out, int device);
void efs_batch_destroy(efs_batch *b);
/* Message of the last failing call on b; with b == NULL, of the last
 * failing efs_batch_create. */
const char *efs_batch_last_error(const efs_batch *b);

/* Field storage of every instance, before efs_batch_finalize:
 * EFS_STORAGE_F32 (default) or EFS_STORAGE_F16 (per-instance,
 * per-component power-of-two scales, FP32 updates; as efs_set_storage). */
int efs_batch_set_storage(efs_batch *b, int storage);

/* Add an instance: init(gridSize, {vv,vi,ii,iv}) of one engine, arrays of
 * 3*nx*ny*nz floats (NIJK).  *id receives the instance number (0, 1, ...). */
int efs_batch_add(efs_batch *b, uint32_t nx, uint32_t ny, uint32_t nz, const float *vv, const float *vi,
                  const float *ii, const float *iv, uint32_t *id);

/* configureExcitation of instance id (efs_cfg_excitation semantics). */
int efs_batch_cfg_excitation(efs_batch *b, uint32_t id, const float *signal, uint32_t len, const uint32_t *delay,
                             const float *amp, const uint32_t *dir, const uint32_t *pos, uint32_t n,
                             uint32_t period);

/* configureMur of instance id, single-list form (efs_cfg_mur semantics). */
int efs_batch_cfg_mur(efs_batch *b, uint32_t id, uint32_t n, const float *coeff, const uint32_t *normal_idx,
                      const uint32_t *shifted_idx);

/* Probe indices of instance id (full field indices of its grid), recorded
 * by the buffered gather (efs_cfg_gather_buffered semantics; the sample
 * interval and capacity are the batch's, set at finalize). */
int efs_batch_cfg_gather(efs_batch *b, uint32_t id, const uint32_t *idx, uint32_t n);

/* Pack every instance into device memory, build the block table and the
 * step graph.  sample_interval / max_samples: the buffered gather of all
 * instances (max_samples 0 = no gather).  No instance can be added after. */
int efs_batch_finalize(efs_batch *b, uint32_t sample_interval, uint32_t max_samples);

/* Autotuner: time `steps` whole steps (rounded up to a multiple of 32, as
 * they run: graph replays or persistent launches) for each candidate shape
 * -- launch form: column tile width 128/256/512 lanes x occupancy 4/5
 * waves per SIMD x 1024..8192 blocks with 4-cell lanes, and 256/512 lanes x
 * 1024..8192 blocks with 8-cell lanes at 3 waves; persistent
 * form: 1, 2, 4 or 8 blocks per resident workgroup, unless the environment
 * sets EFS_BATCH_PERSIST=0 -- and keep the fastest (efs_batch_get_shape
 * reports it).  Changes fields (call before the real run and reset with
 * efs_batch_reset).  *blocks / *waves receive the chosen shape (may be
 * NULL). */
int efs_batch_tune(efs_batch *b, uint32_t steps, uint32_t *blocks, uint32_t *waves);

/* Launch forms of a step (no counterpart in the reference, whose engine
 * submits one compute pass per kernel):
 *   EFS_BATCH_LAUNCH      one launch per half-step and per point pass over
 *                         every instance, 32-step blocks replayed from a
 *                         captured HIP graph;
 *   EFS_BATCH_PERSISTENT  one cooperative launch of resident workgroups
 *                         runs up to 64 steps, the phases of a step
 *                         separated by grid barriers.
 * Both give the same bits. */
#define EFS_BATCH_LAUNCH 0
#define EFS_BATCH_PERSISTENT 1

/* Set the launch shape after efs_batch_finalize: the form, the column tile
 * width (lanes per workgroup), the z cells per lane, the occupancy the
 * sweep kernels are compiled for (waves per SIMD) and the blocks asked of
 * the equal-cost planner.  Valid shapes:
 *   launch form, 4-cell lanes: 128/256/512 lanes, 4 or 5 waves;
 *   launch form, 8-cell lanes (rows are padded to 8 cells unless the
 *   environment sets EFS_BATCH_ZPAD=4): 256/512 lanes, 3 waves;
 *   persistent: 256 lanes of 4 cells, 3 waves.
 * Fields and step count are kept. */
int efs_batch_set_shape(efs_batch *b, int mode, uint32_t tile_threads, uint32_t z_cells, uint32_t waves,
                        uint32_t blocks);
/* The shape in use; *grid = workgroups per launch (persistent: resident
 * workgroups; launch form: the block count). */
int efs_batch_get_shape(efs_batch *b, int *mode, uint32_t *tile_threads, uint32_t *z_cells, uint32_t *waves,
                        uint32_t *blocks, uint32_t *grid);

/* Zero every field, Mur state and gather ring and set numTS to 0. */
int efs_batch_reset(efs_batch *b);

/* Advance every instance by n steps (asynchronous; full graph blocks are
 * replayed, the rest launched directly). */
int efs_batch_iterate(efs_batch *b, uint32_t n);
int efs_batch_sync(efs_batch *b);
int efs_batch_num_ts(efs_batch *b, uint32_t *out);

/* Buffered gather of instance id: out[s*2n + 2k] = volt, [.. + 1] = curr
 * for sample s and probe index k (n = its probe count); *num_samples gets
 * the count.  out may be NULL to query the count only. */
int efs_batch_read_gather(efs_batch *b, uint32_t id, float *out, uint32_t *num_samples);

/* getFields of instance id (3*nx*ny*nz floats each, NIJK, real units). */
int efs_batch_get_fields(efs_batch *b, uint32_t id, float *volt, float *curr);
/* Grid size (dims[3], unpadded) and number of gathered indices of instance
 * id, for bindings that check the caller's buffers. */
int efs_batch_sizes(efs_batch *b, uint32_t id, uint32_t *dims, uint32_t *gather_n);

/* Launch shape in use: blocks per sweep launch, waves per SIMD of the sweep
 * kernels, steps per graph replay, total cells (sum over instances). */
int efs_batch_info(efs_batch *b, uint32_t *blocks, uint32_t *waves, uint32_t *graph_steps, uint64_t *cells);

/* Where the Mur entries went at finalize (no counterpart in the reference; a
 * diagnostic): entries applied by the point kernels, and z-face entries the
 * E sweep applies itself (0 with EFS_BATCH_ZMUR=0).  Both give the same
 * numbers. */
int efs_batch_mur_split(efs_batch *b, uint32_t *listed, uint32_t *in_sweep);

/* Coefficient form of instance id (no counterpart in the reference; a
 * diagnostic): *idx16 = 1 when both sides use 16-bit indices into tables,
 * 0 when they are FP32 arrays. */
int efs_batch_coef_form(efs_batch *b, uint32_t id, uint32_t *idx16);

/* Per-sweep device timing (HIP events on the batch stream): enable, then
 * read the summed milliseconds and counts of the E and H sweep launches
 * made while enabled (graph replays are not split: with profiling on,
 * iterate launches steps directly). */
int efs_batch_profile(efs_batch *b, int enable);
int efs_batch_profile_read(efs_batch *b, double *ms_e, uint32_t *n_e, double *ms_h, uint32_t *n_h);

#ifdef __cplusplus
}
#endif

#endif /* EFS_BATCH_H */
