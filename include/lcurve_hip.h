/*
 * lcurve_hip.h — C ABI of the MI355X (gfx950) light-curve flux engine.
 *
 * Drop-in replacement for the GPU backend of lcurve-rs and for the CPU
 * flux loop behind its light-curve entry points. Every entry point returns
 * an int status (LCH_OK = 0, negative LCH_ERR_* on failure; HIP error codes
 * are never passed through raw) and records a message retrievable with
 * lch_last_error() (per context, and per calling thread). Host buffers are caller-owned;
 * device memory is owned by the context / batch and freed by its destroy
 * call. All pointers are plain host pointers; no HIP or torch types appear.
 *
 * Reference interfaces replaced (paths relative to the lcurve-rs tree):
 *   lch_device_count   <- lcuda_device_count      crates/lcurve-cuda/kernels/flux_kernel.cu:557-562
 *   lch_ctx_create     <- CudaContext::new        crates/lcurve-cuda/src/context.rs:261-276
 *                         lcuda_init               crates/lcurve-cuda/kernels/flux_kernel.cu:494-496
 *   lch_ctx_destroy    <- Drop for CudaContext    crates/lcurve-cuda/src/context.rs:148-252
 *   lch_chisq_batch    <- chisq_batch_gpu         crates/lcurve-cuda/src/batch.rs:73-195
 *                         chisq_batch (CPU)        crates/lcurve/src/orchestration.rs:329-357
 *   lch_light_curve    <- light_curve_comp        crates/lcurve/src/orchestration.rs:36-318
 *   lch_batch_build    <- prepare_model + upload  crates/lcurve-cuda/src/batch.rs:129-156, 227-531
 *   lch_batch_launch   <- lcuda_flux_eval_stream  crates/lcurve-cuda/kernels/flux_kernel.cu:594-615
 *   lch_batch_run      <- launch + D2H + compute_chisq  crates/lcurve-cuda/src/batch.rs:174-222, 533-568
 *   lch_ctx_sync       <- lcuda_stream_sync       crates/lcurve-cuda/kernels/flux_kernel.cu:577-579
 *   lch_model_*        <- Model::from_file / set_param_value / get_param_value
 *                                                 crates/lcurve/src/model.rs:186-532
 *   lch_data_*         <- read_data               crates/lcurve/src/types.rs:229-254
 *   lch_model_upload   <- flatten_points + upload_model_async
 *                                                 crates/lcurve-cuda/src/transfer.rs:134-159,
 *                                                 crates/lcurve-cuda/src/batch.rs:198-231
 *   lch_flux_eval      <- lcuda_flux_eval         crates/lcurve-cuda/kernels/flux_kernel.cu:534-555
 *   lch_model_prepare  <- prepare_model (host)    crates/lcurve-cuda/src/batch.rs:233-501
 *   lch_dpack_*        <- chisq_batch_gpu's flatten + upload + kernel per set, after prepare_model
 *                                                 crates/lcurve-cuda/src/batch.rs:73-136
 * The reference's lcuda_malloc / memcpy / stream / host_alloc exports exist
 * because its Rust host manages device buffers; here the library owns all
 * device memory, so they have no counterpart.
 */
#ifndef LCURVE_HIP_H
#define LCURVE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define LCH_OK 0
#define LCH_ERR_INVALID (-1)     /* bad argument */
#define LCH_ERR_PARSE (-2)       /* model or data file could not be parsed */
#define LCH_ERR_SETUP (-3)       /* model setup failed (e.g. star larger than its Roche lobe) */
#define LCH_ERR_DEVICE (-4)      /* no usable HIP device */
#define LCH_ERR_RUNTIME (-5)     /* HIP runtime failure during a call */
#define LCH_ERR_UNSUPPORTED (-6) /* model feature the engine does not implement */
#define LCH_ERR_NOMEM (-7)       /* host allocation failed */

typedef struct lch_ctx lch_ctx;
typedef struct lch_model lch_model;
typedef struct lch_batch lch_batch;

/* Scalars of one light curve (orchestration.rs:9-28 LcResult). */
typedef struct lch_lc_result {
  double wdwarf, chisq, wnok, logg1, logg2, rv1, rv2;
  double sfac[5];
} lch_lc_result;

typedef struct lch_batch_stats {
  long long n_sets, n_valid, n_times, n_lanes, n_items;
  long long visits;       /* element-visits in the reference's loop order */
  long long record_bytes; /* resident element records and patch moments in HBM */
  long long pairs_evaluated; /* (element, lane) pairs the flux kernel evaluates one
                                by one per launch, padding included */
  long long pairs_by_moments; /* pairs summed in closed form from patch moments
                                 (seen whole, no eclipse, on the whole tile or
                                 on single lanes); the rest of the reference's
                                 pairs are skipped as exactly zero */
  double setup_seconds;   /* CPU setup + packing wall time */
  double upload_seconds;  /* host-to-device copies and table builds */
  int chunk;              /* elements per work item */
  int persistent;         /* always 0 (one workgroup per work item) */
  long long pairs_limb;   /* of pairs_evaluated: pairs of limb lanes (lanes that see
                             a no-eclipse patch only in part, regrouped across tiles) */
  long long limb_chunks;  /* workgroups of the limb pass */
  long long moment_evals; /* moment sums evaluated per launch, counted per lane: one per
                             work item with moment patches (their summed moments),
                             one per patch split per lane */
  long long limb_items;   /* limb items (one patch and up to 64 limb lanes each) */
  long long limb_slots;   /* (element, lane) slots the limb pass computes, padding
                             included: pairs_limb / limb_slots is its lane fill */
} lch_batch_stats;

int lch_version(void);
/* Message of the last failing call: with a context, the last call made with
 * that context that failed; with NULL, the last failing call of the calling
 * thread (the only record for calls that take no context). */
const char* lch_last_error(const lch_ctx* ctx);

/* ---- device context ---- */
int lch_device_count(void);
int lch_ctx_create(int device, lch_ctx** out);
void lch_ctx_destroy(lch_ctx* ctx);
/* Override the work-item size (chunk > 0 elements); persistent must be 0
 * (LCH_ERR_UNSUPPORTED otherwise: there is one launch form). */
int lch_ctx_set_tuning(lch_ctx* ctx, int chunk, int persistent);
/* Limb lanes for batches built from now on: a no-eclipse (patch, tile)
 * pair is split per lane into moments / limb lanes / nothing when at most
 * max_part (1..64, default 48) of its lanes see the patch in part; 0 turns
 * the split off (all 64 lanes of the tile run the element loop). Results
 * agree to FP32 rounding either way. No counterpart in the reference. */
int lch_ctx_set_limb(lch_ctx* ctx, int max_part);
/* Host-packed batches built from now on: 1 (default) decides the lanes the
 * split's cone test leaves undecided from the patch's elements (a lane that
 * sees every element, or none, leaves the limb lanes); 0 leaves them limb
 * lanes, the rule of the device pack (lch_dpack_*), whose tables a host
 * batch then equals bit for bit. No counterpart in the reference. */
int lch_ctx_set_limb_check(lch_ctx* ctx, int on);
int lch_ctx_sync(lch_ctx* ctx);

/* ---- model parameter files ---- */
int lch_model_load_file(const char* path, lch_model** out);
int lch_model_load_text(const char* text, lch_model** out);
int lch_model_copy(const lch_model* m, lch_model** out);
void lch_model_free(lch_model* m);
/* Physical parameters by name; LCH_ERR_INVALID for an unknown name. */
int lch_model_set_param(lch_model* m, const char* name, double value);
int lch_model_get_param(const lch_model* m, const char* name, double* value);
/* The model's `iscale` setting (model.rs:181, 440): 1 if the light curve is
 * fitted per component (lroche prints five scale factors then). */
int lch_model_iscale(const lch_model* m, int* value);
int lch_model_get_pparam(const lch_model* m, const char* name, double* value, double* range,
                         double* dstep, int* vary, int* defined);
/* Null pointers leave the field unchanged. */
int lch_model_set_pparam(lch_model* m, const char* name, const double* value, const double* range,
                         const double* dstep, const int* vary);
int lch_param_count(void);
const char* lch_param_name(int i);

/* ---- data files (time expose flux ferr weight ndiv) ---- */
int lch_data_count(const char* path, int* n);
int lch_data_read(const char* path, int n, double* time, double* expose, double* flux,
                  double* ferr, double* weight, int* ndiv);

/* ---- light curve of one model ----
 * calc_out[n] receives the fluxes (autoscaled when scale != 0). rdata != 0
 * computes chi-squared against flux/ferr/weight as light_curve_comp does. */
int lch_light_curve(lch_ctx* ctx, const lch_model* m, const double* time, const double* expose,
                    const double* flux, const double* ferr, const double* weight, const int* ndiv,
                    int n, int scale, int rdata, double* calc_out, lch_lc_result* res);

/* ---- batched chi-squared ----
 * values is row-major [nsets][ndim]; names[ndim] are parameter names (an
 * unknown name is ignored, as set_param_value's result is in the reference).
 * chisq_out[nsets]; NaN for a set whose setup fails. */
int lch_chisq_batch(lch_ctx* ctx, const lch_model* base, const double* time, const double* expose,
                    const double* flux, const double* ferr, const double* weight, const int* ndiv,
                    int n, const char* const* names, int ndim, const double* values, int nsets,
                    int scale, double* chisq_out);

/* ---- resident batches: build once, launch many times ----
 * flux2d/ferr2d/weight2d, when non-null, give each set its own observations
 * [nsets][n]. flags bit 0 keeps element maps for lch_batch_visibility. */
#define LCH_BATCH_KEEP_MAPS 1
int lch_batch_build(lch_ctx* ctx, const lch_model* base, const double* time, const double* expose,
                    const double* flux, const double* ferr, const double* weight, const int* ndiv,
                    int n, const char* const* names, int ndim, const double* values, int nsets,
                    const double* flux2d, const double* ferr2d, const double* weight2d, int flags,
                    lch_batch** out);
/* Same, with one fully specified model per set (e.g. per-set grid
 * resolution, which is not a physical parameter). */
int lch_batch_build_models(lch_ctx* ctx, const lch_model* const* models, int nsets,
                           const double* time, const double* expose, const double* flux,
                           const double* ferr, const double* weight, const int* ndiv, int n,
                           const double* flux2d, const double* ferr2d, const double* weight2d,
                           int flags, lch_batch** out);
/* Frees the batch's device tables after waiting on the whole device (a
 * wait on the launching stream alone left freed memory allocated on
 * ROCm 7.2). */
void lch_batch_free(lch_batch* b);
/* Replace the observations of every set: flux2d / ferr2d / weight2d are
 * [nsets][n] (a null array keeps the current values). */
int lch_batch_set_obs(lch_ctx* ctx, lch_batch* b, const double* flux2d, const double* ferr2d,
                      const double* weight2d);
/* Enqueue one pass (flux + assembly) on the context stream; no copy back. */
int lch_batch_launch(lch_ctx* ctx, lch_batch* b, int scale);
/* One pass, wait, copy back (any output may be null). flux is [nsets][n]. */
int lch_batch_run(lch_ctx* ctx, lch_batch* b, int scale, double* chisq, double* wnok,
                  double* sfac, double* wdwarf, double* flux);
/* Mean per-pass milliseconds of the flux kernel and of the whole pass over
 * `reps` passes, from HIP events on the context stream. */
int lch_batch_time(lch_ctx* ctx, lch_batch* b, int scale, int reps, double* ms_flux,
                   double* ms_total);
/* Try the candidate chunk sizes / launch forms and keep the fastest. */
int lch_batch_autotune(lch_ctx* ctx, lch_batch* b, int scale);
int lch_batch_stats_get(const lch_batch* b, lch_batch_stats* out);
/* Per-set scale factors of the last lch_batch_run, sfac5[nsets * 5]: the
 * per-component factors of iscale sets (star 1, disc, edge, spot, star 2;
 * orchestration.rs:262-283), else the autoscale factor (or 1) five times
 * (orchestration.rs:284-289), as LcResult.sfac. */
int lch_batch_sfac(const lch_batch* b, double* sfac5);
/* Per-set setup status and per-curve scalars; err may be null. */
int lch_batch_set_info(const lch_batch* b, int set, int* valid, double* logg1, double* logg2,
                       double* rv1, double* rv2, char* err, int errlen);
/* Eclipse visibility computed on the GPU for one grid (0 star1f, 1 star1c,
 * 2 star2f, 3 star2c, 4 disc, 5 edge, 6 spot) of one set at nphase phases:
 * mask[p * nelem + e], element order of the CPU setup. mask may be null to
 * query nelem. Needs LCH_BATCH_KEEP_MAPS. */
int lch_batch_visibility(lch_ctx* ctx, lch_batch* b, int set, int grid, const double* phases,
                         int nphase, unsigned char* mask, int* nelem);

/* ---- device-packed batches: the per-set pack on the GPU ----
 * chisq_batch_gpu (lcurve-cuda/src/batch.rs:73-136) sets up and flattens
 * every model on the CPU, then uploads each flattened set and runs the flux
 * kernel on it. Here the CPU does the setup (prepare_model, the reference's
 * method) and lays out each set (segments, lanes, per-time tables, raw
 * element grids in struct-of-arrays form, staged in pinned memory); every
 * launch then packs all sets on the GPU from those grids (element records
 * sorted into patches, patch cones and moments, work codes, work items and
 * their moment sums) and runs the flux pass and the assembly on them.
 * Tables and results equal those of lch_batch_build_models with the same
 * chunk and limb-lane threshold, bit for bit. Arguments as
 * lch_batch_build_models; chunk: elements per work item, a multiple of 64
 * in 64..4096; limb: 0 no limb lanes, else the largest number of lanes of a
 * tile (1..64) that may see a no-eclipse patch in part for the pair to be
 * split per lane (lch_ctx_set_limb). At most 65536 segments per batch (a
 * set has one per grid and eclipse-pair count, typically 5-12) and under
 * 2^32 elements, (patch, tile) pairs and work-item candidates (chunks x lane
 * tiles). */
typedef struct lch_dpack lch_dpack;
typedef struct lch_dpack_stats {
  long long n_sets, n_valid, n_segs, n_elements, n_pairs, n_patches, n_tiles, n_chunks, n_codes;
  long long n_cand;       /* (chunk, tile) pairs the item kernel considers */
  long long n_items, n_items_plain, n_items_ecl; /* work items of the last pack, by kernel form */
  long long rec_bytes;    /* element records the pack writes */
  long long input_bytes;  /* raw grids and layout tables one lch_dpack_upload copies */
  long long work_bytes;   /* device tables the pack and the flux pass write, per work-table set */
  long long visits;       /* element-visits in the reference's loop order per launch */
  long long pairs_evaluated, pairs_by_moments, moment_evals; /* the flux pass's work at the last
                             lch_dpack_run, as lch_batch_stats counts it */
  long long pairs_limb;   /* of pairs_evaluated: the pairs of limb lanes */
  long long n_limb_chunks, n_limb_items; /* limb tables of the last pack */
  long long split_pairs, split_limb_lanes; /* per-lane split at the last lch_dpack_run: (patch,
                             tile) pairs split, and their lanes that see the patch in part */
  double setup_core_s;    /* CPU setup, summed over the setup threads */
  double layout_core_s;   /* host layout and grid flattening, summed over the setup threads */
  double build_wall_s;    /* both, wall time */
  double concat_s;        /* batch-wide tables from the per-set layouts */
  double upload_s;        /* the last lch_dpack_upload */
  int chunk, limb, threads;
} lch_dpack_stats;
int lch_dpack_build(lch_ctx* ctx, const lch_model* const* models, int nsets, const double* time,
                    const double* expose, const double* flux, const double* ferr, const double* weight,
                    const int* ndiv, int n, const double* flux2d, const double* ferr2d, const double* weight2d,
                    int chunk, int limb, lch_dpack** out);
/* Frees the batch's device tables after waiting on the whole device. */
void lch_dpack_free(lch_dpack* b);
/* Copy the staged grids and tables to the device again (one copy from
 * pinned host memory) and wait; the build already did it once. */
int lch_dpack_upload(lch_ctx* ctx, lch_dpack* b);
/* Work-table sets that launches alternate between: 1 (as built) or 2. With
 * 2, each launch runs on its own stream and its pack runs beside the flux
 * pass of the launch before, at twice the work memory (lch_dpack_stats
 * work_bytes per set). lch_ctx_sync waits for every launch either way. */
int lch_dpack_set_slots(lch_ctx* ctx, lch_dpack* b, int slots);
/* Enqueue the pack, wait for its work-item counts, enqueue the flux pass
 * and the assembly; no copy back. */
int lch_dpack_launch(lch_ctx* ctx, lch_dpack* b, int scale);
/* One launch, wait, copy back (any output may be null; flux [nsets][n]). */
int lch_dpack_run(lch_ctx* ctx, lch_dpack* b, int scale, double* chisq, double* wnok, double* sfac,
                  double* wdwarf, double* flux);
/* Mean milliseconds per launch over `reps` launches: of the pack kernels
 * alone, of the flux pass alone and of the whole launch, from HIP events on
 * the context stream. */
int lch_dpack_time(lch_ctx* ctx, lch_dpack* b, int scale, int reps, double* ms_pack, double* ms_flux,
                   double* ms_total);
/* Replace the observations of every set: [nsets][n] each, a null array
 * keeps the current values. */
int lch_dpack_set_obs(lch_ctx* ctx, lch_dpack* b, const double* flux2d, const double* ferr2d,
                      const double* weight2d);
int lch_dpack_stats_get(const lch_dpack* b, lch_dpack_stats* out);
/* Pack one set on the host (the same setup; the host packer of
 * lch_batch_build_models at this batch's chunk and limb threshold) and
 * compare it with a fresh device pack: diff[0] differing record bytes,
 * [1] patch cones, [2] patch moment values, [3] work codes, [4] segment
 * flags, [5] work items, [6] item moment values and lane-mask groups;
 * diff[7] = the host's item count; [8] differing limb chunks, limb items,
 * lane indices and lane-list entries; diff[9] = the host's limb item count.
 * diff holds 10 entries. */
int lch_dpack_check(lch_ctx* ctx, lch_dpack* b, int set, long long* diff);

/* ---- grid-level boundary: the caller's own surface-element grids ----
 * The reference's GPU backend flattens lcurve's Point grids on the host and
 * sums the flux over them on the device. These entry points take such grids
 * directly, in struct-of-arrays form with every eclipse interval in FP64
 * (CSR), keep them resident in HBM and evaluate comp_light
 * (crates/lcurve/src/flux.rs:7-111) at the caller's phases. Grid order: */
#define LCH_GRID_STAR1F 0
#define LCH_GRID_STAR1C 1 /* n = 0: the coarse grid is the fine grid (the reference's clone, nlat1c == nlat1f) */
#define LCH_GRID_STAR2F 2
#define LCH_GRID_STAR2C 3 /* n = 0: likewise (orchestration.rs:72-79) */
#define LCH_GRID_DISC 4
#define LCH_GRID_EDGE 5
#define LCH_GRID_SPOT 6
/* Optional, read with gint_mode 1 only: the star grids before the opaque-disc
 * eclipses were appended, which the reference's Ginterp calibration sums use
 * (orchestration.rs:93-157); all empty = calibrate on the grids above. */
#define LCH_GRID_STAR1F_CAL 7
#define LCH_GRID_STAR1C_CAL 8
#define LCH_GRID_STAR2F_CAL 9
#define LCH_GRID_STAR2C_CAL 10
#define LCH_NGRID 11

/* One grid: the fields of Point (types.rs:7-21) as arrays. */
typedef struct lch_grid_soa {
  long long n;              /* elements (0: empty grid) */
  const double* posn;       /* [n][3] positions, units of the binary separation */
  const double* dirn;       /* [n][3] unit outward normals */
  const float* area;        /* [n] Point::area; may be NULL (the flux sum does not read it) */
  const float* gravity;     /* [n] Point::gravity; may be NULL */
  const float* flux;        /* [n] Point::flux (area times surface brightness) */
  const long long* ecl_off; /* [n + 1] CSR offsets into ingress / egress, ecl_off[0] = 0;
                               NULL: no eclipses. At most 4 intervals per element. */
  const double* ingress;    /* [ecl_off[n]] eclipse interval starts (Point::eclipse .0) */
  const double* egress;     /* [ecl_off[n]] ends (.1) */
} lch_grid_soa;

/* comp_light's scalar arguments (flux.rs:7-19) and the Ginterp state
 * (types.rs:129-202). */
typedef struct lch_flux_params {
  double iangle;               /* inclination, degrees */
  double q;                    /* mass ratio M2 / M1 */
  int limb1, limb2;            /* limb-darkening law per star: 0 Poly, 1 Claret */
  double ldc1[4], ldc2[4];     /* its four coefficients per star */
  double mucrit1, mucrit2;     /* LDC::see cut-offs */
  double lin_limb_disc, quad_limb_disc;
  double beam_factor1, beam_factor2;
  double spin1, spin2;
  double velocity_scale;       /* km/s; comp_light takes it rounded to f32 */
  int glens1;                  /* nonzero: star-2 elements lensed by star 1 */
  double rlens1;               /* lens radius (orchestration.rs:52-58) */
  double phase1, phase2;       /* Ginterp switch phases */
  /* 0: the Ginterp factors scale11..scale22 below; 1: found on the GPU from
   * the fine and coarse grids as light_curve_comp finds them
   * (orchestration.rs:85-131, comp_star1/2 at flux.rs:151 and 197) */
  int gint_mode;
  double scale11, scale12, scale21, scale22;
} lch_flux_params;

typedef struct lch_grids lch_grids;
/* Validates and uploads grids[LCH_NGRID] (records sorted into patches of 64
 * normals, patch moments); LCH_ERR_INVALID for a malformed grid. The grids
 * stay resident until lch_grids_free. */
int lch_model_upload(lch_ctx* ctx, const lch_grid_soa* grids, const lch_flux_params* params,
                     lch_grids** out);
void lch_grids_free(lch_grids* g);
/* out[i] = slfac[i] * comp_light(phase[i], expose[i], ndiv[i]) + third for
 * i < n: phases with the pdot and Roemer terms applied and exposures in
 * phase units, as light_curve_comp passes them (orchestration.rs:231-251;
 * the reference's GPU path takes the same host-computed values in
 * GpuDatum / GpuFudge, batch.rs:460-492). A call with the same phase,
 * expose and ndiv as the previous call on g reuses its work tables. Not
 * thread-safe for one g. */
int lch_flux_eval(lch_ctx* ctx, lch_grids* g, const double* phase, const double* expose, const int* ndiv,
                  const double* slfac, double third, double* out, int n);

/* No counterpart in the reference: lch_flux_eval for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The call's work tables (lanes,
 * work items, partial sums, per-time tables, results) are carved from a
 * static device array of LCH_STATIC_SCRATCH_BYTES per GPU instead of being
 * allocated: no hipMalloc / hipFree, no stream, event or graph is created
 * (the grids were allocated once by lch_model_upload). A call whose tables
 * exceed the array fails with LCH_ERR_INVALID before any kernel runs (about
 * 512 B per partial row, one row per tile of 64 sub-phases and work chunk
 * of at most 1024 elements, plus ~100 B per time and per sub-phase). All
 * calls on one GPU share the array and the last call's contents stay in it;
 * this process serialises them with a lock per GPU, so they do not race,
 * and each call waits for its own kernels before it returns. Results equal
 * lch_flux_eval's bit for bit. */
#define LCH_STATIC_SCRATCH_BYTES (64ll << 20)
int lch_flux_eval_static(lch_ctx* ctx, lch_grids* g, const double* phase, const double* expose, const int* ndiv,
                         const double* slfac, double third, double* out, int n);

/* The engine's CPU setup of a model (the reference's method: grids,
 * eclipses, brightness; orchestration.rs:36-179) in the form
 * lch_model_upload takes, with the per-time terms of `time`
 * (orchestration.rs:172-241): the host half of the reference's GPU
 * prepare_model. Needs no device. */
typedef struct lch_setup lch_setup;
int lch_model_prepare(const lch_model* m, const double* time, const double* expose, const int* ndiv, int n,
                      lch_setup** out);
/* A view of one grid; its pointers stay valid until lch_setup_free. */
int lch_setup_grid(const lch_setup* s, int grid, lch_grid_soa* view);
/* The flux parameters (gint_mode 1: the engine calibrates as lch_light_curve
 * does). */
int lch_setup_params(const lch_setup* s, lch_flux_params* out);
/* The per-time terms, arrays of n = the setup's time count; any may be NULL. */
int lch_setup_times(const lch_setup* s, int n, double* phase, double* expose, int* ndiv, double* slfac,
                    double* third);
void lch_setup_free(lch_setup* s);

/* ---- learned surrogate ----
 * No counterpart in the reference. A multilayer perceptron (GELU hidden
 * layers, linear output) trained and evaluated on the GPU by BF16 MFMA
 * GEMMs with FP32 accumulation, Adam on FP32 master weights. It maps model
 * parameters to phase-binned light curves and screens MCMC proposals;
 * the exact flux path above stays the judge of every accepted sample.
 * dims = {in, h1, ..., out}: in <= 64, the others multiples of 64;
 * max_batch a multiple of 64. Parameters are ordered W0 [out x in], b0,
 * W1, b1, ... (row-major, FP32). */
typedef struct lch_mlp lch_mlp;
int lch_mlp_create(lch_ctx* ctx, const int* dims, int ndims, unsigned long long seed, int max_batch,
                   lch_mlp** out);
void lch_mlp_free(lch_mlp* m);
long long lch_mlp_param_count(const lch_mlp* m);
int lch_mlp_get_params(const lch_mlp* m, float* out);
int lch_mlp_set_params(lch_mlp* m, const float* in);
/* Resident training set, x [n x in] and y [n x out]. */
int lch_mlp_set_data(lch_mlp* m, const float* x, const float* y, int n);
/* One Adam step per `batch` indices of perm[n] (a trailing partial batch is
 * dropped); mean_loss = mean over steps of the batch MSE. */
int lch_mlp_train_epoch(lch_mlp* m, const int* perm, int n, int batch, float lr, double* mean_loss);
/* MSE loss and its gradient (parameter order) for n rows (n a multiple of
 * 64, at most max_batch), without updating; grads may be null. */
int lch_mlp_loss_grad(lch_mlp* m, const float* x, const float* y, int n, double* loss, float* grads);
/* Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradients. */
int lch_mlp_adam_step(lch_mlp* m, float lr);
int lch_mlp_predict(lch_mlp* m, const float* x, int n, float* y);
/* Mean ms of `reps` training steps on the resident set and of their GEMMs
 * alone (HIP events), and the GEMM flops of one step. */
int lch_mlp_time_step(lch_mlp* m, int batch, int reps, double* ms_step, double* ms_gemm, double* gemm_flop);

/* ---- surrogate log-likelihood ----
 * No counterpart in the reference (whose likelihood is the exact
 * chi-squared of re_scale, crates/lcurve/src/orchestration.rs:361-388).
 * The autoscaled chi-squared of a surrogate curve at the data times, for
 * many parameter rows, in one call on the GPU: frame coordinates
 * u = (p - center) . M (M ndim x ndim row-major), the polynomial baseline
 * sum_f phi_f(u) coef[f][j] (phi: 1, then the monomials of each degree up
 * to `degree` in lexicographic index order; coef [nfeat x nbins]
 * row-major, nfeat from lch_surr_lik_n_features), plus y_std times the
 * network's output j, exponentiated; then re_scale over the points with
 * w > 0: wt = w / ferr^2, s = sum(wt y m) / sum(wt m^2) and
 * logl = -0.5 sum(w ((y - s m) / ferr)^2), or logl = 0 when no point has
 * w > 0 or sum(wt m^2) = 0 (the reference's (1, 0, 0)), all FP64 but the
 * network. ferr must be positive at every point with w > 0.
 * The network must outlive the likelihood and have in == ndim and
 * out >= nbins; training it afterwards changes the likelihood with it.
 * Not thread-safe. */
typedef struct lch_surr_lik lch_surr_lik;
int lch_surr_lik_n_features(int ndim, int degree);
int lch_surr_lik_create(lch_mlp* net, int ndim, const double* center, const double* M, int degree,
                        const double* coef, int nbins, double y_std, const double* y, const double* ferr,
                        const double* w, lch_surr_lik** out);
int lch_surr_lik_eval(lch_surr_lik* s, const double* params, int n, double* logl);
void lch_surr_lik_free(lch_surr_lik* s);

/* ---- diagnostics (CPU only, no device needed) ---- */
/* Packs one model as the engine does and re-evaluates, in FP64 on the CPU,
 * every (element, lane) pair twice: as the per-(patch, tile) work codes
 * direct (skip, closed-form patch moments, no eclipse test, full test) and
 * by brute force. max_rel_diff = max over (lane, segment) of
 * |coded - brute| / sum |terms|; counts[0..3] = (patch, tile) pairs coded
 * full / no-eclipse / skip / moments; split[0..2] = lanes of no-eclipse
 * pairs split per lane into seen whole (moments) / in part (element by
 * element) / not at all (any output may be null). No counterpart in the
 * reference: it checks this engine's own work skipping. */
int lch_selfcheck_patch_codes(const lch_model* m, const double* time, const double* expose,
                              const int* ndiv, int n, double* max_rel_diff, long long* counts,
                              long long* split);

/* Builds both stars' grids (fine and coarse) twice, with and without the
 * certified short cuts of the CPU setup (the eclipse phase scan skips steps
 * whose line of sight provably misses the eclipsing lobe; the surface-radius
 * chop decides midpoints from a certified Taylor bound), and compares them
 * bit for bit. out[0] = elements compared, out[1] = elements that differ (must
 * be 0), out[2] = scan steps skipped, out[3] = scan steps evaluated;
 * ms[0..1] = build time without / with the pre-screen (ms may be null).
 * No counterpart in the reference: it checks this engine's setup speed-ups
 * of crates/roche/src/eclipse.rs:347-373 and surface.rs:82-102. */
int lch_selfcheck_eclipse_prescreen(const lch_model* m, long long* out, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* LCURVE_HIP_H */
