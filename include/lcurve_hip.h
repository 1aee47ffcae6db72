/*
 * lcurve_hip.h — C ABI of the MI355X (gfx950) light-curve flux engine.
 *
 * Drop-in replacement for the GPU backend of lcurve-rs and for the CPU
 * flux loop behind its light-curve entry points. Every entry point returns
 * an int status (LCH_OK = 0, negative LCH_ERR_* on failure; HIP error codes
 * are never passed through raw) and records a message retrievable with
 * lch_last_error() on the calling thread. Host buffers are caller-owned;
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
  int persistent;         /* 1 if the persistent launch form is selected */
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
const char* lch_last_error(void);

/* ---- device context ---- */
int lch_device_count(void);
int lch_ctx_create(int device, lch_ctx** out);
void lch_ctx_destroy(lch_ctx* ctx);
/* Override the launch tuning (chunk > 0 and persistent in {0,1}). */
int lch_ctx_set_tuning(lch_ctx* ctx, int chunk, int persistent);
/* Limb lanes for batches built from now on: a no-eclipse (patch, tile)
 * pair is split per lane into moments / limb lanes / nothing when at most
 * max_part (1..64, default 48) of its lanes see the patch in part; 0 turns
 * the split off (all 64 lanes of the tile run the element loop). Results
 * agree to FP32 rounding either way. No counterpart in the reference. */
int lch_ctx_set_limb(lch_ctx* ctx, int max_part);
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
 * network's output j, exponentiated; then wt = w / ferr^2,
 * s = sum(m wt y) / sum(m^2 wt) (1 if not positive) and
 * logl = -0.5 sum(w ((y - s m) / ferr)^2), all FP64 but the network.
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
