// extern "C" entry points declared in include/lcurve_hip.h.
#include "../../../include/lcurve_hip.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "device_pack.h"
#include "engine.h"
#include "grid_model.h"
#include "scratch.h"
#include "selfcheck.h"
#include "../surrogate/mlp.h"
#include "../surrogate/surr_lik.h"

struct lch_ctx {
  lch::Context impl;
  std::mutex mu;    // guards err
  std::string err;  // message of the last failing call made with this context
  explicit lch_ctx(int dev) : impl(dev) {}
};
struct lch_model {
  lc::Model impl;
};
struct lch_mlp {
  std::unique_ptr<lch::Mlp> impl;
};
struct lch_surr_lik {
  std::unique_ptr<lch::SurrogateLik> impl;
};
struct lch_grids {
  std::unique_ptr<lch::GridModel> impl;
  int device = 0;
};
// CPU setup of one model exported as arrays (lch_model_prepare).
struct lch_setup {
  struct Flat {
    std::vector<double> posn, dirn, ingress, egress;
    std::vector<float> area, gravity, flux;
    std::vector<long long> off;
    long long n = 0;
  } g[LCH_NGRID];
  lch_flux_params params{};
  std::vector<double> phase, expose, slfac;
  std::vector<int> ndiv;
  double third = 0.0;
};
struct lch_dpack {
  std::unique_ptr<lch::DevicePackBatch> impl;
};
struct lch_batch {
  std::unique_ptr<lch::Batch> impl;
  double setup_s = 0, upload_s = 0;
  long long n_valid = 0;
};

namespace {

thread_local std::string g_err;
// The context of the entry point running on this thread, if it takes one.
thread_local lch_ctx* g_ctx = nullptr;

int fail(int code, const std::string& msg) {
  g_err = msg;
  if (g_ctx) {
    std::lock_guard<std::mutex> lk(g_ctx->mu);
    g_ctx->err = msg;
  }
  return code;
}

// Makes fail() record the message in `ctx` too for the rest of the call.
struct CtxScope {
  lch_ctx* prev;
  explicit CtxScope(lch_ctx* c) : prev(g_ctx) { g_ctx = c; }
  ~CtxScope() { g_ctx = prev; }
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Runs f and maps exceptions to status codes.
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const lch::DeviceError& e) {
    return fail(LCH_ERR_RUNTIME, e.what());
  } catch (const lc::LcError& e) {
    std::string m = e.what();
    return fail(m.rfind("parse error", 0) == 0 || m.rfind("I/O error", 0) == 0 ? LCH_ERR_PARSE
                                                                             : LCH_ERR_SETUP,
                m);
  } catch (const std::bad_alloc&) {
    return fail(LCH_ERR_NOMEM, "host allocation failed");
  } catch (const std::invalid_argument& e) {
    return fail(LCH_ERR_INVALID, e.what());
  } catch (const std::exception& e) {
    return fail(LCH_ERR_RUNTIME, e.what());
  }
}

lc::Data make_data(const double* time, const double* expose, const double* flux, const double* ferr,
                   const double* weight, const int* ndiv, int n) {
  if (n <= 0) throw std::invalid_argument("no data points");
  if (!time || !expose || !flux || !ferr || !weight || !ndiv)
    throw std::invalid_argument("null data array");
  lc::Data d(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    d[i].time = time[i];
    d[i].expose = expose[i];
    d[i].flux = flux[i];
    d[i].ferr = ferr[i];
    d[i].weight = weight[i];
    d[i].ndiv = ndiv[i];
  }
  return d;
}

using SetOutArr = std::vector<lch::SetOut>;

// Setup + pack + upload of a batch of models.
lch_batch* build_batch(lch_ctx* ctx, const std::vector<lc::Model>& models,
                       const std::vector<lc::Data>& data, bool keep_maps) {
  auto b = std::make_unique<lch_batch>();
  double t0 = now_s();
  auto packs = lch::build_packs(models, data, 0, keep_maps, ctx->impl.tuning.limb_check);
  double t1 = now_s();
  for (const auto& p : packs) b->n_valid += p->host.valid ? 1 : 0;
  size_t nt = data[0].size();
  b->impl = std::make_unique<lch::Batch>(ctx->impl, std::move(packs), nt);
  b->setup_s = t1 - t0;
  b->upload_s = now_s() - t1;
  return b.release();
}

std::vector<lc::Model> apply_sets(const lc::Model& base, const char* const* names, int ndim,
                                  const double* values, int nsets) {
  if (nsets < 0 || ndim < 0) throw std::invalid_argument("negative batch size");
  if (nsets > 0 && ndim > 0 && (!names || !values)) throw std::invalid_argument("null names/values");
  std::vector<lc::Model> models(static_cast<size_t>(nsets), base);
  for (int s = 0; s < nsets; ++s)
    for (int j = 0; j < ndim; ++j)
      models[s].set_param_value(names[j], values[static_cast<size_t>(s) * ndim + j]);
  return models;
}

}  // namespace

extern "C" {

int lch_version(void) { return 1; }
const char* lch_last_error(const lch_ctx* ctx) {
  if (!ctx) return g_err.c_str();
  // The pointer stays valid until the context's next failing call.
  lch_ctx* c = const_cast<lch_ctx*>(ctx);
  std::lock_guard<std::mutex> lk(c->mu);
  return c->err.c_str();
}

int lch_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int lch_ctx_create(int device, lch_ctx** out) {
  if (!out) return fail(LCH_ERR_INVALID, "null output pointer");
  *out = nullptr;
  try {
    *out = new lch_ctx(device);
    return LCH_OK;
  } catch (const std::exception& e) {
    return fail(LCH_ERR_DEVICE, e.what());
  }
}

void lch_ctx_destroy(lch_ctx* ctx) { delete ctx; }

int lch_ctx_set_tuning(lch_ctx* ctx, int chunk, int persistent) {
  CtxScope scope(ctx);
  if (!ctx || chunk <= 0) return fail(LCH_ERR_INVALID, "bad tuning");
  if (persistent != 0) return fail(LCH_ERR_UNSUPPORTED, "the persistent launch form was removed; use 0");
  ctx->impl.tuning.chunk = static_cast<uint32_t>(chunk);
  return LCH_OK;
}

int lch_ctx_set_limb(lch_ctx* ctx, int max_part) {
  CtxScope scope(ctx);
  if (!ctx || max_part < 0 || max_part > 64) return fail(LCH_ERR_INVALID, "bad limb setting");
  ctx->impl.tuning.limb = max_part;
  return LCH_OK;
}

int lch_ctx_set_limb_check(lch_ctx* ctx, int on) {
  CtxScope scope(ctx);
  if (!ctx) return fail(LCH_ERR_INVALID, "null context");
  ctx->impl.tuning.limb_check = on != 0;
  return LCH_OK;
}

int lch_ctx_sync(lch_ctx* ctx) {
  CtxScope scope(ctx);
  if (!ctx) return fail(LCH_ERR_INVALID, "null context");
  hipError_t e = hipStreamSynchronize(ctx->impl.stream());
  return e == hipSuccess ? LCH_OK : fail(LCH_ERR_RUNTIME, hipGetErrorString(e));
}

int lch_model_load_file(const char* path, lch_model** out) {
  if (!path || !out) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    *out = new lch_model{lc::Model::from_file(path)};
    return LCH_OK;
  });
}

int lch_model_load_text(const char* text, lch_model** out) {
  if (!text || !out) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    *out = new lch_model{lc::Model::from_text(text)};
    return LCH_OK;
  });
}

int lch_model_copy(const lch_model* m, lch_model** out) {
  if (!m || !out) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    *out = new lch_model{m->impl};
    return LCH_OK;
  });
}

void lch_model_free(lch_model* m) { delete m; }

int lch_model_set_param(lch_model* m, const char* name, double value) {
  if (!m || !name) return fail(LCH_ERR_INVALID, "null argument");
  if (!m->impl.set_param_value(name, value))
    return fail(LCH_ERR_INVALID, std::string("unknown parameter: '") + name + "'");
  return LCH_OK;
}

int lch_model_get_param(const lch_model* m, const char* name, double* value) {
  if (!m || !name || !value) return fail(LCH_ERR_INVALID, "null argument");
  const lc::Pparam* p = m->impl.pparam(name);
  if (!p) return fail(LCH_ERR_INVALID, std::string("unknown parameter: '") + name + "'");
  *value = p->value;
  return LCH_OK;
}

int lch_model_iscale(const lch_model* m, int* value) {
  if (!m || !value) return fail(LCH_ERR_INVALID, "null argument");
  *value = m->impl.iscale ? 1 : 0;
  return LCH_OK;
}

int lch_model_get_pparam(const lch_model* m, const char* name, double* value, double* range,
                         double* dstep, int* vary, int* defined) {
  if (!m || !name) return fail(LCH_ERR_INVALID, "null argument");
  const lc::Pparam* p = m->impl.pparam(name);
  if (!p) return fail(LCH_ERR_INVALID, std::string("unknown parameter: '") + name + "'");
  if (value) *value = p->value;
  if (range) *range = p->range;
  if (dstep) *dstep = p->dstep;
  if (vary) *vary = p->vary ? 1 : 0;
  if (defined) *defined = p->defined ? 1 : 0;
  return LCH_OK;
}

int lch_model_set_pparam(lch_model* m, const char* name, const double* value, const double* range,
                         const double* dstep, const int* vary) {
  if (!m || !name) return fail(LCH_ERR_INVALID, "null argument");
  lc::Pparam* p = m->impl.pparam(name);
  if (!p) return fail(LCH_ERR_INVALID, std::string("unknown parameter: '") + name + "'");
  if (value) p->value = *value;
  if (range) p->range = *range;
  if (dstep) p->dstep = *dstep;
  if (vary) p->vary = *vary != 0;
  return LCH_OK;
}

int lch_param_count(void) { return static_cast<int>(lc::pparam_names().size()); }

const char* lch_param_name(int i) {
  const auto& v = lc::pparam_names();
  if (i < 0 || static_cast<size_t>(i) >= v.size()) return nullptr;
  return v[static_cast<size_t>(i)].c_str();
}

int lch_data_count(const char* path, int* n) {
  if (!path || !n) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    *n = static_cast<int>(lc::read_data(path).size());
    return LCH_OK;
  });
}

int lch_data_read(const char* path, int n, double* time, double* expose, double* flux, double* ferr,
                  double* weight, int* ndiv) {
  if (!path) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    lc::Data d = lc::read_data(path);
    if (static_cast<int>(d.size()) != n) throw std::invalid_argument("data size changed");
    for (int i = 0; i < n; ++i) {
      time[i] = d[i].time;
      expose[i] = d[i].expose;
      flux[i] = d[i].flux;
      ferr[i] = d[i].ferr;
      weight[i] = d[i].weight;
      ndiv[i] = d[i].ndiv;
    }
    return LCH_OK;
  });
}

int lch_light_curve(lch_ctx* ctx, const lch_model* m, const double* time, const double* expose,
                    const double* flux, const double* ferr, const double* weight, const int* ndiv,
                    int n, int scale, int rdata, double* calc_out, lch_lc_result* res) {
  CtxScope scope(ctx);
  if (!ctx || !m || !calc_out) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    std::vector<lc::Data> data{make_data(time, expose, flux, ferr, weight, ndiv, n)};
    std::unique_ptr<lch_batch> b(build_batch(ctx, {m->impl}, data, false));
    const lch::SetHost& H = b->impl->set(0);
    if (!H.valid) throw lc::LcError(H.error);
    double chisq = 0, wnok = 0, sfac = 1, wd = 0;
    b->impl->run(scale ? 1 : 0, &chisq, &wnok, &sfac, &wd, calc_out);
    if (res) {
      res->wdwarf = wd;
      res->chisq = (scale || rdata) ? chisq : 0.0;
      res->wnok = (scale || rdata) ? wnok : 0.0;
      res->logg1 = H.logg1;
      res->logg2 = H.logg2;
      res->rv1 = H.rv1;
      res->rv2 = H.rv2;
      const SetOutArr& o = b->impl->last_results();
      for (int j = 0; j < 5; ++j) res->sfac[j] = o.empty() ? 1.0 : o[0].sfac[j];
    }
    return LCH_OK;
  });
}

int lch_chisq_batch(lch_ctx* ctx, const lch_model* base, const double* time, const double* expose,
                    const double* flux, const double* ferr, const double* weight, const int* ndiv,
                    int n, const char* const* names, int ndim, const double* values, int nsets,
                    int scale, double* chisq_out) {
  CtxScope scope(ctx);
  if (!ctx || !base || (nsets > 0 && !chisq_out)) return fail(LCH_ERR_INVALID, "null argument");
  if (ndim <= 0) return fail(LCH_ERR_INVALID, "ndim must be > 0");
  return guarded([&] {
    if (nsets == 0) return LCH_OK;
    std::vector<lc::Data> data{make_data(time, expose, flux, ferr, weight, ndiv, n)};
    auto models = apply_sets(base->impl, names, ndim, values, nsets);
    std::unique_ptr<lch_batch> b(build_batch(ctx, models, data, false));
    b->impl->run(scale ? 1 : 0, chisq_out, nullptr, nullptr, nullptr, nullptr);
    return LCH_OK;
  });
}

namespace {
// One data set shared by every set, or one per set from the [nsets][n]
// arrays given.
std::vector<lc::Data> obs_sets(int nsets, const double* time, const double* expose, const double* flux,
                               const double* ferr, const double* weight, const int* ndiv, int n,
                               const double* flux2d, const double* ferr2d, const double* weight2d) {
  std::vector<lc::Data> data;
  if (flux2d || ferr2d || weight2d) {
    data.resize(static_cast<size_t>(nsets));
    for (int s = 0; s < nsets; ++s) {
      size_t o = static_cast<size_t>(s) * n;
      data[s] = make_data(time, expose, flux2d ? flux2d + o : flux, ferr2d ? ferr2d + o : ferr,
                          weight2d ? weight2d + o : weight, ndiv, n);
    }
  } else {
    data.push_back(make_data(time, expose, flux, ferr, weight, ndiv, n));
  }
  return data;
}

lch_batch* build_with_obs(lch_ctx* ctx, const std::vector<lc::Model>& models, const double* time,
                          const double* expose, const double* flux, const double* ferr,
                          const double* weight, const int* ndiv, int n, const double* flux2d,
                          const double* ferr2d, const double* weight2d, int flags) {
  const std::vector<lc::Data> data = obs_sets(static_cast<int>(models.size()), time, expose, flux, ferr, weight,
                                              ndiv, n, flux2d, ferr2d, weight2d);
  return build_batch(ctx, models, data, (flags & LCH_BATCH_KEEP_MAPS) != 0);
}
}  // namespace

int lch_batch_build(lch_ctx* ctx, const lch_model* base, const double* time, const double* expose,
                    const double* flux, const double* ferr, const double* weight, const int* ndiv,
                    int n, const char* const* names, int ndim, const double* values, int nsets,
                    const double* flux2d, const double* ferr2d, const double* weight2d, int flags,
                    lch_batch** out) {
  CtxScope scope(ctx);
  if (!ctx || !base || !out) return fail(LCH_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    if (nsets <= 0) throw std::invalid_argument("nsets must be > 0");
    auto models = apply_sets(base->impl, names, ndim, values, nsets);
    *out = build_with_obs(ctx, models, time, expose, flux, ferr, weight, ndiv, n, flux2d, ferr2d,
                          weight2d, flags);
    return LCH_OK;
  });
}

int lch_batch_build_models(lch_ctx* ctx, const lch_model* const* models, int nsets,
                           const double* time, const double* expose, const double* flux,
                           const double* ferr, const double* weight, const int* ndiv, int n,
                           const double* flux2d, const double* ferr2d, const double* weight2d,
                           int flags, lch_batch** out) {
  CtxScope scope(ctx);
  if (!ctx || !models || !out) return fail(LCH_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    if (nsets <= 0) throw std::invalid_argument("nsets must be > 0");
    std::vector<lc::Model> ms;
    ms.reserve(static_cast<size_t>(nsets));
    for (int s = 0; s < nsets; ++s) {
      if (!models[s]) throw std::invalid_argument("null model in batch");
      ms.push_back(models[s]->impl);
    }
    *out = build_with_obs(ctx, ms, time, expose, flux, ferr, weight, ndiv, n, flux2d, ferr2d,
                          weight2d, flags);
    return LCH_OK;
  });
}

void lch_batch_free(lch_batch* b) { delete b; }

int lch_dpack_build(lch_ctx* ctx, const lch_model* const* models, int nsets, const double* time,
                    const double* expose, const double* flux, const double* ferr, const double* weight,
                    const int* ndiv, int n, const double* flux2d, const double* ferr2d, const double* weight2d,
                    int chunk, int limb, lch_dpack** out) {
  CtxScope scope(ctx);
  if (!ctx || !models || !out) return fail(LCH_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    if (nsets <= 0) throw std::invalid_argument("nsets must be > 0");
    if (chunk <= 0) throw std::invalid_argument("chunk must be > 0");
    std::vector<lc::Model> ms;
    ms.reserve(static_cast<size_t>(nsets));
    for (int s = 0; s < nsets; ++s) {
      if (!models[s]) throw std::invalid_argument("null model in batch");
      ms.push_back(models[s]->impl);
    }
    const std::vector<lc::Data> data = obs_sets(nsets, time, expose, flux, ferr, weight, ndiv, n, flux2d, ferr2d,
                                                weight2d);
    auto b = std::make_unique<lch_dpack>();
    b->impl = std::make_unique<lch::DevicePackBatch>(ctx->impl, ms, data, static_cast<uint32_t>(chunk), limb);
    *out = b.release();
    return LCH_OK;
  });
}

void lch_dpack_free(lch_dpack* b) { delete b; }

int lch_dpack_upload(lch_ctx* ctx, lch_dpack* b) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->upload();
    return LCH_OK;
  });
}

int lch_dpack_set_slots(lch_ctx* ctx, lch_dpack* b, int slots) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  if (slots < 1 || slots > 2) return fail(LCH_ERR_INVALID, "slots must be 1 or 2");
  return guarded([&] {
    b->impl->set_slots(slots);
    return LCH_OK;
  });
}

int lch_dpack_launch(lch_ctx* ctx, lch_dpack* b, int scale) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->launch(scale ? 1 : 0);
    return LCH_OK;
  });
}

int lch_dpack_run(lch_ctx* ctx, lch_dpack* b, int scale, double* chisq, double* wnok, double* sfac,
                  double* wdwarf, double* flux) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->run(scale ? 1 : 0, chisq, wnok, sfac, wdwarf, flux);
    return LCH_OK;
  });
}

int lch_dpack_time(lch_ctx* ctx, lch_dpack* b, int scale, int reps, double* ms_pack, double* ms_flux,
                   double* ms_total) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  if (reps <= 0) return fail(LCH_ERR_INVALID, "reps must be > 0");
  return guarded([&] {
    b->impl->time(scale ? 1 : 0, reps, ms_pack, ms_flux, ms_total);
    return LCH_OK;
  });
}

int lch_dpack_set_obs(lch_ctx* ctx, lch_dpack* b, const double* flux2d, const double* ferr2d,
                      const double* weight2d) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->set_obs(flux2d, ferr2d, weight2d);
    return LCH_OK;
  });
}

int lch_dpack_stats_get(const lch_dpack* b, lch_dpack_stats* out) {
  if (!b || !out) return fail(LCH_ERR_INVALID, "null argument");
  const lch::DevicePackStats& s = b->impl->stats();
  lch_dpack_stats r{};
  r.n_sets = static_cast<long long>(s.n_sets);
  r.n_valid = static_cast<long long>(s.n_valid);
  r.n_segs = static_cast<long long>(s.n_segs);
  r.n_elements = static_cast<long long>(s.n_elements);
  r.n_pairs = static_cast<long long>(s.n_pairs);
  r.n_patches = static_cast<long long>(s.n_patches);
  r.n_tiles = static_cast<long long>(s.n_tiles);
  r.n_chunks = static_cast<long long>(s.n_chunks);
  r.n_codes = static_cast<long long>(s.n_codes);
  r.n_cand = static_cast<long long>(s.n_cand);
  r.n_items = static_cast<long long>(s.n_items);
  r.n_items_plain = static_cast<long long>(s.n_items_plain);
  r.n_items_ecl = static_cast<long long>(s.n_items_ecl);
  r.rec_bytes = static_cast<long long>(s.rec_bytes);
  r.input_bytes = static_cast<long long>(s.input_bytes);
  r.work_bytes = static_cast<long long>(s.work_bytes);
  r.visits = static_cast<long long>(s.visits);
  r.pairs_evaluated = static_cast<long long>(s.pairs_evaluated);
  r.pairs_by_moments = static_cast<long long>(s.pairs_by_moments);
  r.moment_evals = static_cast<long long>(s.moment_evals);
  r.pairs_limb = static_cast<long long>(s.pairs_limb);
  r.n_limb_chunks = static_cast<long long>(s.n_limb_chunks);
  r.n_limb_items = static_cast<long long>(s.n_limb_items);
  r.split_pairs = static_cast<long long>(s.split_pairs);
  r.split_limb_lanes = static_cast<long long>(s.split_limb_lanes);
  r.setup_core_s = s.setup_core_s;
  r.layout_core_s = s.layout_core_s;
  r.build_wall_s = s.build_wall_s;
  r.concat_s = s.concat_s;
  r.upload_s = s.upload_s;
  r.chunk = static_cast<int>(s.chunk);
  r.limb = static_cast<int>(s.limb);
  r.threads = s.threads;
  *out = r;
  return LCH_OK;
}

int lch_dpack_check(lch_ctx* ctx, lch_dpack* b, int set, long long* diff) {
  CtxScope scope(ctx);
  if (!ctx || !b || !diff) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    uint64_t d[10];
    b->impl->check(set, d);
    for (int i = 0; i < 10; ++i) diff[i] = static_cast<long long>(d[i]);
    return LCH_OK;
  });
}

int lch_batch_set_obs(lch_ctx* ctx, lch_batch* b, const double* flux2d, const double* ferr2d,
                      const double* weight2d) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->set_obs(flux2d, ferr2d, weight2d);
    return LCH_OK;
  });
}

int lch_batch_launch(lch_ctx* ctx, lch_batch* b, int scale) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->launch(scale ? 1 : 0);
    return LCH_OK;
  });
}

int lch_batch_run(lch_ctx* ctx, lch_batch* b, int scale, double* chisq, double* wnok, double* sfac,
                  double* wdwarf, double* flux) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->run(scale ? 1 : 0, chisq, wnok, sfac, wdwarf, flux);
    return LCH_OK;
  });
}

int lch_batch_time(lch_ctx* ctx, lch_batch* b, int scale, int reps, double* ms_flux,
                   double* ms_total) {
  CtxScope scope(ctx);
  if (!ctx || !b || reps <= 0) return fail(LCH_ERR_INVALID, "bad argument");
  return guarded([&] {
    b->impl->time(scale ? 1 : 0, reps, ms_flux, ms_total);
    return LCH_OK;
  });
}

int lch_batch_autotune(lch_ctx* ctx, lch_batch* b, int scale) {
  CtxScope scope(ctx);
  if (!ctx || !b) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    b->impl->autotune(scale ? 1 : 0);
    return LCH_OK;
  });
}

int lch_batch_stats_get(const lch_batch* b, lch_batch_stats* out) {
  if (!b || !out) return fail(LCH_ERR_INVALID, "null argument");
  const lch::Batch& B = *b->impl;
  out->n_sets = B.n_sets();
  out->n_valid = b->n_valid;
  out->n_times = static_cast<long long>(B.n_times());
  out->n_lanes = static_cast<long long>(B.n_lanes());
  out->n_items = static_cast<long long>(B.n_items());
  out->visits = static_cast<long long>(B.visits());
  out->record_bytes = static_cast<long long>(B.record_bytes());
  out->pairs_evaluated = static_cast<long long>(B.pairs_evaluated());
  out->pairs_by_moments = static_cast<long long>(B.pairs_by_moments());
  out->setup_seconds = b->setup_s;
  out->upload_seconds = b->upload_s;
  out->chunk = static_cast<int>(B.chunk());
  out->persistent = 0;
  out->pairs_limb = static_cast<long long>(B.pairs_limb());
  out->limb_chunks = static_cast<long long>(B.n_limb_chunks());
  out->moment_evals = static_cast<long long>(B.moment_evals());
  out->limb_items = static_cast<long long>(B.n_limb_items());
  out->limb_slots = static_cast<long long>(B.limb_slots());
  return LCH_OK;
}

int lch_batch_set_info(const lch_batch* b, int set, int* valid, double* logg1, double* logg2,
                       double* rv1, double* rv2, char* err, int errlen) {
  if (!b || set < 0 || static_cast<uint32_t>(set) >= b->impl->n_sets())
    return fail(LCH_ERR_INVALID, "bad set index");
  const lch::SetHost& H = b->impl->set(static_cast<uint32_t>(set));
  if (valid) *valid = H.valid ? 1 : 0;
  if (logg1) *logg1 = H.logg1;
  if (logg2) *logg2 = H.logg2;
  if (rv1) *rv1 = H.rv1;
  if (rv2) *rv2 = H.rv2;
  if (err && errlen > 0) {
    std::strncpy(err, H.error.c_str(), static_cast<size_t>(errlen - 1));
    err[errlen - 1] = 0;
  }
  return LCH_OK;
}

int lch_batch_visibility(lch_ctx* ctx, lch_batch* b, int set, int grid, const double* phases,
                         int nphase, unsigned char* mask, int* nelem) {
  CtxScope scope(ctx);
  if (!ctx || !b || !nelem) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    uint32_t n = 0;
    b->impl->visibility(set, grid, phases, nphase, mask, &n);
    *nelem = static_cast<int>(n);
    return LCH_OK;
  });
}

}  // extern "C"

int lch_selfcheck_patch_codes(const lch_model* m, const double* time, const double* expose,
                              const int* ndiv, int n, double* max_rel_diff, long long* counts,
                              long long* split) {
  if (!m) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    const std::vector<double> zero(static_cast<size_t>(std::max(n, 0)), 0.0), one(zero.size(), 1.0);
    const lc::Data data = make_data(time, expose, zero.data(), one.data(), one.data(), ndiv, n);
    const lch::SelfCheck c = lch::selfcheck_model(m->impl, data);
    if (max_rel_diff) *max_rel_diff = c.max_rel;
    if (counts)
      for (int i = 0; i < 4; ++i) counts[i] = c.counts[i];
    if (split)
      for (int i = 0; i < 3; ++i) split[i] = c.split_lanes[i];
    return LCH_OK;
  });
}

int lch_selfcheck_eclipse_prescreen(const lch_model* m, long long* out, double* ms) {
  if (!m || !out) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    const lch::PrescreenCheck c = lch::selfcheck_prescreen(m->impl);
    out[0] = c.elements;
    out[1] = c.mismatches;
    out[2] = c.skipped;
    out[3] = c.evaluated;
    if (ms) {
      ms[0] = c.ms_plain;
      ms[1] = c.ms_screened;
    }
    return LCH_OK;
  });
}

int lch_batch_sfac(const lch_batch* b, double* sfac5) {
  if (!b || !sfac5) return fail(LCH_ERR_INVALID, "null argument");
  const SetOutArr& o = b->impl->last_results();
  if (o.size() != b->impl->n_sets()) return fail(LCH_ERR_INVALID, "no results yet: call lch_batch_run first");
  for (size_t s = 0; s < o.size(); ++s)
    for (int j = 0; j < 5; ++j) sfac5[s * 5 + j] = o[s].sfac[j];
  return LCH_OK;
}

// ---------------------------------------------------------------- surrogate
int lch_mlp_create(lch_ctx* ctx, const int* dims, int ndims, unsigned long long seed, int max_batch,
                   lch_mlp** out) {
  CtxScope scope(ctx);
  if (!ctx || !dims || !out || ndims < 2) return fail(LCH_ERR_INVALID, "bad argument");
  return guarded([&] {
    auto m = std::make_unique<lch_mlp>();
    m->impl = std::make_unique<lch::Mlp>(ctx->impl, std::vector<int>(dims, dims + ndims), seed, max_batch);
    *out = m.release();
    return LCH_OK;
  });
}

void lch_mlp_free(lch_mlp* m) { delete m; }

long long lch_mlp_param_count(const lch_mlp* m) {
  return m ? static_cast<long long>(m->impl->param_count()) : -1;
}

int lch_mlp_get_params(const lch_mlp* m, float* out) {
  if (!m || !out) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    m->impl->get_params(out);
    return LCH_OK;
  });
}

int lch_mlp_set_params(lch_mlp* m, const float* in) {
  if (!m || !in) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    m->impl->set_params(in);
    return LCH_OK;
  });
}

int lch_mlp_set_data(lch_mlp* m, const float* x, const float* y, int n) {
  if (!m || !x || !y) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    m->impl->set_data(x, y, n);
    return LCH_OK;
  });
}

int lch_mlp_train_epoch(lch_mlp* m, const int* perm, int n, int batch, float lr, double* mean_loss) {
  if (!m || !perm || n < 0) return fail(LCH_ERR_INVALID, "bad argument");
  return guarded([&] {
    const double l = m->impl->train_epoch(perm, n, batch, lr);
    if (mean_loss) *mean_loss = l;
    return LCH_OK;
  });
}

int lch_mlp_loss_grad(lch_mlp* m, const float* x, const float* y, int n, double* loss, float* grads) {
  if (!m || !x || !y) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    const double l = m->impl->loss_and_grad(x, y, n, grads);
    if (loss) *loss = l;
    return LCH_OK;
  });
}

int lch_mlp_adam_step(lch_mlp* m, float lr) {
  if (!m) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    m->impl->adam_step(lr);
    return LCH_OK;
  });
}

int lch_mlp_predict(lch_mlp* m, const float* x, int n, float* y) {
  if (!m || (n > 0 && (!x || !y))) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    m->impl->predict(x, n, y);
    return LCH_OK;
  });
}

int lch_surr_lik_create(lch_mlp* net, int ndim, const double* center, const double* M, int degree,
                        const double* coef, int nbins, double y_std, const double* y, const double* ferr,
                        const double* w, lch_surr_lik** out) {
  if (!net || !center || !M || !coef || !y || !ferr || !w || !out) return fail(LCH_ERR_INVALID, "null argument");
  return guarded([&] {
    auto s = std::make_unique<lch_surr_lik>();
    s->impl = std::make_unique<lch::SurrogateLik>(*net->impl, ndim, center, M, degree, coef, nbins, y_std, y, ferr,
                                                  w);
    *out = s.release();
    return LCH_OK;
  });
}

int lch_surr_lik_n_features(int ndim, int degree) {
  if (ndim <= 0 || ndim > 64 || degree < 0 || degree > 6) return -1;
  return lch::SurrogateLik::n_features(ndim, degree);
}

int lch_surr_lik_eval(lch_surr_lik* s, const double* params, int n, double* logl) {
  if (!s || n < 0 || (n > 0 && (!params || !logl))) return fail(LCH_ERR_INVALID, "bad argument");
  return guarded([&] {
    s->impl->eval(params, n, logl);
    return LCH_OK;
  });
}

void lch_surr_lik_free(lch_surr_lik* s) { delete s; }

int lch_mlp_time_step(lch_mlp* m, int batch, int reps, double* ms_step, double* ms_gemm, double* gemm_flop) {
  if (!m || reps <= 0 || !ms_step || !ms_gemm || !gemm_flop) return fail(LCH_ERR_INVALID, "bad argument");
  return guarded([&] {
    m->impl->time_step(batch, reps, ms_step, ms_gemm, gemm_flop);
    return LCH_OK;
  });
}

// ------------------------------------------------ grid-level boundary
int lch_model_upload(lch_ctx* ctx, const lch_grid_soa* grids, const lch_flux_params* params,
                     lch_grids** out) {
  CtxScope scope(ctx);
  if (!ctx || !grids || !params || !out) return fail(LCH_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto g = std::make_unique<lch_grids>();
    g->device = ctx->impl.device();
    g->impl = std::make_unique<lch::GridModel>(ctx->impl, grids, *params);
    *out = g.release();
    return LCH_OK;
  });
}

void lch_grids_free(lch_grids* g) { delete g; }

int lch_flux_eval(lch_ctx* ctx, lch_grids* g, const double* phase, const double* expose, const int* ndiv,
                  const double* slfac, double third, double* out, int n) {
  CtxScope scope(ctx);
  if (!ctx || !g) return fail(LCH_ERR_INVALID, "null argument");
  if (g->device != ctx->impl.device()) return fail(LCH_ERR_INVALID, "grids were uploaded on another device");
  return guarded([&] {
    g->impl->eval(phase, expose, ndiv, slfac, third, out, n, nullptr);
    return LCH_OK;
  });
}

int lch_flux_eval_static(lch_ctx* ctx, lch_grids* g, const double* phase, const double* expose, const int* ndiv,
                         const double* slfac, double third, double* out, int n) {
  CtxScope scope(ctx);
  if (!ctx || !g) return fail(LCH_ERR_INVALID, "null argument");
  if (g->device != ctx->impl.device()) return fail(LCH_ERR_INVALID, "grids were uploaded on another device");
  void* ptr = ctx->impl.scratch();
  if (!ptr) return fail(LCH_ERR_RUNTIME, std::string("no static scratch on this device: ") + ctx->impl.scratch_error());
  std::lock_guard<std::mutex> lk(lch::scratch_lock(ctx->impl.device()));
  try {
    lch::BatchMemory mem;
    mem.arena = static_cast<char*>(ptr);
    mem.arena_bytes = lch::kScratchBytes;
    g->impl->eval(phase, expose, ndiv, slfac, third, out, n, &mem);
    return LCH_OK;
  } catch (const lch::ArenaFull& e) {
    return fail(LCH_ERR_INVALID, std::string(e.what()) + " (LCH_STATIC_SCRATCH_BYTES)");
  } catch (...) {
    return guarded([&]() -> int { throw; });
  }
}

int lch_model_prepare(const lch_model* m, const double* time, const double* expose, const int* ndiv, int n,
                      lch_setup** out) {
  if (!m || !out) return fail(LCH_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    if (!time || !expose || !ndiv) throw std::invalid_argument("null data array");
    if (n <= 0) throw std::invalid_argument("no data points");
    const std::vector<double> zero(static_cast<size_t>(n), 0.0), one(zero.size(), 1.0);
    const lc::Data data = make_data(time, expose, zero.data(), one.data(), zero.data(), ndiv, n);
    auto P = lc::prepare_model(m->impl, data);
    if (P->iscale) throw std::invalid_argument("iscale models are fitted per component; the grid boundary sums comp_light");
    auto S = std::make_unique<lch_setup>();
    auto flat = [](const std::vector<lc::Point>& pts, lch_setup::Flat& f) {
      f.n = static_cast<long long>(pts.size());
      f.off.assign(pts.size() + 1, 0);
      for (size_t i = 0; i < pts.size(); ++i) {
        const lc::Point& p = pts[i];
        f.posn.insert(f.posn.end(), {p.posn.x, p.posn.y, p.posn.z});
        f.dirn.insert(f.dirn.end(), {p.dirn.x, p.dirn.y, p.dirn.z});
        f.area.push_back(p.area);
        f.gravity.push_back(p.gravity);
        f.flux.push_back(p.flux);
        for (const auto& e : p.eclipse) {
          f.ingress.push_back(e.first);
          f.egress.push_back(e.second);
        }
        f.off[i + 1] = static_cast<long long>(f.ingress.size());
      }
    };
    flat(P->star1f, S->g[LCH_GRID_STAR1F]);
    if (!P->star1_alias) flat(P->star1c, S->g[LCH_GRID_STAR1C]);
    flat(P->star2f, S->g[LCH_GRID_STAR2F]);
    if (!P->star2_alias) flat(P->star2c, S->g[LCH_GRID_STAR2C]);
    flat(P->disc, S->g[LCH_GRID_DISC]);
    flat(P->edge, S->g[LCH_GRID_EDGE]);
    flat(P->spot, S->g[LCH_GRID_SPOT]);
    if (P->cal_copies) {
      if (P->need_cal1) {
        flat(P->star1f_cal, S->g[LCH_GRID_STAR1F_CAL]);
        flat(P->star1c_cal, S->g[LCH_GRID_STAR1C_CAL]);
      }
      if (P->need_cal2) {
        flat(P->star2f_cal, S->g[LCH_GRID_STAR2F_CAL]);
        flat(P->star2c_cal, S->g[LCH_GRID_STAR2C_CAL]);
      }
    }
    lch_flux_params& fp = S->params;
    const lc::Model& md = m->impl;
    fp.iangle = md.iangle.value;
    fp.q = md.q.value;
    auto ld = [](const lc::LDC& l, int* law, double* a, double* mucrit) {
      *law = l.ltype == lc::LdcType::Claret ? 1 : 0;
      a[0] = l.ldc1;
      a[1] = l.ldc2;
      a[2] = l.ldc3;
      a[3] = l.ldc4;
      *mucrit = l.mucrit;
    };
    ld(P->ldc1, &fp.limb1, fp.ldc1, &fp.mucrit1);
    ld(P->ldc2, &fp.limb2, fp.ldc2, &fp.mucrit2);
    fp.lin_limb_disc = P->lin_limb_disc;
    fp.quad_limb_disc = P->quad_limb_disc;
    fp.beam_factor1 = P->beam_factor1;
    fp.beam_factor2 = P->beam_factor2;
    fp.spin1 = P->spin1;
    fp.spin2 = P->spin2;
    fp.velocity_scale = md.velocity_scale.value;
    fp.glens1 = P->glens1 ? 1 : 0;
    fp.rlens1 = P->rlens1;
    fp.phase1 = P->phase1;
    fp.phase2 = P->phase2;
    fp.gint_mode = 1;
    fp.scale11 = fp.scale12 = fp.scale21 = fp.scale22 = 1.0;
    S->phase = P->phase;
    S->expose = P->expose;
    S->ndiv = P->ndiv;
    S->slfac = P->slfac;
    S->third = P->third;
    *out = S.release();
    return LCH_OK;
  });
}

int lch_setup_grid(const lch_setup* s, int grid, lch_grid_soa* view) {
  if (!s || !view) return fail(LCH_ERR_INVALID, "null argument");
  if (grid < 0 || grid >= LCH_NGRID) return fail(LCH_ERR_INVALID, "grid out of range");
  const lch_setup::Flat& f = s->g[grid];
  lch_grid_soa v{};
  v.n = f.n;
  if (f.n > 0) {
    v.posn = f.posn.data();
    v.dirn = f.dirn.data();
    v.area = f.area.data();
    v.gravity = f.gravity.data();
    v.flux = f.flux.data();
    v.ecl_off = f.off.data();
    v.ingress = f.ingress.data();
    v.egress = f.egress.data();
  }
  *view = v;
  return LCH_OK;
}

int lch_setup_params(const lch_setup* s, lch_flux_params* out) {
  if (!s || !out) return fail(LCH_ERR_INVALID, "null argument");
  *out = s->params;
  return LCH_OK;
}

int lch_setup_times(const lch_setup* s, int n, double* phase, double* expose, int* ndiv, double* slfac,
                    double* third) {
  if (!s) return fail(LCH_ERR_INVALID, "null argument");
  if (n != static_cast<int>(s->phase.size())) return fail(LCH_ERR_INVALID, "n differs from the setup's time count");
  for (int i = 0; i < n; ++i) {
    if (phase) phase[i] = s->phase[i];
    if (expose) expose[i] = s->expose[i];
    if (ndiv) ndiv[i] = s->ndiv[i];
    if (slfac) slfac[i] = s->slfac[i];
  }
  if (third) *third = s->third;
  return LCH_OK;
}

void lch_setup_free(lch_setup* s) { delete s; }
