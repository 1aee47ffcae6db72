// Host side of the surrogate log-likelihood: frame coordinates and
// monomials on the host (a few hundred flops per row), one upload per chunk
// of rows, the network's forward pass and the curve / chi-squared kernels
// on the context stream, one download of the results.
#include "surr_lik.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace lch {

namespace {

// Nondecreasing index tuples of length k over [0, d) in lexicographic
// order (itertools.combinations_with_replacement).
void monomials(int d, int k, int start, std::vector<int>& cur, std::vector<std::vector<int>>& out) {
  if (static_cast<int>(cur.size()) == k) {
    out.push_back(cur);
    return;
  }
  for (int i = start; i < d; ++i) {
    cur.push_back(i);
    monomials(d, k, i, cur, out);
    cur.pop_back();
  }
}

std::vector<std::vector<int>> all_monomials(int d, int degree) {
  std::vector<std::vector<int>> out;
  std::vector<int> cur;
  for (int k = 1; k <= degree; ++k) monomials(d, k, 0, cur, out);
  return out;
}

}  // namespace

int SurrogateLik::n_features(int ndim, int degree) {
  return 1 + static_cast<int>(all_monomials(ndim, degree).size());
}

void SurrogateLik::check(hipError_t e, const char* what) const {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

SurrogateLik::SurrogateLik(Mlp& net, int ndim, const double* center, const double* M, int degree,
                           const double* coef, int nbins, double y_std, const double* y, const double* ferr,
                           const double* w)
    : net_(net), ndim_(ndim), nb_(nbins), y_std_(y_std) {
  if (ndim <= 0 || ndim != net.in_dim()) throw std::invalid_argument("ndim must equal the network's input width");
  if (degree < 0 || degree > 6) throw std::invalid_argument("degree must be in 0..6");
  if (nbins <= 0 || nbins > net.out_dim()) throw std::invalid_argument("nbins must be in 1..network output width");
  // Points of weight <= 0 take no part in re_scale, whatever their error.
  for (int j = 0; j < nbins; ++j)
    if (w[j] > 0.0 && !(ferr[j] > 0.0)) throw std::invalid_argument("ferr must be positive where the weight is");
  mono_ = all_monomials(ndim, degree);
  nf_ = 1 + static_cast<int>(mono_.size());
  if (nf_ > 768) throw std::invalid_argument("too many polynomial features");
  center_.assign(center, center + ndim);
  M_.assign(M, M + static_cast<size_t>(ndim) * ndim);
  chunk_ = std::min(net.max_batch(), 256);
  Context& ctx = net.context();
  check(hipSetDevice(ctx.device()), "hipSetDevice");
  const size_t xb = static_cast<size_t>(chunk_) * ndim * sizeof(float);
  phi_off_ = (xb + 255) / 256 * 256;
  stage_bytes_ = phi_off_ + static_cast<size_t>(chunk_) * nf_ * sizeof(double);
  try {
    init(coef, y, ferr, w);
  } catch (...) {
    release();
    throw;
  }
}

void SurrogateLik::init(const double* coef, const double* y, const double* ferr, const double* w) {
  Context& ctx = net_.context();
  auto dalloc = [&](void** p, size_t bytes) { check(hipMalloc(p, std::max<size_t>(bytes, 16)), "hipMalloc"); };
  dalloc(reinterpret_cast<void**>(&coef_), static_cast<size_t>(nf_) * nb_ * sizeof(double));
  dalloc(reinterpret_cast<void**>(&y_), nb_ * sizeof(double));
  dalloc(reinterpret_cast<void**>(&ferr_), nb_ * sizeof(double));
  dalloc(reinterpret_cast<void**>(&w_), nb_ * sizeof(double));
  dalloc(reinterpret_cast<void**>(&curve_), static_cast<size_t>(chunk_) * nb_ * sizeof(double));
  dalloc(reinterpret_cast<void**>(&out_), chunk_ * sizeof(double));
  dalloc(reinterpret_cast<void**>(&stage_), stage_bytes_);
  check(hipHostMalloc(reinterpret_cast<void**>(&hstage_), stage_bytes_, hipHostMallocDefault), "hipHostMalloc");
  check(hipHostMalloc(reinterpret_cast<void**>(&hout_), chunk_ * sizeof(double), hipHostMallocDefault),
        "hipHostMalloc");
  std::memset(hstage_, 0, stage_bytes_);
  hipStream_t st = ctx.stream();
  check(hipMemcpyAsync(coef_, coef, static_cast<size_t>(nf_) * nb_ * sizeof(double), hipMemcpyHostToDevice, st),
        "hipMemcpyAsync");
  check(hipMemcpyAsync(y_, y, nb_ * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  check(hipMemcpyAsync(ferr_, ferr, nb_ * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  check(hipMemcpyAsync(w_, w, nb_ * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

SurrogateLik::~SurrogateLik() { release(); }

void SurrogateLik::release() {
  Context& ctx = net_.context();
  (void)hipSetDevice(ctx.device());
  // A device-wide wait before the frees: after a stream wait alone the
  // runtime can keep freed memory allocated (lch::Batch::release_all).
  (void)hipDeviceSynchronize();
  for (void* p : {static_cast<void*>(coef_), static_cast<void*>(y_), static_cast<void*>(ferr_),
                  static_cast<void*>(w_), static_cast<void*>(curve_), static_cast<void*>(out_),
                  static_cast<void*>(stage_)})
    if (p) (void)hipFree(p);
  coef_ = y_ = ferr_ = w_ = curve_ = out_ = nullptr;
  stage_ = nullptr;
  if (hstage_) (void)hipHostFree(hstage_);
  if (hout_) (void)hipHostFree(hout_);
  hstage_ = nullptr;
  hout_ = nullptr;
}

void SurrogateLik::eval(const double* params, int n, double* logl) {
  if (n < 0) throw std::invalid_argument("negative row count");
  Context& ctx = net_.context();
  check(hipSetDevice(ctx.device()), "hipSetDevice");
  hipStream_t st = ctx.stream();
  float* hx = reinterpret_cast<float*>(hstage_);
  double* hphi = reinterpret_cast<double*>(hstage_ + phi_off_);
  std::vector<double> u(ndim_);
  for (int r0 = 0; r0 < n; r0 += chunk_) {
    const int rows = std::min(chunk_, n - r0);
    const int padded = (rows + 63) / 64 * 64;
    for (int r = 0; r < rows; ++r) {
      const double* p = params + static_cast<size_t>(r0 + r) * ndim_;
      for (int i = 0; i < ndim_; ++i) {
        double s = 0.0;
        for (int k = 0; k < ndim_; ++k) s += (p[k] - center_[k]) * M_[static_cast<size_t>(k) * ndim_ + i];
        u[i] = s;
        hx[static_cast<size_t>(r) * ndim_ + i] = static_cast<float>(s);
      }
      double* ph = hphi + static_cast<size_t>(r) * nf_;
      ph[0] = 1.0;
      for (size_t f = 0; f < mono_.size(); ++f) {
        double v = u[mono_[f][0]];
        for (size_t t = 1; t < mono_[f].size(); ++t) v *= u[mono_[f][t]];
        ph[f + 1] = v;
      }
    }
    // Rows past `rows` up to `padded` run through the network on zeros.
    std::memset(hx + static_cast<size_t>(rows) * ndim_, 0, static_cast<size_t>(padded - rows) * ndim_ * sizeof(float));
    check(hipMemcpyAsync(stage_, hstage_, phi_off_ + static_cast<size_t>(rows) * nf_ * sizeof(double),
                         hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
    net_.forward_device(reinterpret_cast<const float*>(stage_), padded);
    check(surr::launch_curve(reinterpret_cast<const double*>(stage_ + phi_off_), nf_, coef_, net_.output(),
                             net_.out_dim(), y_std_, nb_, rows, curve_, st),
          "surrogate curve launch");
    check(surr::launch_chisq(curve_, y_, ferr_, w_, nb_, rows, out_, st), "surrogate chisq launch");
    check(hipMemcpyAsync(hout_, out_, rows * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    std::memcpy(logl + r0, hout_, rows * sizeof(double));
  }
}

}  // namespace lch
