// Log-likelihood of a light-curve surrogate against observed data, on the
// GPU. The surrogate's curve at the data times is
//   log(model_j) = sum_f phi_f(u) coef[f][j] + y_std * net(u)_j,
// with u = (p - center) . M the frame coordinates of parameter row p and
// phi the monomials of u up to `degree` (1, u_i, u_i u_j with i <= j, ...,
// each degree in lexicographic order). The likelihood is the autoscaled
// chi-squared of re_scale (crates/lcurve/src/orchestration.rs:361-388):
//   wt = w / ferr^2, s = sum(model wt y) / sum(model^2 wt) (1 if the
//   denominator is not positive), logL = -0.5 sum(w ((y - s model)/ferr)^2).
// One call uploads the frame coordinates and monomials of a chunk of rows,
// runs the network's forward pass, forms the curves and reduces them, all
// on the context stream, and copies back one double per row.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mlp.h"

namespace lch {

class SurrogateLik {
 public:
  // net: in_dim() == ndim, out_dim() >= nbins (outputs past nbins unused).
  // M is ndim x ndim row-major; coef is n_features(ndim, degree) x nbins
  // row-major; y, ferr, w have nbins entries (ferr > 0).
  SurrogateLik(Mlp& net, int ndim, const double* center, const double* M, int degree, const double* coef, int nbins,
               double y_std, const double* y, const double* ferr, const double* w);
  ~SurrogateLik();
  SurrogateLik(const SurrogateLik&) = delete;
  SurrogateLik& operator=(const SurrogateLik&) = delete;

  static int n_features(int ndim, int degree);
  // logl[n] for parameter rows params[n x ndim]. Not thread-safe; the
  // network must not be trained concurrently.
  void eval(const double* params, int n, double* logl);

 private:
  // Device buffers and the uploads of coef, y, ferr and w.
  void init(const double* coef, const double* y, const double* ferr, const double* w);
  void release();  // frees every buffer; safe on a partly built object
  void check(hipError_t e, const char* what) const;

  Mlp& net_;
  int ndim_, nb_, nf_, chunk_;
  double y_std_;
  std::vector<double> center_, M_;
  std::vector<std::vector<int>> mono_;  // variable indices of each monomial past the constant
  double* coef_ = nullptr;
  double* y_ = nullptr;
  double* ferr_ = nullptr;
  double* w_ = nullptr;
  double* curve_ = nullptr;  // [chunk x nbins]
  double* out_ = nullptr;    // [chunk]
  char* stage_ = nullptr;    // device: x (FP32) [chunk x ndim], then phi [chunk x nf]
  char* hstage_ = nullptr;   // pinned host copy of the same layout
  double* hout_ = nullptr;   // pinned host [chunk]
  size_t phi_off_ = 0, stage_bytes_ = 0;
};

namespace surr {
hipError_t launch_curve(const double* phi, int nf, const double* coef, const float* Z, int ldz, double y_std, int nb,
                        int rows, double* curve, hipStream_t st);
hipError_t launch_chisq(const double* curve, const double* y, const double* ferr, const double* w, int nb, int rows,
                        double* out, hipStream_t st);
}  // namespace surr

}  // namespace lch
