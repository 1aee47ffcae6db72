// Kernels of the surrogate log-likelihood: curves of a chunk of parameter
// rows from the polynomial baseline and the network's outputs, then the
// autoscaled chi-squared of each curve (FP64 throughout).
#include <hip/hip_runtime.h>

#include "surr_lik.h"

namespace lch {
namespace surr {

namespace {

constexpr int kRows = 8;  // parameter rows per workgroup of the curve kernel

// curve[r][j] = exp(sum_f phi[r][f] coef[f][j] + y_std Z[r][j]). Grid:
// (column tiles of 256, row groups of kRows). Each thread reads its column
// of coef once for the group's rows; the group's monomials sit in LDS.
__global__ __launch_bounds__(256) void curve_kernel(const double* __restrict__ phi, int nf,
                                                    const double* __restrict__ coef, const float* __restrict__ Z,
                                                    int ldz, double y_std, int nb, int rows,
                                                    double* __restrict__ curve) {
  extern __shared__ double sphi[];  // [kRows x nf]
  const int r0 = blockIdx.y * kRows;
  const int nr = min(kRows, rows - r0);
  for (int i = threadIdx.x; i < kRows * nf; i += 256)
    sphi[i] = i < nr * nf ? phi[static_cast<size_t>(r0) * nf + i] : 0.0;
  __syncthreads();
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= nb) return;
  double acc[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) acc[r] = 0.0;
  for (int f = 0; f < nf; ++f) {
    const double c = coef[static_cast<size_t>(f) * nb + j];
#pragma unroll
    for (int r = 0; r < kRows; ++r) acc[r] = fma(sphi[r * nf + f], c, acc[r]);
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r)
    if (r < nr) {
      const size_t row = static_cast<size_t>(r0 + r);
      curve[row * nb + j] = exp(acc[r] + y_std * static_cast<double>(Z[row * ldz + j]));
    }
}

// Sum over the 256 threads of a workgroup (four 64-wide wavefronts).
__device__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return t;
}

// One workgroup per row: re_scale (orchestration.rs:361-388) of the row's
// curve against the data, then -chi^2 / 2. Points with weight <= 0 take no
// part; without any (or with sum w m^2 / ferr^2 = 0) the reference returns
// scale 1 and chi-squared 0.
__global__ __launch_bounds__(256) void chisq_kernel(const double* __restrict__ curve, const double* __restrict__ y,
                                                    const double* __restrict__ ferr, const double* __restrict__ w,
                                                    int nb, double* __restrict__ out) {
  __shared__ double red[4];
  const double* m = curve + static_cast<size_t>(blockIdx.x) * nb;
  double sdy = 0.0, syy = 0.0, wn = 0.0;
  for (int j = threadIdx.x; j < nb; j += 256) {
    if (w[j] > 0.0) {
      const double wt = w[j] / (ferr[j] * ferr[j]);
      sdy += wt * y[j] * m[j];
      syy += wt * m[j] * m[j];
      wn += w[j];
    }
  }
  sdy = block_sum(sdy, red);
  syy = block_sum(syy, red);
  wn = block_sum(wn, red);
  double chi = 0.0;
  if (wn > 0.0 && syy > 0.0) {
    const double s = sdy / syy;
    for (int j = threadIdx.x; j < nb; j += 256) {
      if (w[j] > 0.0) {
        const double r = (y[j] - m[j] * s) / ferr[j];
        chi += w[j] * (r * r);
      }
    }
    chi = block_sum(chi, red);
  }
  if (threadIdx.x == 0) out[blockIdx.x] = -0.5 * chi;
}

}  // namespace

hipError_t launch_curve(const double* phi, int nf, const double* coef, const float* Z, int ldz, double y_std, int nb,
                        int rows, double* curve, hipStream_t st) {
  if (rows <= 0 || nb <= 0 || nf <= 0 || kRows * nf * sizeof(double) > 48 * 1024) return hipErrorInvalidValue;
  const dim3 grid((nb + 255) / 256, (rows + kRows - 1) / kRows);
  hipLaunchKernelGGL(curve_kernel, grid, dim3(256), kRows * nf * sizeof(double), st, phi, nf, coef, Z, ldz, y_std,
                     nb, rows, curve);
  return hipGetLastError();
}

hipError_t launch_chisq(const double* curve, const double* y, const double* ferr, const double* w, int nb, int rows,
                        double* out, hipStream_t st) {
  if (rows <= 0 || nb <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(chisq_kernel, dim3(rows), dim3(256), 0, st, curve, y, ferr, w, nb, out);
  return hipGetLastError();
}

}  // namespace surr
}  // namespace lch
