// Kernel interface of the surrogate MLP (mlp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace lch {
namespace mlp {

// GEMM epilogues.
enum : int {
  EPI_PLAIN = 0,      // Cf = acc                                   (weight gradients)
  EPI_BIAS = 1,       // Cf = acc + bias[n]                          (output layer)
  EPI_BIAS_GELU = 2,  // Cf = z = acc + bias[n]; Cb = bf16(gelu(z))  (hidden layers)
  EPI_GELU_GRAD = 3,  // Cb = bf16(acc * gelu'(Z[m][n]))             (backward through GELU)
};

// C[M x N] = A[M x K] . B[K x N]. A K-contiguous: A[m * lda + k], else
// A[k * lda + m]; B K-contiguous: B[n * ldb + k], else B[k * ldb + n].
// BF16 operands are raw uint16_t bit patterns.
struct GemmArgs {
  const uint16_t* A = nullptr;
  const uint16_t* B = nullptr;
  int M = 0, N = 0, K = 0, lda = 0, ldb = 0, ldc = 0;
  const float* bias = nullptr;  // EPI_BIAS, EPI_BIAS_GELU
  const float* Z = nullptr;     // EPI_GELU_GRAD: pre-activations [M x ldc]
  float* Cf = nullptr;
  uint16_t* Cb = nullptr;
  // EPI_PLAIN only: K split into ksplit slices whose partial products go to
  // ksplit consecutive M x ldc planes of Cf (reduced by launch_splitk_sum).
  int ksplit = 1;
};

struct AdamParams {
  float lr, beta1, beta2, eps, bc1, bc2;  // bc = 1 - beta^t
};

// Parameter tensors updated by one Adam launch. Tensor k holds elements
// [start[k], start[k+1]) of the launch; wb / wbt (BF16 copy and its
// transpose, rows x cols) are null for bias vectors.
constexpr int ADAM_MAX_TENSORS = 24;
struct AdamTable {
  float* w[ADAM_MAX_TENSORS];
  const float* g[ADAM_MAX_TENSORS];
  float* m[ADAM_MAX_TENSORS];
  float* v[ADAM_MAX_TENSORS];
  uint16_t* wb[ADAM_MAX_TENSORS];
  uint16_t* wbt[ADAM_MAX_TENSORS];
  const uint8_t* mask[ADAM_MAX_TENSORS];
  int rows[ADAM_MAX_TENSORS], cols[ADAM_MAX_TENSORS];
  size_t start[ADAM_MAX_TENSORS + 1];
  int count;
};

hipError_t launch_gemm(const GemmArgs& a, bool a_kc, bool b_kc, int epi, hipStream_t st);
hipError_t launch_loss(const float* y, const float* t, uint16_t* dy, size_t count, double* partial,
                       int blocks, hipStream_t st);
hipError_t launch_gather(const float* X, const float* Y, const int* idx, int batch, int din, int din_pad,
                         int dout, uint16_t* xb, float* yb, hipStream_t st);
hipError_t launch_pack(const float* x, uint16_t* y, size_t n, hipStream_t st);
hipError_t launch_transpose(const uint16_t* in, int rows, int cols, uint16_t* out, hipStream_t st);
hipError_t launch_rowsum(const uint16_t* x, int rows, int cols, float* out, hipStream_t st);
hipError_t launch_splitk_sum(const float* part, int nsplit, size_t n, float* out, hipStream_t st);
hipError_t launch_adam_multi(const AdamTable& t, const AdamParams& p, hipStream_t st);

}  // namespace mlp
}  // namespace lch
