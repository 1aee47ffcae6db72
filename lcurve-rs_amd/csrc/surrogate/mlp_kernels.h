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
};

struct AdamParams {
  float lr, beta1, beta2, eps, bc1, bc2;  // bc = 1 - beta^t
};

hipError_t launch_gemm(const GemmArgs& a, bool a_kc, bool b_kc, int epi, hipStream_t st);
hipError_t launch_loss(const float* y, const float* t, uint16_t* dy, size_t count, double* partial,
                       int blocks, hipStream_t st);
// partial: (rows / 64) x cols floats of work space.
hipError_t launch_bias_grad(const uint16_t* dz, int rows, int cols, float* partial, float* db, hipStream_t st);
hipError_t launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, const uint8_t* mask,
                       size_t n, const AdamParams& p, hipStream_t st);
hipError_t launch_gather(const float* X, const float* Y, const int* idx, int batch, int din, int din_pad,
                         int dout, uint16_t* xb, float* yb, hipStream_t st);
hipError_t launch_pack(const float* x, uint16_t* y, size_t n, hipStream_t st);

}  // namespace mlp
}  // namespace lch
