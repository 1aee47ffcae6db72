// Host side of the light-curve surrogate: a multilayer perceptron
// (GELU hidden layers, linear output) trained and evaluated on the GPU by
// the BF16 MFMA kernels of mlp_kernels.hip. Weights are FP32 masters with a
// BF16 copy for the GEMMs; activations are BF16, accumulation FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../engine/engine.h"
#include "mlp_kernels.h"

namespace lch {

class Mlp {
 public:
  // dims = {in, h1, ..., out}; every dimension but `in` must be a multiple
  // of 64 (`in` is zero-padded to 64). max_batch: rows per training step or
  // prediction chunk, a multiple of 64.
  Mlp(Context& ctx, const std::vector<int>& dims, uint64_t seed, int max_batch);
  ~Mlp();
  Mlp(const Mlp&) = delete;
  Mlp& operator=(const Mlp&) = delete;

  int n_layers() const { return static_cast<int>(dims_.size()) - 1; }
  int in_dim() const { return dims_.front(); }
  int out_dim() const { return dims_.back(); }
  // Parameters in order W0 (out x in), b0, W1, b1, ...; W0 without padding.
  size_t param_count() const;
  void get_params(float* out) const;
  void set_params(const float* in);

  // Resident training set: x [n x in], y [n x out] (FP32, host).
  void set_data(const float* x, const float* y, int n);
  // One Adam step per batch of `perm` (indices into the training set); a
  // trailing partial batch is dropped. Returns the mean loss of the steps.
  double train_epoch(const int* perm, int n, int batch, float lr);
  // Loss and gradients (same order as the parameters) of one batch of host
  // rows, without updating. n must be a multiple of 64 and <= max_batch.
  double loss_and_grad(const float* x, const float* y, int n, float* grads);
  // Adam step with the gradients of the last loss_and_grad / train step.
  void adam_step(float lr);
  // Forward pass of n host rows (any n), y [n x out].
  void predict(const float* x, int n, float* y);
  // Forward pass of `rows` device rows x [rows x in] (rows a multiple of 64,
  // at most max_batch), queued on the context stream without waiting; the
  // outputs are then in output() [rows x out].
  void forward_device(const float* x, int rows);
  const float* output() const { return Z_.back(); }
  int max_batch() const { return max_batch_; }
  Context& context() const { return ctx_; }
  // Mean time of `reps` training steps on resident data, and of their GEMMs
  // alone (HIP events); flops per step of the GEMMs.
  void time_step(int batch, int reps, double* ms_step, double* ms_gemm, double* gemm_flop);

 private:
  void init(uint64_t seed);
  void release();  // frees every device resource; safe on a partly built network
  void forward(int batch);
  void backward(int batch);
  void gemm(const mlp::GemmArgs& a, bool a_kc, bool b_kc, int epi);
  void run_step(int batch, float lr, bool adam, bool gemm_only);
  void check(hipError_t e, const char* what) const;
  void copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) const;

  Context& ctx_;
  std::vector<int> dims_;  // real dims
  std::vector<int> pdims_; // padded dims (in -> 64)
  int max_batch_;
  // Parameters per layer.
  std::vector<float*> W_, b_, gW_, gb_, mW_, vW_, mb_, vb_;
  std::vector<uint16_t*> Wb_, Wbt_;  // BF16 weights [out x in] and transposed [in x out]
  uint8_t* mask0_ = nullptr;  // layer-0 weight mask (padding columns frozen)
  // Activations: A_[l] BF16 input of layer l [B x pdims[l]]; Z_[l] FP32
  // pre-activation of layer l [B x pdims[l+1]]; dZ_[l] BF16 gradient wrt Z_[l].
  std::vector<uint16_t*> A_, dZ_;
  std::vector<float*> Z_;
  float* yb_ = nullptr;      // targets of the batch [B x out]
  float* xin_ = nullptr;     // host-supplied inputs [max_batch x in]
  double* partial_ = nullptr;
  uint16_t* dZt_ = nullptr;  // transposed output gradient of the layer in work
  uint16_t* At_ = nullptr;   // transposed input activations of that layer
  static constexpr int kMaxSplit = 16;
  float* splitk_ = nullptr;  // split-K partial weight gradients [kMaxSplit x wmax x wmax]
  bool gemm_only_ = false;
  int loss_blocks_ = 256;
  // Resident training set.
  float* X_ = nullptr;
  float* Y_ = nullptr;
  int* idx_ = nullptr;
  int* ident_ = nullptr;  // 0, 1, ..., max_batch - 1
  int n_data_ = 0;
  int t_ = 0;  // Adam step count
  hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<void*> allocs_;
  template <class T>
  T* alloc(size_t n);
};

}  // namespace lch
