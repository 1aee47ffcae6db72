// Host side of the surrogate MLP: parameter and activation buffers in HBM,
// the forward / backward / Adam launch sequences, timing.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace lch {

namespace {

// splitmix64 (Steele, Lea, Flood 2014) for reproducible initial weights.
struct SplitMix {
  uint64_t s;
  double uniform() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace

template <class T>
T* Mlp::alloc(size_t n) {
  void* p = nullptr;
  check(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 16)), "hipMalloc");
  allocs_.push_back(p);
  return static_cast<T*>(p);
}

void Mlp::check(hipError_t e, const char* what) const {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

// Every transfer is ordered on the context stream (which does not
// synchronise with the null stream) and waited for.
void Mlp::copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) const {
  check(hipMemcpyAsync(dst, src, bytes, kind, ctx_.stream()), "hipMemcpyAsync");
  check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
}

Mlp::Mlp(Context& ctx, const std::vector<int>& dims, uint64_t seed, int max_batch)
    : ctx_(ctx), dims_(dims), max_batch_(max_batch) {
  if (dims.size() < 2) throw std::invalid_argument("an MLP needs at least one layer");
  if (max_batch <= 0 || max_batch % 64) throw std::invalid_argument("max_batch must be a positive multiple of 64");
  if (dims[0] <= 0 || dims[0] > 64) throw std::invalid_argument("input dimension must be in 1..64");
  for (size_t i = 1; i < dims.size(); ++i)
    if (dims[i] <= 0 || dims[i] % 64) throw std::invalid_argument("layer widths must be positive multiples of 64");
  check(hipSetDevice(ctx_.device()), "hipSetDevice");
  // A failure part-way frees what was allocated: the destructor does not
  // run for a constructor that throws.
  try {
    init(seed);
  } catch (...) {
    release();
    throw;
  }
}

void Mlp::init(uint64_t seed) {
  const int max_batch = max_batch_;
  pdims_ = dims_;
  pdims_[0] = 64;
  const int L = n_layers();
  SplitMix rng{seed};
  for (int l = 0; l < L; ++l) {
    const int din = pdims_[l], dout = pdims_[l + 1], dreal = dims_[l];
    const size_t nw = static_cast<size_t>(dout) * din;
    // PyTorch's default: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for W and b.
    const double bound = 1.0 / std::sqrt(static_cast<double>(dreal));
    std::vector<float> w(nw, 0.0f), b(dout);
    for (int o = 0; o < dout; ++o)
      for (int i = 0; i < dreal; ++i) w[static_cast<size_t>(o) * din + i] = static_cast<float>((2.0 * rng.uniform() - 1.0) * bound);
    for (int o = 0; o < dout; ++o) b[o] = static_cast<float>((2.0 * rng.uniform() - 1.0) * bound);
    W_.push_back(alloc<float>(nw));
    Wb_.push_back(alloc<uint16_t>(nw));
    Wbt_.push_back(alloc<uint16_t>(nw));
    gW_.push_back(alloc<float>(nw));
    mW_.push_back(alloc<float>(nw));
    vW_.push_back(alloc<float>(nw));
    b_.push_back(alloc<float>(dout));
    gb_.push_back(alloc<float>(dout));
    mb_.push_back(alloc<float>(dout));
    vb_.push_back(alloc<float>(dout));
    copy(W_[l], w.data(), nw * sizeof(float), hipMemcpyHostToDevice);
    copy(b_[l], b.data(), dout * sizeof(float), hipMemcpyHostToDevice);
    check(hipMemset(mW_[l], 0, nw * sizeof(float)), "hipMemset");
    check(hipMemset(vW_[l], 0, nw * sizeof(float)), "hipMemset");
    check(hipMemset(mb_[l], 0, dout * sizeof(float)), "hipMemset");
    check(hipMemset(vb_[l], 0, dout * sizeof(float)), "hipMemset");
    check(hipMemset(gW_[l], 0, nw * sizeof(float)), "hipMemset");
    check(hipMemset(gb_[l], 0, dout * sizeof(float)), "hipMemset");
    check(mlp::launch_pack(W_[l], Wb_[l], nw, ctx_.stream()), "pack");
    check(mlp::launch_transpose(Wb_[l], dout, din, Wbt_[l], ctx_.stream()), "transpose");
  }
  {
    const size_t nw = static_cast<size_t>(pdims_[1]) * 64;
    std::vector<uint8_t> m(nw, 0);
    for (int o = 0; o < pdims_[1]; ++o)
      for (int i = 0; i < dims_[0]; ++i) m[static_cast<size_t>(o) * 64 + i] = 1;
    mask0_ = alloc<uint8_t>(nw);
    copy(mask0_, m.data(), nw, hipMemcpyHostToDevice);
  }
  for (int l = 0; l <= L; ++l) A_.push_back(alloc<uint16_t>(static_cast<size_t>(max_batch) * pdims_[l]));
  for (int l = 0; l < L; ++l) {
    Z_.push_back(alloc<float>(static_cast<size_t>(max_batch) * pdims_[l + 1]));
    dZ_.push_back(alloc<uint16_t>(static_cast<size_t>(max_batch) * pdims_[l + 1]));
  }
  yb_ = alloc<float>(static_cast<size_t>(max_batch) * out_dim());
  xin_ = alloc<float>(static_cast<size_t>(max_batch) * in_dim());
  int wmax = 0;
  for (int d : pdims_) wmax = std::max(wmax, d);
  dZt_ = alloc<uint16_t>(static_cast<size_t>(max_batch) * wmax);
  splitk_ = alloc<float>(static_cast<size_t>(kMaxSplit) * wmax * wmax);
  At_ = alloc<uint16_t>(static_cast<size_t>(max_batch) * wmax);
  if (2 * L > mlp::ADAM_MAX_TENSORS) throw std::invalid_argument("too many layers");
  idx_ = alloc<int>(max_batch);
  ident_ = alloc<int>(max_batch);
  {
    std::vector<int> ident(max_batch);
    for (int i = 0; i < max_batch; ++i) ident[i] = i;
    copy(ident_, ident.data(), max_batch * sizeof(int), hipMemcpyHostToDevice);
  }
  partial_ = alloc<double>(loss_blocks_);
  for (hipEvent_t& e : ev_) check(hipEventCreate(&e), "hipEventCreate");
  check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
}

Mlp::~Mlp() { release(); }

void Mlp::release() {
  (void)hipSetDevice(ctx_.device());
  // A device-wide wait before the frees: after a stream wait alone the
  // runtime can keep freed memory allocated (lch::Batch::release_all).
  (void)hipDeviceSynchronize();
  for (hipEvent_t& e : ev_) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (X_) (void)hipFree(X_);
  if (Y_) (void)hipFree(Y_);
  X_ = nullptr;
  Y_ = nullptr;
  for (void* p : allocs_) (void)hipFree(p);
  allocs_.clear();
}

size_t Mlp::param_count() const {
  size_t n = 0;
  for (int l = 0; l < n_layers(); ++l) n += static_cast<size_t>(dims_[l + 1]) * dims_[l] + dims_[l + 1];
  return n;
}

void Mlp::get_params(float* out) const {
  check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
  size_t o = 0;
  for (int l = 0; l < n_layers(); ++l) {
    const int din = pdims_[l], dout = pdims_[l + 1], dreal = dims_[l];
    std::vector<float> w(static_cast<size_t>(dout) * din);
    copy(w.data(), W_[l], w.size() * sizeof(float), hipMemcpyDeviceToHost);
    for (int r = 0; r < dout; ++r)
      for (int i = 0; i < dreal; ++i) out[o++] = w[static_cast<size_t>(r) * din + i];
    copy(out + o, b_[l], dout * sizeof(float), hipMemcpyDeviceToHost);
    o += dout;
  }
}

void Mlp::set_params(const float* in) {
  check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
  size_t o = 0;
  for (int l = 0; l < n_layers(); ++l) {
    const int din = pdims_[l], dout = pdims_[l + 1], dreal = dims_[l];
    std::vector<float> w(static_cast<size_t>(dout) * din, 0.0f);
    for (int r = 0; r < dout; ++r)
      for (int i = 0; i < dreal; ++i) w[static_cast<size_t>(r) * din + i] = in[o++];
    copy(W_[l], w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice);
    copy(b_[l], in + o, dout * sizeof(float), hipMemcpyHostToDevice);
    o += dout;
    check(mlp::launch_pack(W_[l], Wb_[l], w.size(), ctx_.stream()), "pack");
    check(mlp::launch_transpose(Wb_[l], dout, din, Wbt_[l], ctx_.stream()), "transpose");
  }
  check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
}

void Mlp::set_data(const float* x, const float* y, int n) {
  if (n <= 0) throw std::invalid_argument("empty training set");
  check(hipDeviceSynchronize(), "hipDeviceSynchronize");  // before the frees, as in release()
  if (X_) (void)hipFree(X_);
  if (Y_) (void)hipFree(Y_);
  X_ = Y_ = nullptr;
  n_data_ = 0;
  check(hipMalloc(reinterpret_cast<void**>(&X_), static_cast<size_t>(n) * in_dim() * sizeof(float)), "hipMalloc");
  check(hipMalloc(reinterpret_cast<void**>(&Y_), static_cast<size_t>(n) * out_dim() * sizeof(float)), "hipMalloc");
  copy(X_, x, static_cast<size_t>(n) * in_dim() * sizeof(float), hipMemcpyHostToDevice);
  copy(Y_, y, static_cast<size_t>(n) * out_dim() * sizeof(float), hipMemcpyHostToDevice);
  n_data_ = n;
}

void Mlp::gemm(const mlp::GemmArgs& a, bool a_kc, bool b_kc, int epi) {
  check(mlp::launch_gemm(a, a_kc, b_kc, epi, ctx_.stream()), "gemm launch");
}

// A_[0] holds the batch's BF16 inputs; the targets are in yb_.
void Mlp::forward(int batch) {
  const int L = n_layers();
  for (int l = 0; l < L; ++l) {
    mlp::GemmArgs a;
    a.A = A_[l];
    a.B = Wb_[l];
    a.M = batch;
    a.N = pdims_[l + 1];
    a.K = pdims_[l];
    a.lda = pdims_[l];
    a.ldb = pdims_[l];
    a.ldc = pdims_[l + 1];
    a.bias = b_[l];
    a.Cf = Z_[l];
    a.Cb = A_[l + 1];
    gemm(a, true, true, l + 1 < L ? mlp::EPI_BIAS_GELU : mlp::EPI_BIAS);
  }
}

// dZ_[L-1] holds dLoss/dY (BF16). Every GEMM reads K-contiguous operands:
// the weight gradient dZ^T . A uses transposed copies of dZ and A, the
// input gradient dZ . W the transposed BF16 weights Adam keeps.
void Mlp::backward(int batch) {
  const int L = n_layers();
  for (int l = L - 1; l >= 0; --l) {
    const int din = pdims_[l], dout = pdims_[l + 1];
    if (!gemm_only_) {
      check(mlp::launch_transpose(dZ_[l], batch, dout, dZt_, ctx_.stream()), "transpose");
      check(mlp::launch_rowsum(dZt_, dout, batch, gb_[l], ctx_.stream()), "bias grad");
      check(mlp::launch_transpose(A_[l], batch, din, At_, ctx_.stream()), "transpose");
    }
    // gW[out][in] = sum over rows of dZ[row][out] A[row][in]. M x N is
    // small and K = batch long: K is split so that the launch fills the
    // chip, the slices' FP32 partial products are summed in fixed order.
    mlp::GemmArgs g;
    g.A = dZt_;
    g.B = At_;
    g.M = dout;
    g.N = din;
    g.K = batch;
    g.lda = batch;
    g.ldb = batch;
    g.ldc = din;
    int ks = 1;
    while (ks < kMaxSplit && (dout / 64) * (din / 64) * ks < 512 && (batch / 64) % (2 * ks) == 0) ks *= 2;
    g.ksplit = ks;
    g.Cf = ks > 1 ? splitk_ : gW_[l];
    gemm(g, true, true, mlp::EPI_PLAIN);
    if (ks > 1)
      check(mlp::launch_splitk_sum(splitk_, ks, static_cast<size_t>(dout) * din, gW_[l], ctx_.stream()), "split-K sum");
    if (l > 0) {
      // dZ_{l-1} = (dZ_l . W_l) * gelu'(Z_{l-1}); B (k = out, n = in) = Wt[in][out].
      mlp::GemmArgs d;
      d.A = dZ_[l];
      d.B = Wbt_[l];
      d.M = batch;
      d.N = din;
      d.K = dout;
      d.lda = dout;
      d.ldb = dout;
      d.ldc = din;
      d.Z = Z_[l - 1];
      d.Cb = dZ_[l - 1];
      gemm(d, true, true, mlp::EPI_GELU_GRAD);
    }
  }
}

void Mlp::adam_step(float lr) {
  ++t_;
  mlp::AdamParams p;
  p.lr = lr;
  p.beta1 = 0.9f;
  p.beta2 = 0.999f;
  p.eps = 1e-8f;
  p.bc1 = static_cast<float>(1.0 - std::pow(0.9, t_));
  p.bc2 = static_cast<float>(1.0 - std::pow(0.999, t_));
  mlp::AdamTable tab{};
  size_t off = 0;
  int k = 0;
  for (int l = 0; l < n_layers(); ++l) {
    const int din = pdims_[l], dout = pdims_[l + 1];
    tab.w[k] = W_[l];
    tab.g[k] = gW_[l];
    tab.m[k] = mW_[l];
    tab.v[k] = vW_[l];
    tab.wb[k] = Wb_[l];
    tab.wbt[k] = Wbt_[l];
    tab.mask[k] = l == 0 ? mask0_ : nullptr;
    tab.rows[k] = dout;
    tab.cols[k] = din;
    tab.start[k] = off;
    off += static_cast<size_t>(dout) * din;
    ++k;
    tab.w[k] = b_[l];
    tab.g[k] = gb_[l];
    tab.m[k] = mb_[l];
    tab.v[k] = vb_[l];
    tab.wb[k] = nullptr;
    tab.wbt[k] = nullptr;
    tab.mask[k] = nullptr;
    tab.rows[k] = 1;
    tab.cols[k] = dout;
    tab.start[k] = off;
    off += dout;
    ++k;
  }
  tab.start[k] = off;
  tab.count = k;
  check(mlp::launch_adam_multi(tab, p, ctx_.stream()), "adam");
}

void Mlp::run_step(int batch, float lr, bool adam, bool gemm_only) {
  gemm_only_ = gemm_only;
  forward(batch);
  if (!gemm_only) {
    const size_t count = static_cast<size_t>(batch) * out_dim();
    check(mlp::launch_loss(Z_.back(), yb_, dZ_.back(), count, partial_, loss_blocks_, ctx_.stream()), "loss");
  }
  backward(batch);
  gemm_only_ = false;
  if (adam) adam_step(lr);
}

double Mlp::train_epoch(const int* perm, int n, int batch, float lr) {
  if (!X_) throw std::invalid_argument("no training data: call set_data first");
  if (batch <= 0 || batch % 64 || batch > max_batch_) throw std::invalid_argument("bad batch size");
  for (int i = 0; i < n; ++i)
    if (perm[i] < 0 || perm[i] >= n_data_) throw std::invalid_argument("permutation index out of range");
  const int steps = n / batch;
  double loss = 0.0;
  std::vector<double> part(loss_blocks_);
  for (int s = 0; s < steps; ++s) {
    check(hipMemcpyAsync(idx_, perm + static_cast<size_t>(s) * batch, batch * sizeof(int), hipMemcpyHostToDevice,
                         ctx_.stream()), "hipMemcpyAsync");
    check(mlp::launch_gather(X_, Y_, idx_, batch, in_dim(), 64, out_dim(), A_[0], yb_, ctx_.stream()), "gather");
    run_step(batch, lr, true, false);
    check(hipMemcpyAsync(part.data(), partial_, loss_blocks_ * sizeof(double), hipMemcpyDeviceToHost, ctx_.stream()),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
    double sum = 0.0;
    for (double v : part) sum += v;
    loss += sum / (static_cast<double>(batch) * out_dim());
  }
  return steps ? loss / steps : 0.0;
}

double Mlp::loss_and_grad(const float* x, const float* y, int n, float* grads) {
  if (n <= 0 || n % 64 || n > max_batch_) throw std::invalid_argument("rows must be a positive multiple of 64 <= max_batch");
  copy(xin_, x, static_cast<size_t>(n) * in_dim() * sizeof(float), hipMemcpyHostToDevice);
  copy(yb_, y, static_cast<size_t>(n) * out_dim() * sizeof(float), hipMemcpyHostToDevice);
  std::vector<int> ident(n);
  for (int i = 0; i < n; ++i) ident[i] = i;
  copy(idx_, ident.data(), n * sizeof(int), hipMemcpyHostToDevice);
  check(mlp::launch_gather(xin_, nullptr, idx_, n, in_dim(), 64, out_dim(), A_[0], nullptr, ctx_.stream()), "gather");
  run_step(n, 0.0f, false, false);
  std::vector<double> part(loss_blocks_);
  copy(part.data(), partial_, loss_blocks_ * sizeof(double), hipMemcpyDeviceToHost);
  double sum = 0.0;
  for (double v : part) sum += v;
  if (grads) {
    size_t o = 0;
    for (int l = 0; l < n_layers(); ++l) {
      const int din = pdims_[l], dout = pdims_[l + 1], dreal = dims_[l];
      std::vector<float> g(static_cast<size_t>(dout) * din);
      copy(g.data(), gW_[l], g.size() * sizeof(float), hipMemcpyDeviceToHost);
      for (int r = 0; r < dout; ++r)
        for (int i = 0; i < dreal; ++i) grads[o++] = g[static_cast<size_t>(r) * din + i];
      copy(grads + o, gb_[l], dout * sizeof(float), hipMemcpyDeviceToHost);
      o += dout;
    }
  }
  return sum / (static_cast<double>(n) * out_dim());
}

void Mlp::forward_device(const float* x, int rows) {
  if (rows <= 0 || rows % 64 || rows > max_batch_) throw std::invalid_argument("rows must be a positive multiple of 64 <= max_batch");
  check(mlp::launch_gather(x, nullptr, ident_, rows, in_dim(), 64, out_dim(), A_[0], nullptr, ctx_.stream()),
        "gather");
  forward(rows);
}

void Mlp::predict(const float* x, int n, float* y) {
  if (n < 0) throw std::invalid_argument("negative row count");
  const int ch = max_batch_;
  for (int r0 = 0; r0 < n; r0 += ch) {
    const int rows = std::min(ch, n - r0);
    const int padded = (rows + 63) / 64 * 64;
    std::vector<float> xs(static_cast<size_t>(padded) * in_dim(), 0.0f);
    std::memcpy(xs.data(), x + static_cast<size_t>(r0) * in_dim(), static_cast<size_t>(rows) * in_dim() * sizeof(float));
    copy(xin_, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice);
    forward_device(xin_, padded);
    copy(y + static_cast<size_t>(r0) * out_dim(), Z_.back(),
                    static_cast<size_t>(rows) * out_dim() * sizeof(float), hipMemcpyDeviceToHost);
  }
}

void Mlp::time_step(int batch, int reps, double* ms_step, double* ms_gemm, double* gemm_flop) {
  if (!X_) throw std::invalid_argument("no training data: call set_data first");
  if (batch <= 0 || batch % 64 || batch > max_batch_ || batch > n_data_) throw std::invalid_argument("bad batch size");
  std::vector<int> ident(batch);
  for (int i = 0; i < batch; ++i) ident[i] = i;
  copy(idx_, ident.data(), batch * sizeof(int), hipMemcpyHostToDevice);
  check(mlp::launch_gather(X_, Y_, idx_, batch, in_dim(), 64, out_dim(), A_[0], yb_, ctx_.stream()), "gather");
  hipStream_t st = ctx_.stream();
  run_step(batch, 1e-4f, true, false);  // warm-up
  check(hipEventRecord(ev_[0], st), "hipEventRecord");
  for (int r = 0; r < reps; ++r) run_step(batch, 1e-4f, true, false);
  check(hipEventRecord(ev_[1], st), "hipEventRecord");
  for (int r = 0; r < reps; ++r) run_step(batch, 0.0f, false, true);
  check(hipEventRecord(ev_[2], st), "hipEventRecord");
  check(hipEventSynchronize(ev_[2]), "hipEventSynchronize");
  float a = 0, b = 0;
  check(hipEventElapsedTime(&a, ev_[0], ev_[1]), "hipEventElapsedTime");
  check(hipEventElapsedTime(&b, ev_[1], ev_[2]), "hipEventElapsedTime");
  *ms_step = a / reps;
  *ms_gemm = b / reps;
  double f = 0.0;
  for (int l = 0; l < n_layers(); ++l) {
    const double mnk = static_cast<double>(batch) * pdims_[l] * pdims_[l + 1];
    f += 2.0 * mnk * (l > 0 ? 3.0 : 2.0);  // forward, weight gradient, input gradient (not for layer 0)
  }
  *gemm_flop = f;
}

}  // namespace lch
