// HIP kernels of the light-curve surrogate MLP for CDNA4 (gfx950).
//
// gemm_kernel — C[M x N] = A[M x K] . B[K x N] on BF16 MFMA
//   (v_mfma_f32_32x32x16_bf16) with FP32 accumulation. A workgroup of four
//   wavefronts computes a 64 x 64 tile, each wave a 32 x 32 quarter; K is
//   staged through two LDS buffers in steps of 64 with both operands
//   stored K-contiguous there, so every lane's MFMA operand is one 16-byte
//   LDS read, and the global loads of the next two steps are in flight (in
//   registers) while the current step's MFMAs run. Operands
//   may arrive K-contiguous or transposed in global memory (the three GEMMs
//   of a layer's forward and backward passes need both), and the epilogue is
//   fused: bias + GELU for hidden layers, bias only for the output layer,
//   GELU' for the backward pass, plain FP32 for weight gradients.
// loss_kernel / rowsum_kernel / transpose_kernel / adam_multi_kernel /
//   pack_kernel — the MSE loss and its gradient, bias gradients (row sums of
//   the transposed output gradient), BF16 transposes, one Adam launch over
//   all parameter tensors on FP32 master weights writing the BF16 copies,
//   and FP32 -> BF16 packing.
//
// Shapes are checked on the host: M, N and K multiples of 64, leading
// dimensions multiples of 8 elements.
#include <hip/hip_runtime.h>

#include "mlp_kernels.h"

namespace lch {
namespace mlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// K step BK (64 as launched). The LDS row stride is BK + 8 bf16; the pad
// breaks bank conflicts.
constexpr int BM = 64, BN = 64, BK_MIN = 64;
template <int BK>
constexpr int lds_stride() { return BK + 8; }

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// GELU with the error function, and its derivative.
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// One BM x BK (or BN x BK) operand tile: each of the 256 threads fetches
// BK / 32 16-byte pieces into registers (fetch), later written to LDS as
// [row][k] (put). KC: global element (row, k) at g[row * ld + k]; else at
// g[k * ld + row].
template <int BK>
struct Piece {
  uint4 v[BK / 32];
};

template <bool KC, int BK>
__device__ __forceinline__ Piece<BK> fetch(const uint16_t* __restrict__ g, int ld, int row0, int k0) {
  const int t = threadIdx.x;
  Piece<BK> p;
#pragma unroll
  for (int h = 0; h < BK / 32; ++h) {
    if (KC) {
      const int r = t >> 2, kk = (t & 3) * 8 + h * 32;
      p.v[h] = *reinterpret_cast<const uint4*>(g + static_cast<size_t>(row0 + r) * ld + k0 + kk);
    } else {
      const int kk = (t >> 3) + h * 32, r = (t & 7) * 8;
      p.v[h] = *reinterpret_cast<const uint4*>(g + static_cast<size_t>(k0 + kk) * ld + row0 + r);
    }
  }
  return p;
}

template <bool KC, int BK>
__device__ __forceinline__ void put(uint16_t* __restrict__ s, const Piece<BK>& p) {
  constexpr int LS = lds_stride<BK>();
  const int t = threadIdx.x;
#pragma unroll
  for (int h = 0; h < BK / 32; ++h) {
    if (KC) {
      const int r = t >> 2, kk = (t & 3) * 8 + h * 32;
      *reinterpret_cast<uint4*>(s + r * LS + kk) = p.v[h];
    } else {
      const int kk = (t >> 3) + h * 32, r = (t & 7) * 8;
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&p.v[h]);
#pragma unroll
      for (int i = 0; i < 8; ++i) s[(r + i) * LS + kk] = e[i];
    }
  }
}

}  // namespace

// Workgroups are dispatched to the eight XCDs (each with its own L2) in
// turn: workgroup b runs on XCD b % 8. The tile of workgroup b is numbered
// so that XCD x takes a contiguous run of the tiles in (slice, row panel,
// column) order, i.e. whole row panels of A (and whole K slices), read
// into one L2 instead of eight.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  const uint32_t q = n / 8u, r = n % 8u;
  const uint32_t x = b % 8u, i = b / 8u;
  return x * q + (x < r ? x : r) + i;
}

template <bool AKC, bool BKC, int EPI, int BK>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs a) {
  constexpr int LDS_STRIDE = lds_stride<BK>();
  __shared__ __attribute__((aligned(16))) uint16_t As[2][BM * LDS_STRIDE];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN * LDS_STRIDE];
  const uint32_t nx = gridDim.x, ny = gridDim.y;
  const uint32_t tile =
      xcd_tile((blockIdx.z * ny + blockIdx.y) * nx + blockIdx.x, nx * ny * gridDim.z);
  const int bx = static_cast<int>(tile % nx), by = static_cast<int>((tile / nx) % ny),
            bz = static_cast<int>(tile / (nx * ny));
  const int m0 = by * BM, n0 = bx * BN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  // Split K: slice bz covers [kb, kb + K / gridDim.z).
  const int kspan = a.K / static_cast<int>(gridDim.z);
  const int kb = bz * kspan;
  f32x16 acc = {};
  const int nk = kspan / BK;
  // Operand loads run two K steps ahead: two register sets alternate (the
  // loop is unrolled by two so that no register set is copied while its
  // loads are in flight); each step's set is written to the LDS buffer the
  // step before last finished reading.
  auto mma = [&](int buf) {
#pragma unroll
    for (int s = 0; s < BK / 16; ++s) {
      const int kk = s * 16 + (l >> 5) * 8;
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(As[buf] + (wm + (l & 31)) * LDS_STRIDE + kk);
      const bf16x8 bv = *reinterpret_cast<const bf16x8*>(Bs[buf] + (wn + (l & 31)) * LDS_STRIDE + kk);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
  };
  {
    const Piece<BK> pa = fetch<AKC, BK>(a.A, a.lda, m0, kb);
    const Piece<BK> pb = fetch<BKC, BK>(a.B, a.ldb, n0, kb);
    put<AKC, BK>(As[0], pa);
    put<BKC, BK>(Bs[0], pb);
  }
  Piece<BK> a1, b1, a2, b2;  // the sets of steps kt + 1 and kt + 2
  if (nk > 1) {
    a1 = fetch<AKC, BK>(a.A, a.lda, m0, kb + BK);
    b1 = fetch<BKC, BK>(a.B, a.ldb, n0, kb + BK);
  }
  if (nk > 2) {
    a2 = fetch<AKC, BK>(a.A, a.lda, m0, kb + 2 * BK);
    b2 = fetch<BKC, BK>(a.B, a.ldb, n0, kb + 2 * BK);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt += 2) {
    // Step kt from buffer 0; set 1 holds step kt + 1, set 2 step kt + 2.
    mma(0);
    if (kt + 1 < nk) {
      // Buffer 1 was last read in step kt - 1, before the barrier that ended it.
      put<AKC, BK>(As[1], a1);
      put<BKC, BK>(Bs[1], b1);
    }
    if (kt + 3 < nk) {
      a1 = fetch<AKC, BK>(a.A, a.lda, m0, kb + (kt + 3) * BK);
      b1 = fetch<BKC, BK>(a.B, a.ldb, n0, kb + (kt + 3) * BK);
    }
    __syncthreads();
    if (kt + 1 >= nk) break;
    // Step kt + 1 from buffer 1; set 2 holds step kt + 2, set 1 step kt + 3.
    mma(1);
    if (kt + 2 < nk) {
      put<AKC, BK>(As[0], a2);
      put<BKC, BK>(Bs[0], b2);
    }
    if (kt + 4 < nk) {
      a2 = fetch<AKC, BK>(a.A, a.lda, m0, kb + (kt + 4) * BK);
      b2 = fetch<BKC, BK>(a.B, a.ldb, n0, kb + (kt + 4) * BK);
    }
    __syncthreads();
  }
  // Lane l holds C[(r / 4) * 8 + (l / 32) * 4 + r % 4][l % 32] of the wave's
  // 32 x 32 quarter in acc[r].
  const int n = n0 + wn + (l & 31);
  // Split-K slices write FP32 partial products to consecutive M x ldc planes.
  const size_t plane = static_cast<size_t>(bz) * a.M * a.ldc;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + wm + (r >> 2) * 8 + (l >> 5) * 4 + (r & 3);
    const size_t idx = static_cast<size_t>(m) * a.ldc + n + (EPI == EPI_PLAIN ? plane : 0);
    const float v = acc[r];
    if (EPI == EPI_BIAS_GELU) {
      const float z = v + a.bias[n];
      a.Cf[idx] = z;
      a.Cb[idx] = f2bf(gelu(z));
    } else if (EPI == EPI_BIAS) {
      a.Cf[idx] = v + a.bias[n];
    } else if (EPI == EPI_GELU_GRAD) {
      a.Cb[idx] = f2bf(v * gelu_grad(a.Z[idx]));
    } else {
      a.Cf[idx] = v;
    }
  }
}

// MSE over the batch and outputs: loss = mean((y - t)^2); dy = 2 (y - t) / count
// in BF16 for the backward GEMMs. Per-block partial sums in partial[].
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                                   uint16_t* __restrict__ dy, size_t count,
                                                   double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0.0;
  const float g = 2.0f / static_cast<float>(count);
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += 256ull * gridDim.x) {
    const float d = y[i] - t[i];
    s += static_cast<double>(d) * d;
    if (dy) dy[i] = f2bf(g * d);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// BF16 transpose out[c][r] = in[r][c] of a rows x cols matrix (both
// multiples of 64) through a 64 x 64 LDS tile: coalesced 16-byte loads and
// stores, the transpose done in LDS.
__global__ __launch_bounds__(256) void transpose_kernel(const uint16_t* __restrict__ in, int rows, int cols,
                                                        uint16_t* __restrict__ out) {
  __shared__ uint16_t tile[64][66];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int t = threadIdx.x;
  {
    const int r = t >> 2, c = (t & 3) * 16;
    const uint4* src = reinterpret_cast<const uint4*>(in + static_cast<size_t>(r0 + r) * cols + c0 + c);
    const uint4 a = src[0], b = src[1];
    const uint16_t* ha = reinterpret_cast<const uint16_t*>(&a);
    const uint16_t* hb = reinterpret_cast<const uint16_t*>(&b);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      tile[r][c + i] = ha[i];
      tile[r][c + 8 + i] = hb[i];
    }
  }
  __syncthreads();
  {
    const int oc = t >> 2, orr = (t & 3) * 16;  // output row = input column oc
    uint4 a, b;
    uint16_t* ha = reinterpret_cast<uint16_t*>(&a);
    uint16_t* hb = reinterpret_cast<uint16_t*>(&b);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      ha[i] = tile[orr + i][oc];
      hb[i] = tile[orr + 8 + i][oc];
    }
    uint4* dst = reinterpret_cast<uint4*>(out + static_cast<size_t>(c0 + oc) * rows + r0 + orr);
    dst[0] = a;
    dst[1] = b;
  }
}

// Row sums of a BF16 matrix (FP32 out; cols a multiple of 8), one
// wavefront per row with 16-byte loads: the bias gradient from the
// transposed output gradient.
__global__ __launch_bounds__(256) void rowsum_kernel(const uint16_t* __restrict__ x, int rows, int cols,
                                                     float* __restrict__ out) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (row >= rows) return;
  const uint4* p = reinterpret_cast<const uint4*>(x + static_cast<size_t>(row) * cols);
  float s = 0.0f;
  for (int c = l; c < cols / 8; c += 64) {
    const uint4 v = p[c];
    const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int i = 0; i < 8; ++i) s += bf2f(h[i]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (l == 0) out[row] = s;
}

// Adam (Kingma & Ba 2015) over every parameter tensor in one launch: FP32
// master weights, BF16 copy (and its transpose for weight matrices).
// Entries of a tensor's mask that are 0 stay frozen (padding columns of the
// first layer).
__global__ __launch_bounds__(256) void adam_multi_kernel(AdamTable tab, AdamParams p) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= tab.start[tab.count]) return;
  int k = 0;
  while (i >= tab.start[k + 1]) ++k;
  const size_t j = i - tab.start[k];
  if (tab.mask[k] && !tab.mask[k][j]) return;
  float* w = tab.w[k];
  float* m = tab.m[k];
  float* v = tab.v[k];
  const float gi = tab.g[k][j];
  const float mi = p.beta1 * m[j] + (1.0f - p.beta1) * gi;
  const float vi = p.beta2 * v[j] + (1.0f - p.beta2) * gi * gi;
  m[j] = mi;
  v[j] = vi;
  const float wi = w[j] - p.lr * (mi / p.bc1) / (sqrtf(vi / p.bc2) + p.eps);
  w[j] = wi;
  if (tab.wb[k]) {
    const uint16_t h = f2bf(wi);
    tab.wb[k][j] = h;
    const int cols = tab.cols[k], rows = tab.rows[k];
    tab.wbt[k][static_cast<size_t>(j % cols) * rows + j / cols] = h;
  }
}

// Minibatch gather: xb[b][0..din_pad) = bf16(X[idx[b]][0..din)), zero
// padded; yb[b][:] = Y[idx[b]][:].
__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ X, const float* __restrict__ Y,
                                                     const int* __restrict__ idx, int batch, int din,
                                                     int din_pad, int dout, uint16_t* __restrict__ xb,
                                                     float* __restrict__ yb) {
  const int b = blockIdx.x;
  if (b >= batch) return;
  const size_t r = static_cast<size_t>(idx[b]);
  for (int j = threadIdx.x; j < din_pad; j += 256)
    xb[static_cast<size_t>(b) * din_pad + j] = j < din ? f2bf(X[r * din + j]) : static_cast<uint16_t>(0);
  if (Y)
    for (int j = threadIdx.x; j < dout; j += 256) yb[static_cast<size_t>(b) * dout + j] = Y[r * dout + j];
}

__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ x, uint16_t* __restrict__ y,
                                                   size_t n) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// ------------------------------------------------------------ launchers
template <bool AKC, bool BKC, int EPI>
static hipError_t launch_gemm_t(const GemmArgs& a, hipStream_t st) {
  // K steps of 64: 37 KB of LDS, four workgroups per CU (steps of 128 halve
  // that and measured slower at every batch).
  hipLaunchKernelGGL((gemm_kernel<AKC, BKC, EPI, 64>), dim3(a.N / BN, a.M / BM, a.ksplit), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_gemm(const GemmArgs& a, bool a_kc, bool b_kc, int epi, hipStream_t st) {
  if (a.M % BM || a.N % BN || a.K % BK_MIN || a.M <= 0 || a.N <= 0 || a.K <= 0) return hipErrorInvalidValue;
  if (a.ksplit < 1 || (a.K / BK_MIN) % a.ksplit || (a.ksplit > 1 && epi != EPI_PLAIN)) return hipErrorInvalidValue;
  if (a.lda % 8 || a.ldb % 8) return hipErrorInvalidValue;
#define LCH_GEMM_CASE(AK, BK_, E) \
  if (a_kc == AK && b_kc == BK_ && epi == E) return launch_gemm_t<AK, BK_, E>(a, st);
  // The MLP passes K-contiguous operands only (transposed copies are made
  // by transpose_kernel); the strided staging path serves other callers.
  LCH_GEMM_CASE(true, true, EPI_BIAS_GELU)
  LCH_GEMM_CASE(true, true, EPI_BIAS)
  LCH_GEMM_CASE(true, true, EPI_PLAIN)
  LCH_GEMM_CASE(true, true, EPI_GELU_GRAD)
  LCH_GEMM_CASE(false, false, EPI_PLAIN)
#undef LCH_GEMM_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_loss(const float* y, const float* t, uint16_t* dy, size_t count, double* partial,
                       int blocks, hipStream_t st) {
  hipLaunchKernelGGL(loss_kernel, dim3(blocks), dim3(256), 0, st, y, t, dy, count, partial);
  return hipGetLastError();
}

hipError_t launch_gather(const float* X, const float* Y, const int* idx, int batch, int din, int din_pad,
                         int dout, uint16_t* xb, float* yb, hipStream_t st) {
  hipLaunchKernelGGL(gather_kernel, dim3(batch), dim3(256), 0, st, X, Y, idx, batch, din, din_pad, dout, xb, yb);
  return hipGetLastError();
}

hipError_t launch_transpose(const uint16_t* in, int rows, int cols, uint16_t* out, hipStream_t st) {
  if (rows % 64 || cols % 64 || rows <= 0 || cols <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(transpose_kernel, dim3(cols / 64, rows / 64), dim3(256), 0, st, in, rows, cols, out);
  return hipGetLastError();
}

// out[i] = sum over the nsplit planes of part[plane * n + i].
__global__ __launch_bounds__(256) void splitk_sum_kernel(const float* __restrict__ part, int nsplit, size_t n,
                                                         float* __restrict__ out) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int p = 0; p < nsplit; ++p) s += part[static_cast<size_t>(p) * n + i];
  out[i] = s;
}

hipError_t launch_splitk_sum(const float* part, int nsplit, size_t n, float* out, hipStream_t st) {
  hipLaunchKernelGGL(splitk_sum_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, part,
                     nsplit, n, out);
  return hipGetLastError();
}

hipError_t launch_rowsum(const uint16_t* x, int rows, int cols, float* out, hipStream_t st) {
  if (cols % 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rowsum_kernel, dim3((rows + 3) / 4), dim3(256), 0, st, x, rows, cols, out);
  return hipGetLastError();
}

hipError_t launch_adam_multi(const AdamTable& t, const AdamParams& p, hipStream_t st) {
  if (t.count <= 0 || t.count > ADAM_MAX_TENSORS) return hipErrorInvalidValue;
  const size_t n = t.start[t.count];
  hipLaunchKernelGGL(adam_multi_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, t, p);
  return hipGetLastError();
}

hipError_t launch_pack(const float* x, uint16_t* y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(pack_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, x, y, n);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace lch
