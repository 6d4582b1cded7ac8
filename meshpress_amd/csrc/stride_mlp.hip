// Learned coder/predictor proposal: a 32 -> 128 -> 128 -> 176 MLP with
// BF16 weights and activations and FP32 accumulation on the gfx950 matrix
// cores (v_mfma_f32_16x16x32_bf16), plus its backward pass, a grouped
// softmax cross-entropy and an Adam step.
//
// Parameter layout (float32 master copy, n_params = 43,440):
//   W1 32 x 128 | W2 128 x 128 | W3 128 x 176 | b1 128 | b2 128 | b3 176
// (W row-major, input index major).  The first 43,008 entries (the
// weights) also exist as a bf16 copy that the GEMMs read.
//
// Output heads (logit columns): three predictor heads of 17 classes (one
// per axis, columns 0..50) and six coder heads of 20 classes (one per
// residual stream dx dy dz px py pz, columns 51..170); columns 171..175 pad
// the width to a multiple of 16 and never enter a softmax.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/stride_hip.h"
#include "stride_err.h"

namespace stride {
namespace mlp {

constexpr int kIn = 32, kHid = 128, kOut = 176;
constexpr int kHeads = 9;
constexpr int kNW = kIn * kHid + kHid * kHid + kHid * kOut;   // 43,008
constexpr int kNP = kNW + kHid + kHid + kOut;                 // 43,440
__constant__ int kHeadOff[kHeads + 1] = {0, 17, 34, 51, 71, 91, 111, 131, 151, 171};

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN stays a quiet NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum Epi : int { kEpiF32 = 0, kEpiBiasReluBf16 = 1, kEpiMaskBf16 = 2 };

struct GemmArgs {
  const uint16_t* A;   // op(A)[m][k] = transA ? A[k * lda + m] : A[m * lda + k]
  const uint16_t* B;   // op(B)[k][n] = transB ? B[n * ldb + k] : B[k * ldb + n]
  int lda, ldb, transA, transB;
  int M, N, K;
  int epi;
  const float* bias;     // N floats or null (kEpiF32, kEpiBiasReluBf16)
  const uint16_t* mask;  // kEpiMaskBf16: keep C[m][n] where mask[m * ldm + n] > 0
  int ldm;
  float* Cf;             // kEpiF32 output
  uint16_t* Cb;          // bf16 outputs
  int ldc;
};

// C = op(A) op(B): a 256-thread workgroup owns a 64 x 64 tile of C; wave w
// owns rows 16w..16w+15 as four 16 x 16 MFMA tiles.  Each K step of 32
// stages the 64 x 32 slice of op(A) and the 32 x 64 slice of op(B) in LDS
// (zero outside the matrices), then every lane feeds its fragments of the
// gfx950 16 x 16 x 32 bf16 MFMA, eight k values each:
//   A: row = lane % 16, k = 8 (lane / 16) + i
//   B: col = lane % 16, k = 8 (lane / 16) + i
//   D: row = 4 (lane / 16) + i, col = lane % 16
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  __shared__ uint16_t As[64][32 + 4];
  __shared__ uint16_t Bs[32][64 + 4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  f4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    // stage: 2048 elements of each operand, 8 per thread
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = t + 256 * r;
      {  // A slice: row e / 32, col e % 32
        const int mm = e >> 5, kk = e & 31;
        const int gm = m0 + mm, gk = k0 + kk;
        uint16_t v = 0;
        if (gm < g.M && gk < g.K)
          v = g.transA ? g.A[(size_t)gk * g.lda + gm] : g.A[(size_t)gm * g.lda + gk];
        As[mm][kk] = v;
      }
      {  // B slice: row e / 64, col e % 64
        const int kk = e >> 6, nn = e & 63;
        const int gk = k0 + kk, gn = n0 + nn;
        uint16_t v = 0;
        if (gk < g.K && gn < g.N)
          v = g.transB ? g.B[(size_t)gn * g.ldb + gk] : g.B[(size_t)gk * g.ldb + gn];
        Bs[kk][nn] = v;
      }
    }
    __syncthreads();
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    bf8 a;
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = __builtin_bit_cast(__bf16, As[16 * w + fr][fk + i]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bf8 b;
#pragma unroll
      for (int i = 0; i < 8; ++i) b[i] = __builtin_bit_cast(__bf16, Bs[fk + i][16 * j + fr]);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    }
    __syncthreads();
  }
  const int col_in = lane & 15, row_in = (lane >> 4) * 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + 16 * j + col_in;
    if (n >= g.N) continue;
    const float bn = g.bias ? g.bias[n] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + 16 * w + row_in + i;
      if (m >= g.M) continue;
      const float x = acc[j][i];
      if (g.epi == kEpiF32) {
        g.Cf[(size_t)m * g.ldc + n] = x + bn;
      } else if (g.epi == kEpiBiasReluBf16) {
        const float y = x + bn;
        g.Cb[(size_t)m * g.ldc + n] = f2bf(y > 0.f ? y : 0.f);
      } else {
        const bool keep = bf2f(g.mask[(size_t)m * g.ldm + n]) > 0.f;
        g.Cb[(size_t)m * g.ldc + n] = f2bf(keep ? x : 0.f);
      }
    }
  }
}

// Grouped softmax cross-entropy: per sample and head h with label y
// (255 = head ignored): loss += logsumexp(z_h) - z_h[y];
// dz_h = scale * (softmax(z_h) - onehot(y)), written as bf16.
__global__ __launch_bounds__(256) void xent_kernel(const float* __restrict__ z,
                                                   const uint8_t* __restrict__ labels, int64_t n,
                                                   float scale, uint16_t* __restrict__ dz,
                                                   float* __restrict__ loss) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const float* zr = z + b * kOut;
  uint16_t* dr = dz + b * kOut;
  float L = 0.f;
  for (int h = 0; h < kHeads; ++h) {
    const int o = kHeadOff[h], c = kHeadOff[h + 1] - o;
    const int y = labels[b * kHeads + h];
    if (y == 255 || y >= c) {
      for (int j = 0; j < c; ++j) dr[o + j] = 0;
      continue;
    }
    float mx = zr[o];
    for (int j = 1; j < c; ++j) mx = fmaxf(mx, zr[o + j]);
    float s = 0.f;
    for (int j = 0; j < c; ++j) s += expf(zr[o + j] - mx);
    L += logf(s) + mx - zr[o + y];
    const float inv = 1.f / s;
    for (int j = 0; j < c; ++j) {
      const float p = expf(zr[o + j] - mx) * inv;
      dr[o + j] = f2bf(scale * (p - (j == y ? 1.f : 0.f)));
    }
  }
  for (int j = kHeadOff[kHeads]; j < kOut; ++j) dr[j] = 0;
  if (loss) loss[b] = L;
}

// Inference heads: predictor heads as probabilities (n x 51 float), coder
// heads as arg-max class (lowest index on ties) and its probability.
__global__ __launch_bounds__(256) void heads_kernel(const float* __restrict__ z, int64_t n,
                                                    float* __restrict__ pred_prob,
                                                    uint8_t* __restrict__ cls,
                                                    float* __restrict__ cls_prob) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const float* zr = z + b * kOut;
  for (int h = 0; h < kHeads; ++h) {
    const int o = kHeadOff[h], c = kHeadOff[h + 1] - o;
    float mx = zr[o];
    int am = 0;
    for (int j = 1; j < c; ++j)
      if (zr[o + j] > mx) { mx = zr[o + j]; am = j; }
    float s = 0.f;
    for (int j = 0; j < c; ++j) s += expf(zr[o + j] - mx);
    const float inv = 1.f / s;
    if (h < 3) {
      for (int j = 0; j < c; ++j) pred_prob[b * 51 + o + j] = expf(zr[o + j] - mx) * inv;
    } else {
      cls[b * 6 + (h - 3)] = (uint8_t)am;
      cls_prob[b * 6 + (h - 3)] = inv;   // exp(0) / s
    }
  }
}

// Bias gradient: column sums of a bf16 (M x N) matrix, fixed row order.
__global__ __launch_bounds__(256) void colsum_kernel(const uint16_t* __restrict__ G, int64_t M,
                                                     int N, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int64_t m = 0; m < M; ++m) s += bf2f(G[m * N + c]);
  out[c] = s;
}

// Adam (Kingma & Ba 2015) with bias correction folded into c1 = 1/(1-b1^t),
// c2 = 1/(1-b2^t); refreshes the bf16 copy of the weights.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, float* __restrict__ m,
                                                   float* __restrict__ v,
                                                   const float* __restrict__ g,
                                                   uint16_t* __restrict__ wb, int64_t n, int64_t n_w,
                                                   float lr, float b1, float b2, float eps,
                                                   float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i] = pi;
  if (i < n_w) wb[i] = f2bf(pi);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float* __restrict__ x, int64_t n,
                                                      uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

static int launch_gemm(const GemmArgs& g, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0 || g.K < 0) return set_error(STRIDE_E_ARG, "gemm: bad shape");
  const unsigned gx = (unsigned)((g.N + 63) / 64), gy = (unsigned)((g.M + 63) / 64);
  if ((g.M + 63) / 64 > 0x7FFFFFFF) return set_error(STRIDE_E_ARG, "gemm: M too large");
  hipLaunchKernelGGL(gemm_kernel, dim3(gx, gy), dim3(256), 0, st, g);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? STRIDE_OK : set_error(STRIDE_E_HIP, hipGetErrorString(e));
}

static GemmArgs gemm(const uint16_t* A, int lda, int tA, const uint16_t* B, int ldb, int tB, int M,
                     int N, int K, int epi, const float* bias, const uint16_t* mask, int ldm,
                     float* Cf, uint16_t* Cb, int ldc) {
  GemmArgs g;
  g.A = A; g.B = B; g.lda = lda; g.ldb = ldb; g.transA = tA; g.transB = tB;
  g.M = M; g.N = N; g.K = K; g.epi = epi; g.bias = bias; g.mask = mask; g.ldm = ldm;
  g.Cf = Cf; g.Cb = Cb; g.ldc = ldc;
  return g;
}

static int last(void) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? STRIDE_OK : set_error(STRIDE_E_HIP, hipGetErrorString(e));
}

}  // namespace mlp
}  // namespace stride

using namespace stride::mlp;
using stride::set_error;

#define MLP_TRY(x)              \
  do {                          \
    int rc_ = (x);              \
    if (rc_ != STRIDE_OK) return rc_; \
  } while (0)

extern "C" {

int stride_mlp_sizes(int32_t* n_in, int32_t* n_hidden, int32_t* n_out, int32_t* n_weights,
                     int32_t* n_params) {
  if (n_in) *n_in = kIn;
  if (n_hidden) *n_hidden = kHid;
  if (n_out) *n_out = kOut;
  if (n_weights) *n_weights = kNW;
  if (n_params) *n_params = kNP;
  return STRIDE_OK;
}

int stride_mlp_gemm(const uint16_t* d_a, int32_t lda, int32_t trans_a, const uint16_t* d_b,
                    int32_t ldb, int32_t trans_b, int32_t M, int32_t N, int32_t K, int32_t epi,
                    const float* d_bias, const uint16_t* d_mask, int32_t ldm, float* d_cf,
                    uint16_t* d_cb, int32_t ldc, void* stream) {
  if (!d_a || !d_b || M <= 0 || N <= 0 || K <= 0 || epi < 0 || epi > 2 ||
      (epi == kEpiF32 && !d_cf) || (epi != kEpiF32 && !d_cb) || (epi == kEpiMaskBf16 && !d_mask) ||
      lda <= 0 || ldb <= 0 || ldc < N)
    return set_error(STRIDE_E_ARG, "stride_mlp_gemm: invalid argument");
  if ((trans_a ? lda < M : lda < K) || (trans_b ? ldb < K : ldb < N) ||
      (epi == kEpiMaskBf16 && ldm < N))
    return set_error(STRIDE_E_ARG, "stride_mlp_gemm: leading dimension too small");
  return launch_gemm(gemm(d_a, lda, trans_a, d_b, ldb, trans_b, M, N, K, epi, d_bias, d_mask, ldm,
                          d_cf, d_cb, ldc),
                     (hipStream_t)stream);
}

int stride_mlp_forward(const uint16_t* d_x, int64_t n, const uint16_t* d_wb, const float* d_params,
                       uint16_t* d_h1, uint16_t* d_h2, float* d_z, void* stream) {
  if (n <= 0 || n > 0x7FFFFFFF || !d_x || !d_wb || !d_params || !d_h1 || !d_h2 || !d_z)
    return set_error(STRIDE_E_ARG, "stride_mlp_forward: invalid argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = (int)n;
  const uint16_t *W1 = d_wb, *W2 = W1 + kIn * kHid, *W3 = W2 + kHid * kHid;
  const float *b1 = d_params + kNW, *b2 = b1 + kHid, *b3 = b2 + kHid;
  MLP_TRY(launch_gemm(gemm(d_x, kIn, 0, W1, kHid, 0, N, kHid, kIn, kEpiBiasReluBf16, b1, nullptr,
                           0, nullptr, d_h1, kHid), st));
  MLP_TRY(launch_gemm(gemm(d_h1, kHid, 0, W2, kHid, 0, N, kHid, kHid, kEpiBiasReluBf16, b2,
                           nullptr, 0, nullptr, d_h2, kHid), st));
  return launch_gemm(gemm(d_h2, kHid, 0, W3, kOut, 0, N, kOut, kHid, kEpiF32, b3, nullptr, 0, d_z,
                          nullptr, kOut), st);
}

int stride_mlp_xent(const float* d_z, const uint8_t* d_labels, int64_t n, float scale,
                    uint16_t* d_dz, float* d_loss, void* stream) {
  if (n <= 0 || !d_z || !d_labels || !d_dz) return set_error(STRIDE_E_ARG, "stride_mlp_xent: invalid argument");
  hipLaunchKernelGGL(xent_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d_z, d_labels, n, scale, d_dz, d_loss);
  return last();
}

int stride_mlp_backward(const uint16_t* d_x, int64_t n, const uint16_t* d_wb, const uint16_t* d_h1,
                        const uint16_t* d_h2, const uint16_t* d_dz, uint16_t* d_dh2,
                        uint16_t* d_dh1, float* d_grad, void* stream) {
  if (n <= 0 || n > 0x7FFFFFFF || !d_x || !d_wb || !d_h1 || !d_h2 || !d_dz || !d_dh2 || !d_dh1 ||
      !d_grad)
    return set_error(STRIDE_E_ARG, "stride_mlp_backward: invalid argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = (int)n;
  const uint16_t *W2 = d_wb + kIn * kHid, *W3 = W2 + kHid * kHid;
  float *gW1 = d_grad, *gW2 = gW1 + kIn * kHid, *gW3 = gW2 + kHid * kHid;
  float *gb1 = d_grad + kNW, *gb2 = gb1 + kHid, *gb3 = gb2 + kHid;
  // layer 3: dW3 = H2^T dZ, db3 = colsum dZ, dH2 = (dZ W3^T) * (H2 > 0)
  MLP_TRY(launch_gemm(gemm(d_h2, kHid, 1, d_dz, kOut, 0, kHid, kOut, N, kEpiF32, nullptr, nullptr,
                           0, gW3, nullptr, kOut), st));
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, st, d_dz, n, kOut, gb3);
  MLP_TRY(last());
  MLP_TRY(launch_gemm(gemm(d_dz, kOut, 0, W3, kOut, 1, N, kHid, kOut, kEpiMaskBf16, nullptr, d_h2,
                           kHid, nullptr, d_dh2, kHid), st));
  // layer 2
  MLP_TRY(launch_gemm(gemm(d_h1, kHid, 1, d_dh2, kHid, 0, kHid, kHid, N, kEpiF32, nullptr, nullptr,
                           0, gW2, nullptr, kHid), st));
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, st, d_dh2, n, kHid, gb2);
  MLP_TRY(last());
  MLP_TRY(launch_gemm(gemm(d_dh2, kHid, 0, W2, kHid, 1, N, kHid, kHid, kEpiMaskBf16, nullptr, d_h1,
                           kHid, nullptr, d_dh1, kHid), st));
  // layer 1
  MLP_TRY(launch_gemm(gemm(d_x, kIn, 1, d_dh1, kHid, 0, kIn, kHid, N, kEpiF32, nullptr, nullptr, 0,
                           gW1, nullptr, kHid), st));
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, st, d_dh1, n, kHid, gb1);
  return last();
}

int stride_mlp_adam(float* d_params, float* d_m, float* d_v, const float* d_grad, uint16_t* d_wb,
                    int64_t n_params, float lr, float beta1, float beta2, float eps, int32_t step,
                    void* stream) {
  if (!d_params || !d_m || !d_v || !d_grad || !d_wb || n_params != kNP || step < 1 ||
      !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
    return set_error(STRIDE_E_ARG, "stride_mlp_adam: invalid argument");
  const double p1 = std::pow((double)beta1, (double)step), p2 = std::pow((double)beta2, (double)step);
  const float c1 = (float)(1.0 / (1.0 - p1)), c2 = (float)(1.0 / (1.0 - p2));
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n_params + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d_params, d_m, d_v, d_grad, d_wb, n_params,
                     (int64_t)kNW, lr, beta1, beta2, eps, c1, c2);
  return last();
}

int stride_mlp_to_bf16(const float* d_x, int64_t n, uint16_t* d_y, void* stream) {
  if (n < 0 || (n && (!d_x || !d_y))) return set_error(STRIDE_E_ARG, "stride_mlp_to_bf16: invalid argument");
  if (n == 0) return STRIDE_OK;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d_x, n, d_y);
  return last();
}

int stride_mlp_heads(const float* d_z, int64_t n, float* d_pred_prob, uint8_t* d_cls,
                     float* d_cls_prob, void* stream) {
  if (n <= 0 || !d_z || !d_pred_prob || !d_cls || !d_cls_prob)
    return set_error(STRIDE_E_ARG, "stride_mlp_heads: invalid argument");
  hipLaunchKernelGGL(heads_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d_z, n, d_pred_prob, d_cls, d_cls_prob);
  return last();
}

}  // extern "C"
