// Encoder search kernels for gfx950.
//
//  * stream costs / selection: one wave64 per residual stream; each lane
//    folds its values into the 20 candidate body costs (Rice k = 0..11,
//    Exp-Golomb k = 0..7, residual_entropy.py:39-54 exact integer form),
//    then a cross-lane reduction and the tie-break of _pick_best_k
//    (paradelta_v5_dup.py:159-171, paradelta_codec.py:131-146).
//  * IRLP candidate cost: one workgroup per meshlet, one lane per candidate
//    predictor; the meshlet's samples are wave-uniform (scalar loads), each
//    lane evaluates the parallelogram residual of its candidate and the 20
//    coder costs, keeps the minimum and adds it to the candidate's total
//    (_cost_seeds_njit, encoder/_irlp_fit.py:120-183).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/stride_hip.h"
#include "stride_err.h"

namespace stride {

constexpr int kRiceK = 12;
constexpr int kEgK = 8;
constexpr int kCoders = kRiceK + kEgK;

__device__ __forceinline__ uint32_t zz32(int32_t r) {
  return r < 0 ? ((uint32_t)(-(int64_t)r) << 1) - 1u : ((uint32_t)r << 1);
}
__device__ __forceinline__ uint64_t zz64(int64_t r) {
  return r < 0 ? ((uint64_t)(-r) << 1) - 1u : ((uint64_t)r << 1);
}
// floor(log2(x)) for x >= 1
__device__ __forceinline__ uint32_t ilog2_64(uint64_t x) { return 63u - (uint32_t)__builtin_clzll(x); }

__device__ __forceinline__ void accumulate(uint64_t u, uint64_t (&acc)[kCoders]) {
#pragma unroll
  for (int k = 0; k < kRiceK; ++k) acc[k] += u >> k;
#pragma unroll
  for (int k = 0; k < kEgK; ++k) acc[kRiceK + k] += 2u * ilog2_64((u >> k) + 1);
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, o, 64);
    hi = __shfl_xor(hi, o, 64);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// Body bits of candidate j for n values given the accumulated sum.
__device__ __forceinline__ uint64_t body_bits(int j, uint64_t acc, uint64_t n) {
  if (j < kRiceK) return acc + n * (uint64_t)(1 + j);
  const int k = j - kRiceK;
  return acc + n * (uint64_t)(1 + k);
}

__global__ __launch_bounds__(256) void stream_cost_kernel(const int32_t* __restrict__ vals,
                                                          const int64_t* __restrict__ off,
                                                          int64_t n_streams,
                                                          int64_t* __restrict__ costs,
                                                          uint8_t* __restrict__ tagk,
                                                          int64_t* __restrict__ best_bits) {
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (s >= n_streams) return;
  const int64_t b = off[s], e = off[s + 1];
  uint64_t acc[kCoders];
#pragma unroll
  for (int j = 0; j < kCoders; ++j) acc[j] = 0;
  for (int64_t i = b + lane; i < e; i += 64) accumulate(zz32(vals[i]), acc);
  const uint64_t n = (uint64_t)(e - b);
  uint64_t tot[kCoders];
#pragma unroll
  for (int j = 0; j < kCoders; ++j) tot[j] = body_bits(j, wave_sum(acc[j]), n);
  if (lane == 0) {
    if (costs) {
#pragma unroll
      for (int j = 0; j < kCoders; ++j) costs[s * kCoders + j] = (int64_t)tot[j];
    }
    if (tagk) {
      int rk = 0, ek = 0;
      uint64_t rb = tot[0], eb = tot[kRiceK];
      for (int k = 1; k < kRiceK; ++k)
        if (tot[k] < rb) { rb = tot[k]; rk = k; }
      for (int k = 1; k < kEgK; ++k)
        if (tot[kRiceK + k] < eb) { eb = tot[kRiceK + k]; ek = k; }
      if (n == 0) { rb = 0; eb = 0; rk = 0; ek = 0; }
      const bool rice = rb <= eb;
      tagk[2 * s] = rice ? 1 : 2;
      tagk[2 * s + 1] = (uint8_t)(rice ? rk : ek);
      if (best_bits) best_bits[s] = (int64_t)(rice ? rb : eb);
    }
  }
}

__global__ __launch_bounds__(64) void irlp_cost_kernel(const int32_t* __restrict__ A,
                                                       const int32_t* __restrict__ B,
                                                       const int32_t* __restrict__ C,
                                                       const int32_t* __restrict__ Vv,
                                                       const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ cand, int S,
                                                       int K, unsigned long long* __restrict__ cost,
                                                       int64_t* __restrict__ per_m) {
  const int64_t m = blockIdx.x;
  const int sc = blockIdx.y * 64 + threadIdx.x;
  const int64_t b = off[m], e = off[m + 1];
  if (sc >= S) return;
  if (e == b) {
    if (per_m) per_m[m * S + sc] = 0;
    return;
  }
  const int64_t n0 = cand[3 * sc], n1 = cand[3 * sc + 1], n2 = cand[3 * sc + 2];
  const int64_t half = K ? ((int64_t)1 << (K - 1)) : 0;
  uint64_t acc[kCoders];
#pragma unroll
  for (int j = 0; j < kCoders; ++j) acc[j] = 0;
  for (int64_t i = b; i < e; ++i) {
    int64_t s = n0 * A[i] + n1 * B[i] + n2 * C[i];
    int64_t pred = K ? ((s + half) >> K) : s;
    accumulate(zz64((int64_t)Vv[i] - pred), acc);
  }
  const uint64_t n = (uint64_t)(e - b);
  uint64_t best = ~0ull;
#pragma unroll
  for (int j = 0; j < kCoders; ++j) {
    uint64_t t = body_bits(j, acc[j], n);
    best = t < best ? t : best;
  }
  if (per_m)
    per_m[m * S + sc] = (int64_t)best;
  else
    atomicAdd(&cost[sc], (unsigned long long)best);
}

__global__ void zero_i64_kernel(int64_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace stride

using namespace stride;

extern "C" int stride_stream_costs(const int32_t* d_vals, const int64_t* d_off, int64_t n_streams,
                                   int64_t* d_costs, void* stream) {
  if (n_streams < 0 || (n_streams && (!d_vals || !d_off || !d_costs)))
    return set_error(STRIDE_E_ARG, "stride_stream_costs: null argument");
  if (n_streams == 0) return STRIDE_OK;
  const int64_t blocks = (n_streams + 3) / 4;
  hipLaunchKernelGGL(stream_cost_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, d_vals, d_off, n_streams, d_costs, nullptr, nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(STRIDE_E_HIP, hipGetErrorString(e));
  return STRIDE_OK;
}

extern "C" int stride_select_codes(const int32_t* d_vals, const int64_t* d_off, int64_t n_streams,
                                   uint8_t* d_tagk, int64_t* d_bits, void* stream) {
  if (n_streams < 0 || (n_streams && (!d_vals || !d_off || !d_tagk)))
    return set_error(STRIDE_E_ARG, "stride_select_codes: null argument");
  if (n_streams == 0) return STRIDE_OK;
  const int64_t blocks = (n_streams + 3) / 4;
  hipLaunchKernelGGL(stream_cost_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, d_vals, d_off, n_streams, nullptr, d_tagk, d_bits);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(STRIDE_E_HIP, hipGetErrorString(e));
  return STRIDE_OK;
}

extern "C" int stride_irlp_cost(const int32_t* d_a, const int32_t* d_b, const int32_t* d_c,
                                const int32_t* d_v, const int64_t* d_off, int64_t n_m,
                                const int32_t* d_cand, int32_t S, int32_t K, int64_t* d_cost,
                                void* stream) {
  if (n_m < 0 || S < 0 || K < 0 || K > 30 || !d_cost || (S && !d_cand) ||
      (n_m && (!d_a || !d_b || !d_c || !d_v || !d_off)))
    return set_error(STRIDE_E_ARG, "stride_irlp_cost: invalid argument");
  if (S == 0) return STRIDE_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_i64_kernel, dim3((S + 255) / 256), dim3(256), 0, st, d_cost, (int)S);
  if (n_m > 0) {
    if (n_m > 0x7FFFFFFF) return set_error(STRIDE_E_ARG, "stride_irlp_cost: too many meshlets");
    hipLaunchKernelGGL(irlp_cost_kernel, dim3((unsigned)n_m, (S + 63) / 64), dim3(64), 0, st, d_a,
                       d_b, d_c, d_v, d_off, d_cand, (int)S, (int)K,
                       reinterpret_cast<unsigned long long*>(d_cost), nullptr);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(STRIDE_E_HIP, hipGetErrorString(e));
  return STRIDE_OK;
}

extern "C" int stride_irlp_cost_meshlets(const int32_t* d_a, const int32_t* d_b, const int32_t* d_c,
                                         const int32_t* d_v, const int64_t* d_off, int64_t n_m,
                                         const int32_t* d_cand, int32_t S, int32_t K,
                                         int64_t* d_cost_m, void* stream) {
  if (n_m < 0 || S < 0 || K < 0 || K > 30 || (S && !d_cand) ||
      (n_m && (!d_a || !d_b || !d_c || !d_v || !d_off || !d_cost_m)))
    return set_error(STRIDE_E_ARG, "stride_irlp_cost_meshlets: invalid argument");
  if (S == 0 || n_m == 0) return STRIDE_OK;
  if (n_m > 0x7FFFFFFF) return set_error(STRIDE_E_ARG, "stride_irlp_cost_meshlets: too many meshlets");
  hipLaunchKernelGGL(irlp_cost_kernel, dim3((unsigned)n_m, (S + 63) / 64), dim3(64), 0,
                     (hipStream_t)stream, d_a, d_b, d_c, d_v, d_off, d_cand, (int)S, (int)K, nullptr,
                     d_cost_m);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(STRIDE_E_HIP, hipGetErrorString(e));
  return STRIDE_OK;
}
