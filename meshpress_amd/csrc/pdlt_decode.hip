// gfx950 PDLT v5 decoder: host set-up (header, boundary table, pre-pass,
// wave groups) and the decode kernels.
//
// One wave decodes a group of up to 16 consecutive meshlets whose stream
// bits are staged in LDS.  Lane l serves meshlet slot l & 15 in role l >> 4:
//   phase A  role 0 (token lane): boundary refs, strips -> walk, triangles;
//   phase B  all roles: boundary reconstructions gathered into LDS;
//   phase C  roles 1..3 (axis lanes): boundary centroid and the float64
//            LIN3 chain of one axis along the walk;
//   phase D  all 64 lanes: interior vertices written out.
// Boundary-table vertices are written by a separate element-wise kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/stride_hip.h"
#include "pdlt_decode.h"
#include "stride_err.h"

namespace stride {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pdlt_boundary_kernel(const double* __restrict__ bnd,
                                                            uint64_t n3, PdltParams P,
                                                            float* __restrict__ V) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n3;
       i += (uint64_t)gridDim.x * 256) {
    V[i] = pdlt_out(pdlt_bnd(bnd[i], P), P, (int)(i % 3));
  }
}

// Phases B-D of a group (after the walk, seen set, walk length and refs of
// every meshlet are in its LDS block): boundary gather, axis chains, output.
// T: the reconstruction type, double (float64 chain) or float (float32 chain,
// half the LDS per vertex); a block holds three T arrays of n_local, then
// refs, walk and seen set.
template <class T>
__device__ __forceinline__ void pdlt_group_tail(const uint32_t* __restrict__ lds_c, uint32_t* lds,
                                                const PdltGroup& G, const PdltEntry& e, bool active,
                                                int role, int lane, int64_t rel,
                                                const PdltEntry* __restrict__ tab,
                                                const double* __restrict__ bnd, const PdltParams& P,
                                                float* __restrict__ V) {
  constexpr uint32_t RW = 3 * sizeof(T) / 4;   // block words of reconstruction per local
  const uint32_t nl = e.n_bnd + e.n_int;
  uint32_t* blk = lds + G.blk_base + e.blk_off;
  T* recon = reinterpret_cast<T*>(blk);
  const uint32_t* refs = blk + RW * nl;
  const uint32_t* walk = refs + e.n_bnd;
  const uint32_t* n_emit = walk + e.n_int + 8;
  // boundary reconstructions: the four lanes of a slot gather its refs
  if (active) {
    const uint32_t nb3 = 3 * e.n_bnd;
    for (uint32_t base = (uint32_t)role; base < nb3; base += 32) {
      T x[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t idx = base + 4 * u;
        if (idx < nb3) {
          const uint32_t i = idx / 3, d = idx - 3 * i;
          x[u] = (T)pdlt_bnd(bnd[3ull * refs[i] + d], P);
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t idx = base + 4 * u;
        if (idx < nb3) {
          const uint32_t i = idx / 3, d = idx - 3 * i;
          recon[d * nl + i] = x[u];
        }
      }
    }
  }
  __syncthreads();
  if (active && role > 0)
    pdlt_axis(lds_c, rel, e, role - 1, P, recon + (uint32_t)(role - 1) * nl, walk, *n_emit);
  __syncthreads();
  // interior vertices, one meshlet at a time, 64 coordinates per step
  for (uint32_t j = 0; j < G.m1 - G.m0; ++j) {
    const PdltEntry ej = tab[G.m0 + j];
    const uint32_t nlj = ej.n_bnd + ej.n_int;
    const uint32_t* bj = lds + G.blk_base + ej.blk_off;
    const T* rj = reinterpret_cast<const T*>(bj);
    const uint32_t* sj = bj + RW * nlj + ej.n_bnd + ej.n_int;
    float* vo = V + 3ull * ej.int_base;
    const uint32_t n3 = 3 * ej.n_int;
    for (uint32_t t = lane; t < n3; t += 64) {
      const uint32_t i = t / 3, d = t - 3 * i, v = ej.n_bnd + i;
      const bool reached = ((sj[v >> 5] >> (v & 31)) & 1u) != 0;
      vo[t] = pdlt_out(reached ? (double)rj[d * nlj + v] : 0.0, P, (int)d);
    }
  }
}

// Single-pass decode: the token lanes run inside the group.  R: the
// reconstruction type (pdlt_group_tail).
template <class R>
__global__ __launch_bounds__(64) void pdlt_decode_kernel(
    const uint32_t* __restrict__ words, const PdltEntry* __restrict__ tab,
    const PdltGroup* __restrict__ groups, const double* __restrict__ bnd, PdltParams P,
    float* __restrict__ V, uint32_t* __restrict__ T) {
  extern __shared__ uint64_t lds64[];
  uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
  const int lane = threadIdx.x;
  const int slot = lane & 15, role = lane >> 4;
  const PdltGroup G = groups[blockIdx.x];
  stage_words(lds, words, G.w_lo, G.n_words, (uint32_t)lane, 64);
  const uint32_t m = G.m0 + (uint32_t)slot;
  const bool active = m < G.m1;
  PdltEntry e = {};
  if (active) e = tab[m];
  const uint32_t nl = e.n_bnd + e.n_int;
  uint32_t* blk = lds + G.blk_base + e.blk_off;
  uint32_t* refs = blk + (3 * sizeof(R) / 4) * nl;
  uint32_t* walk = refs + e.n_bnd;
  uint32_t* seen = walk + e.n_int;
  const int64_t rel = -32 * (int64_t)G.w_lo;
  __syncthreads();
  if (active && role == 0)
    pdlt_tokens(lds, rel, e, refs, walk, BitSeen{seen}, seen + 8, T + 3ull * e.t_off);
  __syncthreads();
  pdlt_group_tail<R>(lds, lds, G, e, active, role, lane, rel, tab, bnd, P, V);
}

// Two-pass decode, pass 1: one token lane per meshlet, up to 64 meshlets
// per wave (refs in LDS for the global triangle ids; the lanes' seen sets
// interleaved in LDS, lane i's word q at seen_base + 64 q + i, so one access
// of the wave touches distinct banks); the walk, seen set and walk length go
// to HBM.
__global__ __launch_bounds__(64) void pdlt_tokens_kernel(
    const uint32_t* __restrict__ words, const PdltEntry* __restrict__ tab,
    const PdltGroup* __restrict__ groups, uint32_t n_b, uint32_t* __restrict__ walk_g,
    uint32_t* __restrict__ seen_g, uint32_t* __restrict__ T) {
  extern __shared__ uint64_t lds64[];
  uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
  const int lane = threadIdx.x;
  const PdltGroup G = groups[blockIdx.x];
  stage_words(lds, words, G.w_lo, G.n_words, (uint32_t)lane, 64);
  __syncthreads();
  const uint32_t m = G.m0 + (uint32_t)lane;
  if (m < G.m1) {
    const PdltEntry e = tab[m];
    uint32_t* refs = lds + G.blk_base + e.tok_ref_off;
    uint32_t* seen = lds + G.seen_base + lane;
    uint32_t n_emit = 0;
    pdlt_tokens<WalkSlot>(lds, -32 * (int64_t)G.w_lo, e, refs,
                          walk_g + pdlt_walk_off(e.int_base, m, n_b),
                          BitSeen{seen, kTokenSeenStride}, &n_emit, T + 3ull * e.t_off);
    uint32_t* so = seen_g + (uint64_t)kSeenWords * m;
#pragma unroll
    for (int q = 0; q < 8; ++q) so[q] = seen[q * kTokenSeenStride];
    so[8] = n_emit;
  }
}

// Two-pass decode, pass 2: the group's blocks take the walk, seen set and
// walk length from HBM and the refs decoded again, then phases B-D.
template <class R>
__global__ __launch_bounds__(64) void pdlt_chain_kernel(
    const uint32_t* __restrict__ words, const PdltEntry* __restrict__ tab,
    const PdltGroup* __restrict__ groups, const double* __restrict__ bnd, PdltParams P,
    uint32_t n_b, const uint32_t* __restrict__ walk_g, const uint32_t* __restrict__ seen_g,
    float* __restrict__ V) {
  extern __shared__ uint64_t lds64[];
  uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
  const int lane = threadIdx.x;
  const int slot = lane & 15, role = lane >> 4;
  const PdltGroup G = groups[blockIdx.x];
  stage_words(lds, words, G.w_lo, G.n_words, (uint32_t)lane, 64);
  const uint32_t m = G.m0 + (uint32_t)slot;
  const bool active = m < G.m1;
  PdltEntry e = {};
  if (active) e = tab[m];
  const uint32_t nl = e.n_bnd + e.n_int;
  uint32_t* blk = lds + G.blk_base + e.blk_off;
  uint32_t* refs = blk + (3 * sizeof(R) / 4) * nl;
  uint32_t* walk = refs + e.n_bnd;
  uint32_t* seen = walk + e.n_int;
  const int64_t rel = -32 * (int64_t)G.w_lo;
  if (active) {
    const uint32_t* ws = walk_g + pdlt_walk_off(e.int_base, m, n_b);
    for (uint32_t i = (uint32_t)role; i < e.n_int; i += 4) walk[i] = ws[i];
    const uint32_t* ss = seen_g + (uint64_t)kSeenWords * m;
    for (uint32_t i = (uint32_t)role; i < kSeenWords; i += 4) seen[i] = ss[i];
  }
  __syncthreads();
  if (active && role == 0) pdlt_refs(lds, rel, e, refs);
  __syncthreads();
  pdlt_group_tail<R>(lds, lds, G, e, active, role, lane, rel, tab, bnd, P, V);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define PDLT_HIP(expr)                                                                \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return set_error(STRIDE_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

PdltSet::~PdltSet() {
  (void)hipFree(d_words);
  (void)hipFree(d_tab);
  (void)hipFree(d_groups);
  (void)hipFree(d_bnd);
  (void)hipFree(d_v_own);
  (void)hipFree(d_t_own);
  (void)hipFree(d_tok_groups);
  (void)hipFree(d_walk);
  (void)hipFree(d_seen);
}

size_t PdltSet::side_bytes() const {
  return H.tab.size() * sizeof(PdltEntry) + groups.size() * sizeof(PdltGroup) +
         H.bnd.size() * sizeof(double) +
         (two_pass ? tok_groups.size() * sizeof(PdltGroup) : 0);
}

int PdltSet::build(const uint8_t* d, size_t len) {
  std::string err;
  int rc = pdlt_parse(d, len, H, err);
  if (rc) return set_error(rc, err);
  // two-pass scratch: the walks (one word per interior vertex and one per
  // meshlet) and the seen sets with walk lengths
  const uint64_t n_walk = H.n_v - H.n_b + H.n_m;
  PDLT_HIP(hipMalloc(&d_walk, 4 * std::max<uint64_t>(1, n_walk)));
  PDLT_HIP(hipMalloc(&d_seen, 4ull * kSeenWords * std::max<uint32_t>(1, H.n_m)));
  PDLT_HIP(hipMalloc(&d_words, H.words.size() * 4));
  PDLT_HIP(hipMemcpy(d_words, H.words.data(), H.words.size() * 4, hipMemcpyHostToDevice));
  if (!H.bnd.empty()) {
    PDLT_HIP(hipMalloc(&d_bnd, H.bnd.size() * sizeof(double)));
    PDLT_HIP(hipMemcpy(d_bnd, H.bnd.data(), H.bnd.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (!H.tab.empty()) PDLT_HIP(hipMalloc(&d_tab, H.tab.size() * sizeof(PdltEntry)));
  return set_layout(group_size, lds_budget, tok_lds_budget);
}

int PdltSet::set_layout(int gs, int budget, int tok_budget) {
  std::string err;
  std::vector<PdltGroup> gr, tg;
  uint32_t lds = 0, tlds = 0;
  std::vector<PdltEntry> tab = H.tab;
  // the float32 chain keeps three float32 per vertex, the float64 chain three float64
  int rc = pdlt_layout(tab, gs, budget, gr, lds, err, chain32 ? 3u : 6u);
  if (!rc) rc = pdlt_layout_tokens(tab, tok_budget, tok_group, tg, tlds, err);
  if (rc) return set_error(rc, err);
  // the new device tables are built aside and take over only once every
  // allocation and copy has succeeded; until then the set is not ready
  ready = false;
  PdltGroup* ng = nullptr;
  PdltGroup* nt = nullptr;
  hipError_t e = hipSuccess;
  if (!gr.empty()) {
    e = hipMalloc(&ng, gr.size() * sizeof(PdltGroup));
    if (e == hipSuccess)
      e = hipMemcpy(ng, gr.data(), gr.size() * sizeof(PdltGroup), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !tg.empty()) {
    e = hipMalloc(&nt, tg.size() * sizeof(PdltGroup));
    if (e == hipSuccess)
      e = hipMemcpy(nt, tg.data(), tg.size() * sizeof(PdltGroup), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !tab.empty())
    e = hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(PdltEntry), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(ng);
    (void)hipFree(nt);
    return set_error(STRIDE_E_HIP, std::string("PDLT launch layout: ") + hipGetErrorString(e));
  }
  H.tab.swap(tab);
  groups.swap(gr);
  tok_groups.swap(tg);
  lds_bytes = lds;
  tok_lds_bytes = tlds;
  group_size = gs;
  lds_budget = budget;
  tok_lds_budget = tok_budget;
  (void)hipFree(d_groups);
  (void)hipFree(d_tok_groups);
  d_groups = ng;
  d_tok_groups = nt;
  ready = true;
  return STRIDE_OK;
}

int PdltSet::own_outputs() {
  if (!d_v_own && H.n_v) PDLT_HIP(hipMalloc(&d_v_own, 12 * H.n_v));
  if (!d_t_own && H.n_t) PDLT_HIP(hipMalloc(&d_t_own, 12 * H.n_t));
  return STRIDE_OK;
}

int PdltSet::decode(float* V, uint32_t* T, int world, hipStream_t s) {
  if ((!V && H.n_v) || (!T && H.n_t)) return set_error(STRIDE_E_ARG, "null output pointer");
  if (!ready)
    return set_error(STRIDE_E_HIP, "the decoder has no launch layout (the last set_tune failed)");
  PdltParams p = H.params;
  p.world = world ? 1 : 0;
  p.chain32 = chain32;
  if (H.n_b) {
    const uint64_t n3 = 3ull * H.n_b;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n3 + 255) / 256, 8192);
    hipLaunchKernelGGL(pdlt_boundary_kernel, dim3(grid), dim3(256), 0, s, d_bnd, n3, p, V);
    PDLT_HIP(hipGetLastError());
  }
  if (groups.empty()) return STRIDE_OK;
  if (two_pass) {
    hipLaunchKernelGGL(pdlt_tokens_kernel, dim3((uint32_t)tok_groups.size()), dim3(64), tok_lds_bytes,
                       s, d_words, d_tab, d_tok_groups, H.n_b, d_walk, d_seen, T);
    PDLT_HIP(hipGetLastError());
    if (chain32)
      hipLaunchKernelGGL(pdlt_chain_kernel<float>, dim3((uint32_t)groups.size()), dim3(64),
                         lds_bytes, s, d_words, d_tab, d_groups, d_bnd, p, H.n_b, d_walk, d_seen, V);
    else
      hipLaunchKernelGGL(pdlt_chain_kernel<double>, dim3((uint32_t)groups.size()), dim3(64),
                         lds_bytes, s, d_words, d_tab, d_groups, d_bnd, p, H.n_b, d_walk, d_seen, V);
  } else {
    if (chain32)
      hipLaunchKernelGGL(pdlt_decode_kernel<float>, dim3((uint32_t)groups.size()), dim3(64),
                         lds_bytes, s, d_words, d_tab, d_groups, d_bnd, p, V, T);
    else
      hipLaunchKernelGGL(pdlt_decode_kernel<double>, dim3((uint32_t)groups.size()), dim3(64),
                         lds_bytes, s, d_words, d_tab, d_groups, d_bnd, p, V, T);
  }
  PDLT_HIP(hipGetLastError());
  return STRIDE_OK;
}

}  // namespace stride
