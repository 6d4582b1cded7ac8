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
#include "stride_wave.h"

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

// Boundary reconstructions of a whole group, from the refs the token pass
// wrote to HBM (refs_g, the group's meshlets back to back from the first
// one's ref_base): all 64 lanes take the group's 3 x refs coordinates in
// order, kStageDepth refs and then kStageDepth table reads in flight per
// lane, and store each into its meshlet's block.  Slot lane j < m1 - m0
// holds the entry of meshlet m0 + j.
template <class T>
__device__ __forceinline__ void pdlt_gather_group(uint32_t* lds, const PdltGroup& G,
                                                  const PdltEntry& e, int lane,
                                                  const uint32_t* __restrict__ refs_g,
                                                  const double* __restrict__ bnd,
                                                  const PdltParams& P) {
  const uint32_t nmesh = G.m1 - G.m0;
  const uint32_t rb = (uint32_t)__builtin_amdgcn_readlane((int)e.ref_base, 0);
  const uint32_t pre = e.ref_base - rb;   // the slot's first ref within the group
  const uint32_t nl = e.n_bnd + e.n_int;
  const uint32_t n3 =
      3u * (uint32_t)__builtin_amdgcn_readlane((int)(pre + e.n_bnd), (int)(nmesh - 1));
  for (uint32_t q0 = 0; q0 < n3; q0 += 64 * kStageDepth) {
    uint32_t ref[kStageDepth];
    T x[kStageDepth];
#pragma unroll
    for (uint32_t u = 0; u < kStageDepth; ++u) {
      const uint32_t q = q0 + 64 * u + (uint32_t)lane;
      ref[u] = refs_g[rb + (q < n3 ? q / 3 : 0u)];
    }
#pragma unroll
    for (uint32_t u = 0; u < kStageDepth; ++u) {
      const uint32_t q = q0 + 64 * u + (uint32_t)lane;
      const uint32_t d = q < n3 ? q - 3 * (q / 3) : 0u;
      x[u] = (T)pdlt_bnd(bnd[3ull * ref[u] + d], P);
    }
#pragma unroll
    for (uint32_t u = 0; u < kStageDepth; ++u) {
      const uint32_t q = q0 + 64 * u + (uint32_t)lane;
      const uint32_t r = q < n3 ? q / 3 : 0u, d = q - 3 * r;
      // the ref's meshlet: the last slot whose first ref is at or below r
      uint32_t j = 0;
      for (uint32_t k = 1; k < nmesh; ++k)
        j += r >= (uint32_t)__builtin_amdgcn_readlane((int)pre, (int)k) ? 1u : 0u;
      const uint32_t nlj = (uint32_t)__shfl((int)nl, (int)j, 64);
      const uint32_t boj = (uint32_t)__shfl((int)e.blk_off, (int)j, 64);
      const uint32_t prj = (uint32_t)__shfl((int)pre, (int)j, 64);
      if (q < n3) reinterpret_cast<T*>(lds + G.blk_base + boj)[d * nlj + (r - prj)] = x[u];
    }
  }
}

// Phases B-D of a group (after the walk, seen set, walk length and refs of
// every meshlet are in its LDS block): boundary gather, axis chains, output.
// T: the reconstruction type, double (float64 chain) or float (float32 chain,
// half the LDS per vertex); a block holds three T arrays of n_local, then
// refs, walk and seen set.  kRefsG: the boundary gather reads the refs from
// HBM (pdlt_gather_group) and the blocks' refs slot, if pdlt_chain_block_refs
// keeps one, is not used; else the four lanes of a slot gather from its
// block's refs.
template <class T, bool kRefsG = false>
__device__ __forceinline__ void pdlt_group_tail(const uint32_t* __restrict__ lds_c, uint32_t* lds,
                                                const PdltGroup& G, const PdltEntry& e, bool active,
                                                int role, int lane, int64_t rel,
                                                const PdltEntry* __restrict__ tab,
                                                const double* __restrict__ bnd, const PdltParams& P,
                                                float* __restrict__ V,
                                                const uint32_t* __restrict__ refs_g = nullptr) {
  constexpr uint32_t RW = 3 * sizeof(T) / 4;   // block words of reconstruction per local
  const uint32_t nl = e.n_bnd + e.n_int;
  uint32_t* blk = lds + G.blk_base + e.blk_off;
  T* recon = reinterpret_cast<T*>(blk);
  const uint32_t* refs = blk + RW * nl;
  constexpr bool kSlot = !kRefsG || pdlt_chain_block_refs(sizeof(T) == 4);
  const uint32_t* walk = refs + (kSlot ? e.n_bnd : 0u);
  const uint32_t* n_emit = walk + e.n_int + 8;
  if constexpr (kRefsG) pdlt_gather_group<T>(lds, G, e, lane, refs_g, bnd, P);
  // boundary reconstructions: the four lanes of a slot gather its refs
  if (!kRefsG && active) {
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
  // interior vertices, one meshlet at a time, 64 coordinates per step; the
  // meshlet's fields come from its slot lane (lane j), not a table read
  for (uint32_t j = 0; j < G.m1 - G.m0; ++j) {
    const uint32_t nbj = (uint32_t)__builtin_amdgcn_readlane((int)e.n_bnd, (int)j);
    const uint32_t nij = (uint32_t)__builtin_amdgcn_readlane((int)e.n_int, (int)j);
    const uint32_t boj = (uint32_t)__builtin_amdgcn_readlane((int)e.blk_off, (int)j);
    const uint32_t ibj = (uint32_t)__builtin_amdgcn_readlane((int)e.int_base, (int)j);
    const uint32_t nlj = nbj + nij;
    const uint32_t* bj = lds + G.blk_base + boj;
    const T* rj = reinterpret_cast<const T*>(bj);
    const uint32_t* sj = bj + RW * nlj + (kSlot ? nbj : 0u) + nij;
    float* vo = V + 3ull * ibj;
    const uint32_t n3 = 3 * nij;
    for (uint32_t t = lane; t < n3; t += 64) {
      const uint32_t i = t / 3, d = t - 3 * i, v = nbj + i;
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

// The token wave's refs (LDS words [blk_base, seen_base): its meshlets' refs
// back to back) to the refs buffer from its first meshlet's ref_base.
__device__ __forceinline__ void pdlt_refs_out(const uint32_t* __restrict__ lds, const PdltGroup& G,
                                              uint32_t rb, uint32_t* __restrict__ refs_g,
                                              int lane) {
  const uint32_t n = G.seen_base - G.blk_base;
  for (uint32_t i = (uint32_t)lane; i < n; i += 64) refs_g[rb + i] = lds[G.blk_base + i];
}

// Two-pass decode, pass 1: one token lane per meshlet, up to 64 meshlets
// per wave (refs in LDS for the global triangle ids; the lanes' seen sets
// interleaved in LDS, lane i's word q at seen_base + 64 q + i, so one access
// of the wave touches distinct banks); the walk, seen set and walk length go
// to HBM.
__global__ __launch_bounds__(64) void pdlt_tokens_kernel(
    const uint32_t* __restrict__ words, const PdltEntry* __restrict__ tab,
    const PdltGroup* __restrict__ groups, uint32_t n_b, uint32_t* __restrict__ walk_g,
    uint32_t* __restrict__ seen_g, uint32_t* __restrict__ refs_g, uint32_t* __restrict__ T) {
  extern __shared__ uint64_t lds64[];
  uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
  const int lane = threadIdx.x;
  const PdltGroup G = groups[blockIdx.x];
  const uint32_t rb = tab[G.m0].ref_base;   // read early: the refs go out last
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
  __syncthreads();
  pdlt_refs_out(lds, G, rb, refs_g, lane);
}

// Two-pass decode, pass 1 with token chains: a wave takes consecutive
// meshlets (up to 64 chains), whose bits from the first header to the end of
// the last token section are staged in LDS.  Chunk lanes decode the
// boundary refs into LDS, 16 refs a lane (RefChunk); phase 1: one lane per chain writes a 16-bit
// record per token (ChainLane); phase 2: the whole wave takes the meshlets
// one after another, resolves the FIFO placeholders, finds first sightings
// (an LDS atomic min of generation << 16 | token per local id) and forms
// every token's triangle (global ids: refs for boundary locals, int_base
// onward for interior ones) and walk entry from the records; an interior
// local's first sighting enters the walk.  The seen set (boundary locals
// and every local a token names) and the walk length go to HBM as the
// token lanes write them.  LDS: staged words | refs | records (each
// meshlet's padded to an even count) | final FIFO of each chain (16 bytes) |
// first sightings (256 words).
__global__ __launch_bounds__(64) void pdlt_chain_tokens_kernel(
    const uint32_t* __restrict__ words, const PdltEntry* __restrict__ tab,
    const PdltGroup* __restrict__ groups, const ChainEntry* __restrict__ chains,
    const RefChunk* __restrict__ rchunks, uint32_t k0, uint32_t n_b,
    uint32_t* __restrict__ walk_g, uint32_t* __restrict__ seen_g, uint32_t* __restrict__ refs_g,
    uint32_t* __restrict__ T) {
  extern __shared__ uint64_t lds64[];
  uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
  const int lane = threadIdx.x;
  const PdltGroup G = groups[blockIdx.x];
  const uint32_t nmesh = G.m1 - G.m0;
  const uint32_t c0 = G.pad[0], nch = G.pad[1] - G.pad[0];
  PdltEntry e{};
  if ((uint32_t)lane < nmesh) e = tab[G.m0 + (uint32_t)lane];
  ChainEntry ce{};
  if ((uint32_t)lane < nch) ce = chains[c0 + (uint32_t)lane];
  // the wave's boundary-ref chunks: from its first meshlet's first chunk to
  // its last meshlet's last
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)e.chunk_base, 0);
  const uint32_t nrc =
      (uint32_t)__builtin_amdgcn_readlane((int)(e.chunk_base + pdlt_ref_chunks(e.n_bnd)),
                                          (int)(nmesh - 1)) - r0;
  RefChunk rk{};
  if ((uint32_t)lane < nrc) rk = rchunks[r0 + (uint32_t)lane];
  stage_words(lds, words, G.w_lo, G.n_words, (uint32_t)lane, 64);
  const uint32_t n_local = e.n_bnd + e.n_int;
  const uint32_t ntok = (uint32_t)lane < nmesh ? pdlt_tokens_of(e) : 0u;
  const uint32_t lm = pdlt_chain_tokens(e, k0);
  const uint32_t nchain = chain_count(ntok, lm);
  uint32_t rtotal = 0;
  const uint32_t roff = wave_excl_scan((ntok + 1u) & ~1u, rtotal);
  uint32_t* refs = lds + G.blk_base;
  uint16_t* recs = reinterpret_cast<uint16_t*>(lds + G.seen_base);
  uint32_t* fst = lds + G.seen_base + rtotal / 2;   // 4 words per chain
  uint32_t* first = fst + 4 * 64;
  // the chain lane's meshlet: the count of chain-index-0 entries up to the
  // lane, less one (the wave's chains run meshlet by meshlet)
  const bool chain_lane = (uint32_t)lane < nch;
  const uint64_t starts = __ballot(chain_lane && ((ce.state >> 24) & 15u) == 0u);
  const uint32_t mi =
      chain_lane ? lanes_below(starts) - (((ce.state >> 24) & 15u) != 0u ? 1u : 0u) : 0u;
  const uint32_t c_roff = __shfl(roff, (int)mi, 64);
  const uint32_t c_bit = __shfl(e.strips_bit, (int)mi, 64);
  const uint32_t c_nl = __shfl(n_local, (int)mi, 64);
  const uint32_t c_sf = __shfl(e.n_strips, (int)mi, 64);
  const uint32_t c_ntok = __shfl(ntok, (int)mi, 64);
  const uint32_t c_lm = __shfl(lm, (int)mi, 64);
  // the chunk lane's meshlet, found as the chain lane's
  const bool chunk_lane = (uint32_t)lane < nrc;
  const uint32_t rci = rk.pos >> 16;
  const uint64_t rstarts = __ballot(chunk_lane && rci == 0u);
  const uint32_t ri = chunk_lane ? lanes_below(rstarts) - (rci != 0u ? 1u : 0u) : 0u;
  const uint32_t r_start = __shfl(e.start_bit, (int)ri, 64);
  const uint32_t r_nb = __shfl(e.n_bnd, (int)ri, 64);
  const uint32_t r_off = __shfl(e.tok_ref_off, (int)ri, 64);
  const int64_t rel = -32 * (int64_t)G.w_lo;
  __syncthreads();
  if (chunk_lane) pdlt_refs_chunk(lds, rel, r_start, r_nb, rk, refs + r_off);
  if (chain_lane) {
    LdsCursor cur{lds, (uint32_t)((int64_t)c_bit + rel) + (ce.pos & 0xFFFFu)};
    ChainLane<LdsCursor, true> cl;
    cl.init(cur, ce, chain_end(ce, c_ntok, c_lm), c_nl, c_sf);
    RecSink sink{recs + c_roff};
    if (wave_any(cl.raw_in_fifo)) chain_run<true>(cl, cur, sink);
    else chain_run<false>(cl, cur, sink);
    fst[4 * lane + 0] = (uint32_t)cl.fifo.lo;
    fst[4 * lane + 1] = (uint32_t)(cl.fifo.lo >> 32);
    fst[4 * lane + 2] = (uint32_t)cl.fifo.hi;
    fst[4 * lane + 3] = (uint32_t)(cl.fifo.hi >> 32);
  }
  __syncthreads();
  pdlt_refs_out(lds, G, (uint32_t)__builtin_amdgcn_readlane((int)e.ref_base, 0), refs_g, lane);
  uint8_t* fb = reinterpret_cast<uint8_t*>(fst);
  for (uint32_t i = (uint32_t)lane; i < 256; i += 64) first[i] = 0xFFFFFFFFu;
  uint32_t ctotal = 0;
  const uint32_t cfirst = wave_excl_scan((uint32_t)lane < nmesh ? nchain : 0u, ctotal);
  wave_sync();
  resolve_chain_fifos(fb, nmesh, n_local, cfirst, nchain);
  for (uint32_t j = 0; j < nmesh; ++j) {
    const uint32_t nl = (uint32_t)__builtin_amdgcn_readlane((int)n_local, (int)j);
    const uint32_t nb = (uint32_t)__builtin_amdgcn_readlane((int)e.n_bnd, (int)j);
    const uint32_t nt = (uint32_t)__builtin_amdgcn_readlane((int)ntok, (int)j);
    const uint32_t rb = (uint32_t)__builtin_amdgcn_readlane((int)roff, (int)j);
    const uint32_t ib = (uint32_t)__builtin_amdgcn_readlane((int)e.int_base, (int)j);
    const uint32_t toff = (uint32_t)__builtin_amdgcn_readlane((int)e.t_off, (int)j);
    const uint32_t ro = (uint32_t)__builtin_amdgcn_readlane((int)e.tok_ref_off, (int)j);
    const uint32_t cb = (uint32_t)__builtin_amdgcn_readlane((int)cfirst, (int)j);
    const uint32_t m = G.m0 + j;
    const uint32_t gen = (63u - j) << 16;
    uint16_t* rj = recs + rb;
    for (uint32_t t = lane; t < nt; t += 64) {
      const uint32_t r = rj[t];
      const uint32_t v = r & 0xFFu;
      const uint32_t ci = r >> 12;
      const uint32_t pv = fb[16 * (cb + (ci ? ci - 1 : 0u)) + ((v - nl) & 15u)];
      const uint32_t rv = v >= nl ? pv : v;   // placeholders only in chains c >= 1
      rj[t] = (uint16_t)((r & 0xFF00u) | rv);
      atomicMin(first + rv, gen | t);
    }
    wave_sync();
    // output ids: boundary locals through their refs, interior locals after
    // int_base (branch-free: an interior local reads the meshlet's first
    // ref slot, inside the LDS block, and drops it)
    const uint32_t* rf = refs + ro;
    const uint32_t ibase = ib - nb;
    auto gid = [&](uint32_t x) {
      const bool bx = x < nb;
      const uint32_t g = rf[bx ? x : 0u];
      return bx ? g : ibase + x;
    };
    uint32_t* wout = walk_g + pdlt_walk_off(ib, m, n_b);
    uint32_t* tout = T + 3ull * toff;
    uint32_t wbase = 0, tbase = 0;
    int cv1 = 0, cv2 = 0;
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < nt; t0 += 64) {
      const uint32_t t = t0 + (uint32_t)lane;
      const bool act = t < nt;
      const uint32_t rr = rj[act ? t : nt - 1];
      const uint32_t r = act ? rr : kRecRoot;
      const uint32_t v = r & 0xFFu;
      const bool fresh = act && v >= nb && first[v] == (gen | t);
      const int iv1 = wave_shr1((int)v, cv1);
      const int iv2 = wave_shr1(iv1, cv2);
      const uint32_t key =
          (act && (r & kRecEc) == 0) ? ((t << 8) | ((uint32_t)iv2 & 0xFFu)) + 1u : 0u;
      const uint32_t inc = wave_incl_max(key);
      const uint32_t s = max((uint32_t)wave_shr1((int)inc, 0), carry);
      cv1 = __builtin_amdgcn_readlane((int)v, 63);
      cv2 = __builtin_amdgcn_readlane(iv1, 63);
      carry = max(carry, (uint32_t)__builtin_amdgcn_readlane((int)inc, 63));
      const uint32_t v1 = (uint32_t)iv1 & 0xFFu;
      const uint32_t v2 = (uint32_t)iv2 & 0xFFu;
      const uint32_t p0 = (s - 1u) & 0xFFu;
      uint32_t a, b;
      const uint32_t entry = chain_entry(r, v1, v2, p0, a, b);
      const bool emit = act && chain_emits(r);
      const uint64_t bf = __ballot(fresh), be = __ballot(emit);
      const uint32_t widx = wbase + lanes_below(bf), tidx = tbase + lanes_below(be);
      wbase += (uint32_t)__popcll(bf);
      tbase += (uint32_t)__popcll(be);
      if (fresh) wout[widx] = entry;
      if (emit) {
        const uint32_t ga = gid(a), gb = gid(b), gv = gid(v);
        uint32_t* o = tout + (size_t)(3u * tidx);
        o[0] = ga;
        o[1] = gb;
        o[2] = gv;
      }
    }
    // the seen set: boundary locals and every local a token names
    uint32_t* so = seen_g + (uint64_t)kSeenWords * m;
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      const uint32_t i = 64u * q + (uint32_t)lane;
      const uint64_t sb = __ballot(i < nb || (first[i] >> 16) == 63u - j);
      if (lane == 0) {
        so[2 * q] = (uint32_t)sb;
        so[2 * q + 1] = (uint32_t)(sb >> 32);
      }
    }
    if (lane == 0) so[8] = wbase;
    wave_sync();
  }
}

// Two-pass decode, pass 2: the group's blocks take the walk, seen set and
// walk length from HBM and the refs decoded again, then phases B-D.
template <class R>
__global__ __launch_bounds__(64) void pdlt_chain_kernel(
    const uint32_t* __restrict__ words, const PdltEntry* __restrict__ tab,
    const PdltGroup* __restrict__ groups, const double* __restrict__ bnd, PdltParams P,
    uint32_t n_b, const uint32_t* __restrict__ walk_g, const uint32_t* __restrict__ seen_g,
    const uint32_t* __restrict__ refs_g, float* __restrict__ V) {
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
  // the block: reconstructions, the refs slot where pdlt_chain_block_refs
  // keeps one (the refs come from HBM), walk, seen set
  uint32_t* blk = lds + G.blk_base + e.blk_off;
  uint32_t* walk = blk + (3 * sizeof(R) / 4) * nl +
                   (pdlt_chain_block_refs(sizeof(R) == 4) ? e.n_bnd : 0u);
  uint32_t* seen = walk + e.n_int;
  const int64_t rel = -32 * (int64_t)G.w_lo;
  if (active) {
    // the slot's four lanes copy its walk, seen set and walk length
    copy_strided(walk, walk_g + pdlt_walk_off(e.int_base, m, n_b), e.n_int, (uint32_t)role, 4);
    copy_strided(seen, seen_g + (uint64_t)kSeenWords * m, kSeenWords, (uint32_t)role, 4);
  }
  pdlt_group_tail<R, true>(lds, lds, G, e, active, role, lane, rel, tab, bnd, P, V, refs_g);
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
  (void)hipFree(d_chains);
  (void)hipFree(d_rchunks);
  (void)hipFree(d_refs);
}

size_t PdltSet::side_bytes() const {
  return H.tab.size() * sizeof(PdltEntry) + groups.size() * sizeof(PdltGroup) +
         H.bnd.size() * sizeof(double) +
         (two_pass ? tok_groups.size() * sizeof(PdltGroup) : 0) +
         (two_pass && chain_waves ? n_chains * sizeof(ChainEntry) + n_rchunks * sizeof(RefChunk)
                                  : 0);
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
  PDLT_HIP(hipMalloc(&d_refs, 4 * std::max<uint64_t>(1, H.n_refs)));
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
  // (the chain pass of the two-pass decode reads the refs from HBM; its
  // blocks keep the refs slot where pdlt_chain_block_refs says so)
  int rc = pdlt_layout(tab, gs, budget, gr, lds, err, chain32 ? 3u : 6u,
                       two_pass == 0 || pdlt_chain_block_refs(chain32 != 0));
  // token chains where every meshlet fits a chain token wave, else the
  // token lanes
  std::vector<ChainEntry> ch;
  std::vector<RefChunk> rch;
  bool cw = false;
  if (!rc && tok_chains > 0) {
    rc = pdlt_layout_chains(tab, H.words, tok_chains, tok_group, tg, ch, rch, tlds, err);
    cw = rc == STRIDE_OK;
    if (rc == STRIDE_E_UNSUPPORTED) rc = STRIDE_OK;
  }
  // (token lanes in place of chains take 64 meshlets a wave)
  if (!rc && !cw)
    rc = pdlt_layout_tokens(tab, tok_budget, tok_chains > 0 ? 64 : tok_group, tg, tlds, err);
  if (rc) return set_error(rc, err);
  // the new device tables are built aside and take over only once every
  // allocation and copy has succeeded; until then the set is not ready
  ready = false;
  PdltGroup* ng = nullptr;
  PdltGroup* nt = nullptr;
  ChainEntry* nc = nullptr;
  RefChunk* nr = nullptr;
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
  if (e == hipSuccess && !ch.empty()) {
    e = hipMalloc(&nc, ch.size() * sizeof(ChainEntry));
    if (e == hipSuccess)
      e = hipMemcpy(nc, ch.data(), ch.size() * sizeof(ChainEntry), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !rch.empty()) {
    e = hipMalloc(&nr, rch.size() * sizeof(RefChunk));
    if (e == hipSuccess)
      e = hipMemcpy(nr, rch.data(), rch.size() * sizeof(RefChunk), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !tab.empty())
    e = hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(PdltEntry), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(ng);
    (void)hipFree(nt);
    (void)hipFree(nc);
    (void)hipFree(nr);
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
  (void)hipFree(d_chains);
  (void)hipFree(d_rchunks);
  d_groups = ng;
  d_tok_groups = nt;
  d_chains = nc;
  d_rchunks = nr;
  chain_waves = cw;
  n_chains = (uint32_t)ch.size();
  n_rchunks = (uint32_t)rch.size();
  layout_two_pass = two_pass;
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
  if (layout_two_pass != two_pass)
    return set_error(STRIDE_E_ARG, "the launch layout was made for the other decode mode");
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
    if (chain_waves)
      hipLaunchKernelGGL(pdlt_chain_tokens_kernel, dim3((uint32_t)tok_groups.size()), dim3(64),
                         tok_lds_bytes, s, d_words, d_tab, d_tok_groups, d_chains, d_rchunks,
                         (uint32_t)tok_chains, H.n_b, d_walk, d_seen, d_refs, T);
    else
      hipLaunchKernelGGL(pdlt_tokens_kernel, dim3((uint32_t)tok_groups.size()), dim3(64),
                         tok_lds_bytes, s, d_words, d_tab, d_tok_groups, H.n_b, d_walk, d_seen,
                         d_refs, T);
    PDLT_HIP(hipGetLastError());
    if (chain32)
      hipLaunchKernelGGL(pdlt_chain_kernel<float>, dim3((uint32_t)groups.size()), dim3(64),
                         lds_bytes, s, d_words, d_tab, d_groups, d_bnd, p, H.n_b, d_walk, d_seen,
                         d_refs, V);
    else
      hipLaunchKernelGGL(pdlt_chain_kernel<double>, dim3((uint32_t)groups.size()), dim3(64),
                         lds_bytes, s, d_words, d_tab, d_groups, d_bnd, p, H.n_b, d_walk, d_seen,
                         d_refs, V);
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
