// Wave-level helpers of the chain token passes (DPR4 and PDLT v5): wave
// ordering, prefix scans and DPP moves over one wave64, the chain lanes'
// record sink and the resolution of their FIFO placeholders.  Device code
// only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "stride_lane.h"

namespace stride {

// Orders one wave's LDS hand-offs (a workgroup barrier would also wait for
// the wave's global stores in flight).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive prefix sum over the wave; total = the sum over all lanes.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t& total) {
  const int lane = threadIdx.x & 63;
  uint32_t s = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(s, o, 64);
    if (lane >= o) s += t;
  }
  total = __shfl(s, 63, 64);
  return s - x;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Cross-lane moves by DPP (GFX9 data-parallel primitives: no LDS traffic).
// update_dpp(old, src, ctrl, row_mask, bank_mask, bound_ctrl) returns src
// from the lane that ctrl names, or `old` where there is none.
//   0x111..0x118 row_shr:1..8, 0x142 row_bcast:15, 0x143 row_bcast:31,
//   0x138 wave_shr:1 (lane i takes lane i - 1, lane 0 takes `old`).
__device__ __forceinline__ int wave_shr1(int x, int lane0) {
  return __builtin_amdgcn_update_dpp(lane0, x, 0x138, 0xF, 0xF, false);
}
// Inclusive max over lanes 0..lane of values >= 0 (0 = none): lanes with no
// source read 0 (bound_ctrl), so each step is one max with a DPP operand.
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t x) {
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false));
  return x;
}

struct RecSink {
  uint16_t* __restrict__ r;
  __device__ __forceinline__ void operator()(uint32_t t, uint32_t rec) const { r[t] = (uint16_t)rec; }
};

template <bool kFind, class Lane>
__device__ __forceinline__ void chain_run(Lane& cl, LdsCursor& cur, RecSink& sink) {
  while (!cl.done()) cl.template step<kFind>(cur, sink);
}

// FIFO placeholders, four meshlets at a time (16 lanes each, one per rank):
// the true FIFO at the start of chain c + 1 replaces chain c's final FIFO
// (fb: 16 bytes per chain, newest first); its placeholders (n_local + r)
// stand for entries of the true FIFO at the start of chain c (chain 0's
// final FIFO holds local ids only).  Meshlet lane j holds n_local, its
// first chain in the wave (cfirst) and its chain count (nchain).
__device__ __forceinline__ void resolve_chain_fifos(uint8_t* __restrict__ fb, uint32_t nmesh,
                                                    uint32_t n_local, uint32_t cfirst,
                                                    uint32_t nchain) {
  const int lane = threadIdx.x & 63;
  for (uint32_t j0 = 0; j0 < nmesh; j0 += 4) {
    const uint32_t jm = j0 + ((uint32_t)lane >> 4);
    const uint32_t js = jm < nmesh ? jm : j0;
    const uint32_t mnl = __shfl(n_local, (int)js, 64);
    const uint32_t mcb = __shfl(cfirst, (int)js, 64);
    const uint32_t mmc = jm < nmesh ? __shfl(nchain, (int)js, 64) : 0u;
    uint32_t cmax = 0;
    for (uint32_t q = j0; q < j0 + 4 && q < nmesh; ++q)
      cmax = max(cmax, (uint32_t)__builtin_amdgcn_readlane((int)nchain, (int)q));
    const uint32_t rk = (uint32_t)lane & 15u;
    for (uint32_t c = 1; c + 1 < cmax; ++c) {
      if (c + 1 < mmc) {
        const uint32_t x = fb[16 * (mcb + c) + rk];
        fb[16 * (mcb + c) + rk] =
            x >= mnl ? fb[16 * (mcb + c - 1) + ((x - mnl) & 15u)] : (uint8_t)x;
      }
      wave_sync();
    }
  }
}

}  // namespace stride
