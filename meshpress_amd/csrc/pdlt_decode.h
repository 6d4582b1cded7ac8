// PDLT v5 decode set: the device side of the gfx950 PDLT v5 decoder
// (pdlt_decode.hip), used by the C ABI glue (stride_capi.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "pdlt_host.h"

namespace stride {

struct PdltSet {
  PdltHost H;
  std::vector<PdltGroup> groups;        // chain (or single-pass) groups
  std::vector<PdltGroup> tok_groups;    // token groups of the two-pass decode
  uint32_t lds_bytes = 0;               // dynamic LDS of the group launch
  uint32_t tok_lds_bytes = 0;           // dynamic LDS of the token launch
  int group_size = 16;
  int lds_budget = 20480;
  int tok_lds_budget = 16384;
  int tok_group = 4;                    // meshlets per token-pass wave
  int tok_chains = 16;                  // > 0: token chains (about this many a meshlet)
  bool chain_waves = false;             // tok_groups are chain token waves (pdlt_layout_chains)
  uint32_t n_chains = 0;                // entries at d_chains
  int chain32 = 0;                      // 1: the encoder's float32 chain
  int two_pass = 1;                     // 1: token pass + chain pass, 0: one kernel
  int layout_two_pass = -1;             // two_pass when the layout was made (its blocks differ)
  // device
  uint32_t* d_words = nullptr;
  PdltEntry* d_tab = nullptr;
  PdltGroup* d_groups = nullptr;
  PdltGroup* d_tok_groups = nullptr;
  ChainEntry* d_chains = nullptr;       // parser state at every token chain's start
  RefChunk* d_rchunks = nullptr;        // boundary-ref chunks of the chain token pass
  uint32_t n_rchunks = 0;
  uint32_t* d_walk = nullptr;           // two-pass walks (n_v - n_b words)
  uint32_t* d_seen = nullptr;           // two-pass seen sets + walk lengths (kSeenWords per meshlet)
  uint32_t* d_refs = nullptr;           // two-pass boundary refs (n_refs words)
  double* d_bnd = nullptr;              // boundary table, normalised f64 (n_b x 3)
  float* d_v_own = nullptr;
  uint32_t* d_t_own = nullptr;
  // false while no launch layout is complete on the device (a failed
  // set_layout leaves it false and decode() refuses to launch)
  bool ready = false;

  ~PdltSet();
  // Parses, validates and uploads.  STRIDE_OK or a STRIDE_E_* status with
  // the message set.
  int build(const uint8_t* data, size_t len);
  int set_layout(int group_size, int lds_budget, int tok_lds_budget);
  // world = 1: world coordinates; 0: normalised.
  int decode(float* d_verts, uint32_t* d_tris, int world, hipStream_t stream);
  int own_outputs();
  size_t side_bytes() const;
};

}  // namespace stride
