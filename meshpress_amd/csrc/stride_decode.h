// Host-side launch interface of the STRIDE decode kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "stride_device.h"

namespace stride {

struct DecodeConfig {
  int mode = 4;           // 0 = one wave per meshlet, 1 = groups (one lane per meshlet),
                          // 2 = split groups (token lane + three axis lanes per meshlet),
                          // 4 = two passes: token lanes (tok_group meshlets per wave,
                          //     walks to HBM), then the axis lanes of the split groups
  int waves = 1;          // mode 0: waves per workgroup (1, 2, 4); one meshlet per wave
  int stage_words = 512;  // mode 0: LDS words staged per meshlet (256, 512, 1024)
  int grid = 0;           // 0 = one workgroup per unit of work, else persistent grid
  int use_queue = 0;      // mode 0: persistent grid pulls meshlets from an atomic counter
  int group_size = 16;    // mode 1: meshlets per group (lanes), 1..64; mode 2: 1..16;
                          // mode 4: 1..21 (kAxisSlots)
  int lds_bytes = 16384;  // modes 1, 2, 4: LDS per group (stage words + codes / walk blocks)
  int tok_group = 32;     // mode 4: meshlets per token-pass wave, 1..64
  int tok_chains = 0;     // mode 4: token chains per meshlet (0 = one token lane per
                          //     meshlet, else 1..16: the chain token pass)
};

struct DecodeLaunch {
  const uint8_t* buf;
  const MeshletEntry* tab;
  const MeshParams* meshes;
  const uint32_t* idx;    // optional meshlet index list
  uint32_t* overflow;     // optional: set to 1 when a code exceeds 32 bits
  uint32_t n_m;
  float* V;
  uint32_t* T;
  uint32_t* queue;        // zeroed counter when use_queue
  int use_queue;
  uint32_t grid;
};

struct GroupLaunch {
  const uint8_t* buf;
  const MeshletEntry* tab;
  const MeshParams* meshes;
  const GroupEntry* groups;
  uint32_t n_groups;
  float* V;
  uint32_t* T;
  uint32_t* overflow;     // optional, one flag per group
  uint32_t lds_bytes;
  uint32_t grid;          // 0 = one workgroup per group
  int split;              // mode 2 kernel
  int narrow;             // mode 2: uint16 codes (every mesh has g_bits <= 16)
  uint32_t* queue;        // mode 2: zeroed work-queue counter (persistent grid), or null
  // mode 4 (two passes): token groups of up to 64 meshlets and the walks
  const GroupEntry* tgroups;
  uint32_t n_tgroups;
  uint32_t tlds_bytes;
  uint32_t* walk;         // one word per local vertex (mesh v_base + v_off)
  // token chains: the token groups are chain waves (GroupEntry pad[0..1] =
  // their chain range), k0 the chains per meshlet they were cut for; null =
  // one token lane per meshlet
  const ChainEntry* chains;
  uint32_t k0;
};

hipError_t launch_group(const GroupLaunch& d, hipStream_t stream);
hipError_t launch_scan(const uint8_t* buf, const MeshletEntry* seeds, const MeshParams* meshes,
                       uint32_t n_m, MeshletScan* out, hipStream_t stream);
hipError_t launch_decode(const DecodeLaunch& d, const DecodeConfig& c, hipStream_t stream);
hipError_t launch_chain_scan(const uint8_t* buf, const MeshletEntry* tab, const MeshParams* meshes,
                             const uint32_t* cbase, uint32_t n_m, uint32_t k0, ChainEntry* out,
                             hipStream_t stream);
size_t decode_lds_bytes(const DecodeConfig& c);

}  // namespace stride
