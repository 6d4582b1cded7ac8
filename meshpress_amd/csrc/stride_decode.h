// Host-side launch interface of the STRIDE decode kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "stride_device.h"

namespace stride {

struct DecodeConfig {
  int waves = 1;          // waves per workgroup (1, 2, 4); one meshlet per wave
  int stage_words = 512;  // LDS words staged per meshlet (256, 512, 1024)
  int grid = 0;           // 0 = one wave per meshlet, else persistent grid (workgroups)
  int use_queue = 0;      // persistent grid pulls meshlets from an atomic counter
};

struct DecodeLaunch {
  const uint8_t* buf;
  const MeshletEntry* tab;
  const MeshParams* meshes;
  uint32_t n_m;
  float* V;
  uint32_t* T;
  uint32_t* queue;        // zeroed counter when use_queue
  int use_queue;
  uint32_t grid;
};

hipError_t launch_scan(const uint8_t* buf, const MeshletEntry* seeds, const MeshParams* meshes,
                       uint32_t n_m, MeshletScan* out, hipStream_t stream);
hipError_t launch_decode(const DecodeLaunch& d, const DecodeConfig& c, hipStream_t stream);
size_t decode_lds_bytes(const DecodeConfig& c);

}  // namespace stride
