// STRIDE (DPR4 v3) decode kernels for gfx950.
//
//  * stride_scan_kernel   index pre-pass, one lane per meshlet: parses a
//    whole meshlet once to derive the side table the DPR4 stream does not
//    store (anchor offset, the six residual substream offsets, n_kind0) and
//    validates every field against the meshlet's bit range.
//  * stride_decode_kernel fused decode, one wave64 per meshlet: stage the
//    meshlet's bits in LDS with coalesced loads, walk the strips and the
//    reuse FIFO, decode the six Rice/Exp-Golomb substreams on six lanes,
//    run the integer IRLP reconstruction chain, then dequantise in FP64 and
//    write vertices and indices as coalesced dword streams.
//
// Semantics follow decode_dup (encoder/paradelta_v5_dup.py:471-618).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "stride_bits.h"
#include "stride_lane.h"
#include "stride_wave.h"
#include "stride_decode.h"
#include "stride_device.h"

namespace stride {

// ---------------------------------------------------------------------------
// Index pre-pass
// ---------------------------------------------------------------------------
__global__ void stride_scan_kernel(const uint8_t* __restrict__ buf,
                                   const MeshletEntry* __restrict__ seeds,
                                   const MeshParams* __restrict__ meshes, uint32_t n_m,
                                   MeshletScan* __restrict__ out) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_m) return;
  const MeshletEntry e = seeds[m];
  const MeshParams& P = meshes[e.mesh];
  GlobalWords src{reinterpret_cast<const uint32_t*>(buf + P.blob_byte), (e.end_bit + 31u) / 32u + 2u};
  MeshletScan s;
  scan_meshlet(src, e.start_bit, e.end_bit, P.g_bits, s);
  out[m] = s;
}

// ---------------------------------------------------------------------------
// Fused decode
// ---------------------------------------------------------------------------
template <int STAGE_WORDS>
struct WaveLds {
  uint32_t stage[STAGE_WORDS];
  uint32_t tris[kMaxTris];      // a | b << 8 | c << 16
  uint32_t walk[kMaxLocal];     // v | a << 8 | b << 16 | c << 24; refs 0xFFFFFF = root
  int32_t resid[3][kMaxLocal];  // per axis: delta slots, then para slots
  int4 codes[kMaxLocal];
  int32_t anchor[4];
};


// Phase A (lane 0): strips + FIFO -> local triangles and the walk.
template <class Src, int SW>
__device__ __forceinline__ uint32_t conn_phase(const Src& src, uint32_t bit, uint32_t n_local,
                                               uint32_t n_strips, WaveLds<SW>& L) {
  BitReader<Src> r;
  r.init(src, bit + 48);
  int ib = 1;
  while ((1u << ib) < n_local + 1) ++ib;
  Fifo fifo;
  fifo.clear();
  Seen256 seen;
  seen.clear();
  uint32_t walk = 0, tc = 0;
  // strip roots enter the walk as v * 0x01010101 (walk_root)
  auto token = [&]() -> uint32_t {
    uint32_t flag = r.bits(1);
    uint32_t v;
    if (flag == 0) {
      v = fifo.hit(r.bits(5) & 15u);
    } else {
      v = r.bits(ib) & 0xFFu;
      fifo.raw(v);
    }
    return v;
  };
  for (uint32_t s = 0; s < n_strips; ++s) {
    uint32_t len = r.bits(16);
    uint32_t p0 = token(), p1 = token(), p2 = token();
    L.tris[tc & (kMaxTris - 1)] = p0 | (p1 << 8) | (p2 << 16);
    ++tc;
    if (!seen.test_set(p0)) L.walk[walk++ & (kMaxLocal - 1)] = byte4(p0);
    if (!seen.test_set(p1)) L.walk[walk++ & (kMaxLocal - 1)] = byte4(p1);
    if (!seen.test_set(p2)) L.walk[walk++ & (kMaxLocal - 1)] = byte4(p2);
    for (uint32_t t = 1; t < len; ++t) {
      uint32_t ec = r.bits(1);
      uint32_t nv = token();
      uint32_t a, b, c;
      if (ec == 0) {
        a = p1; b = p2; c = p0;
        p0 = p1; p1 = p2; p2 = nv;
      } else {
        a = p0; b = p2; c = p1;
        p1 = p2; p2 = nv;
      }
      L.tris[tc & (kMaxTris - 1)] = a | (b << 8) | (nv << 16);
      ++tc;
      if (!seen.test_set(nv)) L.walk[walk++ & (kMaxLocal - 1)] = nv | (a << 8) | (b << 16) | (c << 24);
    }
  }
  return walk;
}

// Phase B (lanes 0..5): substream s -> residuals of axis s % 3.
template <class Src, int SW>
__device__ __forceinline__ void resid_phase(const Src& src, const MeshletEntry& e, int s,
                                            uint32_t gsum, WaveLds<SW>& L) {
  BitReader<Src> h;
  h.init(src, e.resid_bit + gsum + 16u * (uint32_t)s);
  const uint32_t tag = h.bits(8);
  const int k = (int)(h.bits(8) & 31u);
  const uint32_t n_delta = e.n_kind0 - 1, n_para = e.n_local - e.n_kind0;
  const uint32_t cnt = s < 3 ? n_delta : n_para;
  int32_t* dst = L.resid[s % 3] + (s < 3 ? 0 : n_delta);
  const uint32_t limit = e.end_bit - e.start_bit;
  // select the substream start without indexing a private array by lane
  const uint32_t sb = s == 0 ? e.sub_bit[0] : s == 1 ? e.sub_bit[1] : s == 2 ? e.sub_bit[2]
                    : s == 3 ? e.sub_bit[3] : s == 4 ? e.sub_bit[4] : e.sub_bit[5];
  BitReader<Src> r;
  r.init(src, sb);
  if (tag == 1) {
    for (uint32_t i = 0; i < cnt; ++i) dst[i] = (int32_t)unzigzag(r.rice(k, limit));
  } else {
    for (uint32_t i = 0; i < cnt; ++i) dst[i] = (int32_t)unzigzag(r.exp_golomb(k, limit));
  }
}

// irlp64 / fits32: stride_lane.h

// Phase C (lane 0): walk-order reconstruction (paradelta_v5_dup.py:581-600).
template <int SW>
__device__ __forceinline__ bool chain_phase(const MeshletEntry& e, const MeshParams& P,
                                            int32_t ax, int32_t ay, int32_t az,
                                            WaveLds<SW>& L) {
  bool ok = true;
  const uint32_t n_delta = e.n_kind0 - 1;
  int32_t px = 0, py = 0, pz = 0;
  uint32_t cd = 0, cp = 0;
  bool first = true;
  const int32_t* rx = L.resid[0];
  const int32_t* ry = L.resid[1];
  const int32_t* rz = L.resid[2];
  for (uint32_t i = 0; i < e.n_local; ++i) {
    uint32_t pk = L.walk[i];
    uint32_t v = pk & 0xFFu;
    int32_t cx, cy, cz;
    if (walk_root(pk)) {
      if (first) {
        cx = ax; cy = ay; cz = az;
        first = false;
      } else {
        int64_t x = (int64_t)px + rx[cd], y = (int64_t)py + ry[cd], z = (int64_t)pz + rz[cd];
        ok &= fits32(x) & fits32(y) & fits32(z);
        cx = (int32_t)x; cy = (int32_t)y; cz = (int32_t)z;
        ++cd;
      }
    } else {
      int4 A = L.codes[(pk >> 8) & 0xFFu];
      int4 B = L.codes[(pk >> 16) & 0xFFu];
      int4 C = L.codes[pk >> 24];
      uint32_t slot = (n_delta + cp) & (kMaxLocal - 1);
      int64_t x = irlp64(A.x, B.x, C.x, P.pred_n[0], P.pred_n[1], P.pred_n[2], P.pred_K[0]) + rx[slot];
      int64_t y = irlp64(A.y, B.y, C.y, P.pred_n[3], P.pred_n[4], P.pred_n[5], P.pred_K[1]) + ry[slot];
      int64_t z = irlp64(A.z, B.z, C.z, P.pred_n[6], P.pred_n[7], P.pred_n[8], P.pred_K[2]) + rz[slot];
      ok &= fits32(x) & fits32(y) & fits32(z);
      cx = (int32_t)x; cy = (int32_t)y; cz = (int32_t)z;
      ++cp;
    }
    L.codes[v] = make_int4(cx, cy, cz, 0);
    px = cx; py = cy; pz = cz;
  }
  return ok;
}

// decode_dup dequantisation, in FP64 with the reference's operation order:
// ((g_min + (code / mx) * g_range) * scale + center) rounded to f32.
__device__ __forceinline__ float dequant(int32_t code, float gmin, float grange, int gbits,
                                         float scale, float center) {
  const uint32_t mx = (1u << gbits) - 1u;
  double vn;
  if (mx == 0) {
    vn = (double)gmin;
  } else {
    double q = __ddiv_rn((double)code, (double)mx);
    vn = __dadd_rn((double)gmin, __dmul_rn(q, (double)grange));
  }
  return (float)__dadd_rn(__dmul_rn(vn, (double)scale), (double)center);
}

// dequant for codes below 2^16 and g_bits <= 16, with r = RN(1 / mx) per
// axis: the quotient by quot_mx (equal to the division in that range), the
// rest as dequant.
__device__ __forceinline__ float dequant16(uint32_t code, float gmin, float grange, double mx,
                                           double r, float scale, float center) {
  const double vn = mx == 0.0 ? (double)gmin
                              : __dadd_rn((double)gmin, __dmul_rn(quot_mx((double)code, mx, r),
                                                                  (double)grange));
  return (float)__dadd_rn(__dmul_rn(vn, (double)scale), (double)center);
}

// Dequantisation of the split kernels' code arrays: the wave writes meshlet
// e's vertices (x, y, z) one per lane.  With uint16 codes, one reciprocal
// per axis and mesh (kept across a group's meshlets) replaces the division
// per coordinate.
struct SplitDequant {
  uint32_t mesh = ~0u;
  double mx[3] = {0.0, 0.0, 0.0}, r[3] = {0.0, 0.0, 0.0};

  // a meshlet of n local vertices at v_off in mesh mesh_id (parameters P)
  template <bool NARROW, class CodeT>
  __device__ __forceinline__ void meshlet(const CodeT* __restrict__ cb, uint32_t n, uint32_t v_off,
                                          uint32_t mesh_id, const MeshParams& P,
                                          float* __restrict__ V, int lane) {
    const uint32_t cs = split_code_stride(n, NARROW);
    float* vo = V + 3 * (P.v_base + v_off);
    if (NARROW && mesh_id != mesh) {
      mesh = mesh_id;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mx[a] = (double)((1u << P.g_bits[a]) - 1u);
        r[a] = mx[a] == 0.0 ? 0.0 : __ddiv_rn(1.0, mx[a]);
      }
    }
    for (uint32_t v = (uint32_t)lane; v < n; v += 64) {
      float x, y, z;
      if constexpr (NARROW) {
        x = dequant16(cb[v], P.g_min[0], P.g_range[0], mx[0], r[0], P.scale, P.center[0]);
        y = dequant16(cb[cs + v], P.g_min[1], P.g_range[1], mx[1], r[1], P.scale, P.center[1]);
        z = dequant16(cb[2 * cs + v], P.g_min[2], P.g_range[2], mx[2], r[2], P.scale,
                      P.center[2]);
      } else {
        x = dequant((int32_t)cb[v], P.g_min[0], P.g_range[0], P.g_bits[0], P.scale, P.center[0]);
        y = dequant((int32_t)cb[cs + v], P.g_min[1], P.g_range[1], P.g_bits[1], P.scale,
                    P.center[1]);
        z = dequant((int32_t)cb[2 * cs + v], P.g_min[2], P.g_range[2], P.g_bits[2], P.scale,
                    P.center[2]);
      }
      // one pointer for the vertex: the three stores merge into one
      // 12-byte store (32-bit index sums 3v, 3v + 1, 3v + 2 do not)
      float* p = vo + 3ull * v;
      p[0] = x;
      p[1] = y;
      p[2] = z;
    }
  }
};

template <int WAVES, int SW>
__global__ __launch_bounds__(64 * WAVES) void stride_decode_kernel(
    const uint8_t* __restrict__ buf, const MeshletEntry* __restrict__ tab,
    const MeshParams* __restrict__ meshes, const uint32_t* __restrict__ idx, uint32_t n_m,
    float* __restrict__ V, uint32_t* __restrict__ T, uint32_t* __restrict__ queue,
    int use_queue, uint32_t* __restrict__ overflow) {
  __shared__ WaveLds<SW> lds_all[WAVES];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  WaveLds<SW>& L = lds_all[wib];
  const uint32_t total_waves = gridDim.x * WAVES;
  uint32_t m = blockIdx.x * WAVES + wib;
  if (use_queue) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(queue, 1u);
    m = __builtin_amdgcn_readfirstlane(t);
  }
  while (m < n_m) {
    const MeshletEntry e = tab[idx ? idx[m] : m];
    const MeshParams P = meshes[e.mesh];
    const uint32_t* blob = reinterpret_cast<const uint32_t*>(buf + P.blob_byte);
    const uint32_t w0 = e.start_bit >> 5;
    const uint32_t nw = ((e.end_bit + 31u) >> 5) + 2u - w0;
    const bool staged = nw <= (uint32_t)SW;
    if (staged) {
      stage_words(L.stage, blob, w0, nw, lane, 64);
    }
    wave_sync();
    const uint32_t base = w0 << 5;
    const uint32_t gsum = (uint32_t)(P.g_bits[0] + P.g_bits[1] + P.g_bits[2]);
    if (staged) {
      LdsWords src{L.stage, nw};
      MeshletEntry el = e;
      el.start_bit -= base;
      el.resid_bit -= base;
      el.end_bit -= base;
#pragma unroll
      for (int j = 0; j < 6; ++j) el.sub_bit[j] -= base;
      if (lane == 0) conn_phase(src, el.start_bit, e.n_local, e.n_strips & kStripMask, L);
      if (lane < 6) resid_phase(src, el, lane, gsum, L);
      if (lane == 6) {
        BitReader<LdsWords> r;
        r.init(src, el.resid_bit);
        L.anchor[0] = (int32_t)r.bits(P.g_bits[0]);
        L.anchor[1] = (int32_t)r.bits(P.g_bits[1]);
        L.anchor[2] = (int32_t)r.bits(P.g_bits[2]);
      }
    } else {
      GlobalWords src{blob, ((e.end_bit + 31u) >> 5) + 2u};
      if (lane == 0) conn_phase(src, e.start_bit, e.n_local, e.n_strips & kStripMask, L);
      if (lane < 6) resid_phase(src, e, lane, gsum, L);
      if (lane == 6) {
        BitReader<GlobalWords> r;
        r.init(src, e.resid_bit);
        L.anchor[0] = (int32_t)r.bits(P.g_bits[0]);
        L.anchor[1] = (int32_t)r.bits(P.g_bits[1]);
        L.anchor[2] = (int32_t)r.bits(P.g_bits[2]);
      }
    }
    wave_sync();
    if (lane == 0) {
      if (!chain_phase(e, P, L.anchor[0], L.anchor[1], L.anchor[2], L) && overflow) *overflow = 1u;
    }
    wave_sync();
    // Phase D: coalesced dword streams of xyz floats and uint32 indices.
    {
      float* vo = V + 3 * (P.v_base + e.v_off);
      const uint32_t nv3 = 3 * e.n_local;
      for (uint32_t i = lane; i < nv3; i += 64) {
        const uint32_t vtx = i / 3, a = i - 3 * vtx;
        const int4 c = L.codes[vtx];
        const int32_t code = a == 0 ? c.x : (a == 1 ? c.y : c.z);
        const float gmin = a == 0 ? P.g_min[0] : (a == 1 ? P.g_min[1] : P.g_min[2]);
        const float grng = a == 0 ? P.g_range[0] : (a == 1 ? P.g_range[1] : P.g_range[2]);
        const int gb = a == 0 ? P.g_bits[0] : (a == 1 ? P.g_bits[1] : P.g_bits[2]);
        const float ce = a == 0 ? P.center[0] : (a == 1 ? P.center[1] : P.center[2]);
        vo[i] = dequant(code, gmin, grng, gb, P.scale, ce);
      }
      uint32_t* to = T + 3 * (P.t_base + e.t_off);
      const uint32_t nt3 = 3 * e.n_tris;
      for (uint32_t i = lane; i < nt3; i += 64) {
        const uint32_t tri = i / 3, a = i - 3 * tri;
        to[i] = e.v_off + ((L.tris[tri] >> (8 * a)) & 0xFFu);
      }
    }
    wave_sync();
    if (use_queue) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(queue, 1u);
      m = __builtin_amdgcn_readfirstlane(t);
    } else {
      m += total_waves;
    }
  }
}

// ---------------------------------------------------------------------------
// Group decode: one wave per group of consecutive meshlets, one lane per
// meshlet; the per-lane decoder is decode_lane (stride_lane.h).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(64) void stride_decode_group_kernel(
    const uint8_t* __restrict__ buf, const MeshletEntry* __restrict__ tab,
    const MeshParams* __restrict__ meshes, const GroupEntry* __restrict__ groups,
    uint32_t n_groups, float* __restrict__ V, uint32_t* __restrict__ T,
    uint32_t* __restrict__ overflow) {
  extern __shared__ uint64_t lds64[];
  uint32_t* stage = reinterpret_cast<uint32_t*>(lds64);
  const int lane = threadIdx.x;
  const uint32_t* words = reinterpret_cast<const uint32_t*>(buf);
  for (uint32_t g = blockIdx.x; g < n_groups; g += gridDim.x) {
    const GroupEntry G = groups[g];
    stage_words(stage, words, G.w_lo, G.n_words, lane, 64);
    uint64_t* codes = lds64 + G.codes_off;
    const uint32_t m = G.m0 + lane;
    const bool active = m < G.m1;
    MeshletEntry e;
    if (active) e = tab[m];
    uint32_t total = 0;
    const uint32_t coff = wave_excl_scan(active ? e.n_local : 0u, total);
    for (uint32_t i = lane; i < total; i += 64) codes[i] = kUnseen;
    __syncthreads();
    if (active) {
      const MeshParams& P = meshes[e.mesh];
      const uint64_t abs_bit = P.blob_byte * 8ull + e.start_bit;
      const uint32_t rel = (uint32_t)(abs_bit - 32ull * G.w_lo);
      uint32_t* tout = T + 3 * (P.t_base + e.t_off);
      if (!decode_lane(stage, G.n_words, rel, e, P, codes + coff, tout) && overflow)
        overflow[g] = 1u;
    }
    __syncthreads();
    // dequantise the group's vertices, one meshlet at a time, coalesced
    uint32_t cbase = 0;
    for (uint32_t j = G.m0; j < G.m1; ++j) {
      const MeshletEntry ej = tab[j];
      const MeshParams& Pj = meshes[ej.mesh];
      float* vo = V + 3 * (Pj.v_base + ej.v_off);
      const uint32_t nv3 = 3 * ej.n_local;
      for (uint32_t i = lane; i < nv3; i += 64) {
        const uint32_t vtx = i / 3, a = i - 3 * vtx;
        const int32_t code = field21(codes[cbase + vtx], (int)a);
        const float gmin = a == 0 ? Pj.g_min[0] : (a == 1 ? Pj.g_min[1] : Pj.g_min[2]);
        const float grng = a == 0 ? Pj.g_range[0] : (a == 1 ? Pj.g_range[1] : Pj.g_range[2]);
        const int gb = a == 0 ? Pj.g_bits[0] : (a == 1 ? Pj.g_bits[1] : Pj.g_bits[2]);
        const float ce = a == 0 ? Pj.center[0] : (a == 1 ? Pj.center[1] : Pj.center[2]);
        vo[i] = dequant(code, gmin, grng, gb, Pj.scale, ce);
      }
      cbase += ej.n_local;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Split decode: one wave per group of up to 16 consecutive meshlets; lane
// l works on meshlet l % 16 with role l / 16 (0: tokens, 1..3: axis x, y,
// z).  LDS: the group's staged words, then per meshlet a block with the
// walk (n_local words) and the x, y, z code arrays (int32, or uint16 when
// NARROW: every mesh of the set has at most 16 bits per axis).
// ---------------------------------------------------------------------------
template <bool NARROW>
__global__ __launch_bounds__(64) void stride_decode_split_kernel(
    const uint8_t* __restrict__ buf, const MeshletEntry* __restrict__ tab,
    const MeshParams* __restrict__ meshes, const GroupEntry* __restrict__ groups,
    uint32_t n_groups, float* __restrict__ V, uint32_t* __restrict__ T,
    uint32_t* __restrict__ overflow, uint32_t* __restrict__ queue) {
  using CodeT = typename std::conditional<NARROW, uint16_t, int32_t>::type;
  extern __shared__ uint32_t lds32[];
  const int lane = threadIdx.x;
  const int slot = lane & 15, role = lane >> 4;
  const uint32_t* words = reinterpret_cast<const uint32_t*>(buf);
  // groups come from the device work queue (a counter zeroed before the
  // launch) or, without one, in grid-stride order; every wave leaves once
  // the index passes n_groups
  auto next = [&](uint32_t cur) -> uint32_t {
    if (!queue) return cur + gridDim.x;
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(queue, 1u);
    return __builtin_amdgcn_readfirstlane(t);
  };
  for (uint32_t g = queue ? next(0) : blockIdx.x; g < n_groups; g = next(g)) {
    const GroupEntry G = groups[g];
    stage_words(lds32, words, G.w_lo, G.n_words, lane, 64);
    const uint32_t m = G.m0 + (uint32_t)slot;
    const bool active = m < G.m1;
    MeshletEntry e;
    if (active) e = tab[m];
    const uint32_t nl = active ? e.n_local : 0u;
    uint32_t total = 0;
    const uint32_t pre0 = wave_excl_scan(role == 0 ? split_block_words(nl, NARROW) : 0u, total);
    const uint32_t pre = __shfl(pre0, slot, 64);       // block offset of this lane's meshlet
    uint32_t* blk = lds32 + G.codes_off + pre;
    CodeT* codes0 = reinterpret_cast<CodeT*>(blk + nl);
    const uint32_t cstride = split_code_stride(nl, NARROW);
    // zero the blocks: the x-code array is the token lanes' seen set
    for (uint32_t i = lane; i < total; i += 64) lds32[G.codes_off + i] = 0u;
    __syncthreads();
    if (active && role == 0) {
      const MeshParams& P = meshes[e.mesh];
      const uint64_t abs_bit = P.blob_byte * 8ull + e.start_bit;
      const uint32_t rel = (uint32_t)(abs_bit - 32ull * G.w_lo);
      ScratchSeen<CodeT> seen{codes0};
      split_tokens(lds32, rel, e, blk, seen, T + 3 * (P.t_base + e.t_off));
    }
    __syncthreads();
    if (active && role > 0) {
      const MeshParams& P = meshes[e.mesh];
      const uint64_t abs_bit = P.blob_byte * 8ull + e.start_bit;
      const uint32_t rel = (uint32_t)(abs_bit - 32ull * G.w_lo);
      CodeT* codes = codes0 + cstride * (uint32_t)(role - 1);
      if (!split_axis(lds32, G.n_words, rel, e, P, role - 1, blk, codes) && overflow)
        overflow[g] = 1u;
    }
    __syncthreads();
    // dequantise, one meshlet at a time, one vertex (x, y, z) per lane; a
    // group lies in one mesh, and each meshlet's size and output offset come
    // from its slot lane
    SplitDequant dq;
    const uint32_t gmesh = (uint32_t)__shfl((int)(active ? e.mesh : 0u), 0, 64);
    const MeshParams& Pg = meshes[gmesh];
    const uint32_t voff = active ? e.v_off : 0u;
    for (uint32_t j = 0; j < G.m1 - G.m0; ++j) {
      const uint32_t pj = __shfl(pre0, (int)j, 64), nj = __shfl(nl, (int)j, 64);
      dq.meshlet<NARROW>(reinterpret_cast<const CodeT*>(lds32 + G.codes_off + pj + nj), nj,
                         __shfl(voff, (int)j, 64), gmesh, Pg, V, lane);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Two-pass split decode.  Pass 1: one token lane per meshlet, up to 64
// meshlets per wave, seen set in LDS after the staged words (an HBM-streamed
// cursor measured slower: its refills sit on the serial token chain); the
// walks go to HBM and the triangles to the output.  Pass 2: the split
// groups' axis lanes.  The wave stages only each meshlet's residual section
// (anchor, coder pairs, substreams: about half of its bits) in LDS after its
// code arrays, then each axis lane runs its chain from there (split_axis_g)
// with the walk read from HBM four entries ahead, then the dequantisation.
// With no token bits or walk in LDS, every meshlet of the mesh is resident
// at once; with the substreams in LDS, the only global loads of the chains
// are the walk reads, issued a group of four codes before their use.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void stride_tokens_kernel(
    const uint8_t* __restrict__ buf, const MeshletEntry* __restrict__ tab,
    const MeshParams* __restrict__ meshes, const GroupEntry* __restrict__ groups,
    uint32_t* __restrict__ walk_g, uint32_t* __restrict__ T) {
  extern __shared__ uint32_t lds32[];
  const int lane = threadIdx.x;
  const GroupEntry G = groups[blockIdx.x];
  const uint32_t* words = reinterpret_cast<const uint32_t*>(buf);
  // the lane's entry and the group's mesh (a token group lies in one mesh)
  // are read while the staging loads are in flight
  const uint32_t m = G.m0 + (uint32_t)lane;
  MeshletEntry e{};
  if (m < G.m1) e = tab[m];
  const MeshParams& P = meshes[tab[G.m0].mesh];
  stage_words(lds32, words, G.w_lo, G.n_words, lane, 64);
  __syncthreads();
  if (m < G.m1) {
    const uint64_t abs_bit = P.blob_byte * 8ull + e.start_bit;
    const uint32_t rel = (uint32_t)(abs_bit - 32ull * G.w_lo);
    // a constant stride: the word index scales by a shift, not a multiply
    BitSeen seen{lds32 + G.n_words + lane, kTokenSeenStride};
    seen.preset(0);
    split_tokens<WalkSlot>(lds32, rel, e, walk_g + walk_slot(P, e, m), seen,
                           T + 3 * (P.t_base + e.t_off));
  }
}

// ---------------------------------------------------------------------------
// Chain token pass (mode 4 with token chains).  The chain-start scan runs
// when the launch layout is built: one lane per meshlet re-parses its tokens
// and records the parser state at every chain cut (chain_starts).  A decode
// then runs one wave per token wave of consecutive meshlets of one mesh (up
// to 64 chains), whose bits are staged in LDS:
//   phase 1: one lane per chain writes a 16-bit record per token (ChainLane);
//   phase 2: the whole wave takes the meshlets one after another: it
//     resolves the FIFO placeholders, finds first sightings (an LDS atomic
//     min of the token index per local id), and forms every token's
//     triangle and walk entry from the records (chain_entry); prefix counts
//     over the wave (ballot + mbcnt) place the triangles, stored as one
//     contiguous run per 64 tokens, and the walk entries.
// LDS: staged words | records (each meshlet's padded to an even count) |
// the final FIFO of each chain (16 bytes) | first sightings (256 words).
// ---------------------------------------------------------------------------
__global__ void stride_chain_scan_kernel(const uint8_t* __restrict__ buf,
                                         const MeshletEntry* __restrict__ tab,
                                         const MeshParams* __restrict__ meshes,
                                         const uint32_t* __restrict__ cbase, uint32_t n_m,
                                         uint32_t k0, ChainEntry* __restrict__ out) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_m) return;
  const uint32_t cb = cbase[m];
  if (cb == 0xFFFFFFFFu) return;
  const MeshletEntry e = tab[m];
  const MeshParams& P = meshes[e.mesh];
  GlobalWords src{reinterpret_cast<const uint32_t*>(buf + P.blob_byte), (e.end_bit + 31u) / 32u + 2u};
  chain_starts(src, e.start_bit, e, k0, out + cb);
}


__global__ __launch_bounds__(64) void stride_chain_tokens_kernel(
    const uint8_t* __restrict__ buf, const MeshletEntry* __restrict__ tab,
    const MeshParams* __restrict__ meshes, const GroupEntry* __restrict__ groups,
    const ChainEntry* __restrict__ chains, uint32_t k0, uint32_t* __restrict__ walk_g,
    uint32_t* __restrict__ T) {
  extern __shared__ uint32_t lds32[];
  const int lane = threadIdx.x;
  const GroupEntry G = groups[blockIdx.x];
  const uint32_t* words = reinterpret_cast<const uint32_t*>(buf);
  const uint32_t nmesh = G.m1 - G.m0;
  const uint32_t c0 = G.pad[0], nch = G.pad[1] - G.pad[0];
  // meshlet lanes and chain lanes; their table reads overlap the staging
  MeshletEntry e{};
  if ((uint32_t)lane < nmesh) e = tab[G.m0 + (uint32_t)lane];
  ChainEntry ce{};
  if ((uint32_t)lane < nch) ce = chains[c0 + (uint32_t)lane];
  const MeshParams& P = meshes[tab[G.m0].mesh];
  stage_words(lds32, words, G.w_lo, G.n_words, lane, 64);
  const uint32_t ntok = (uint32_t)lane < nmesh ? meshlet_tokens(e) : 0u;
  // the meshlet's chain count, once per lane (phase 2 reads it per meshlet)
  const uint32_t nchain = chain_count(ntok, chain_tokens(ntok, e.n_local, e.n_strips, k0));
  uint32_t rtotal = 0;
  const uint32_t roff = wave_excl_scan((ntok + 1u) & ~1u, rtotal);
  uint16_t* recs = reinterpret_cast<uint16_t*>(lds32 + G.n_words);
  uint32_t* fst = lds32 + G.n_words + rtotal / 2;   // 4 words per chain
  uint32_t* first = fst + 4 * 64;
  // the chain lane's meshlet: the wave's chains run meshlet by meshlet, each
  // meshlet's from chain index 0, so it is the count of index-0 chains up to
  // the lane, less one; its fields come from its meshlet lane
  const bool chain_lane = (uint32_t)lane < nch;
  const uint64_t starts = __ballot(chain_lane && ((ce.state >> 24) & 15u) == 0u);
  const uint32_t mi =
      chain_lane ? lanes_below(starts) - (((ce.state >> 24) & 15u) != 0u ? 1u : 0u) : 0u;
  const uint32_t c_roff = __shfl(roff, (int)mi, 64);
  const uint32_t c_start = __shfl(e.start_bit, (int)mi, 64);
  const uint32_t c_nl = __shfl(e.n_local, (int)mi, 64);
  const uint32_t c_sf = __shfl(e.n_strips, (int)mi, 64);
  const uint32_t c_ntok = __shfl(ntok, (int)mi, 64);
  __syncthreads();
  if (chain_lane) {
    const uint64_t abs_bit = P.blob_byte * 8ull + c_start;
    LdsCursor cur{lds32, (uint32_t)(abs_bit - 32ull * G.w_lo) + (ce.pos & 0xFFFFu)};
    ChainLane<LdsCursor> cl;
    cl.init(cur, ce, chain_end(ce, c_ntok, chain_tokens(c_ntok, c_nl, c_sf, k0)), c_nl, c_sf);
    RecSink sink{recs + c_roff};
    if (wave_any(cl.raw_in_fifo)) chain_run<true>(cl, cur, sink);
    else chain_run<false>(cl, cur, sink);
    // the final FIFO, ranks 0..15 (newest first) as bytes
    fst[4 * lane + 0] = (uint32_t)cl.fifo.lo;
    fst[4 * lane + 1] = (uint32_t)(cl.fifo.lo >> 32);
    fst[4 * lane + 2] = (uint32_t)cl.fifo.hi;
    fst[4 * lane + 3] = (uint32_t)(cl.fifo.hi >> 32);
  }
  __syncthreads();
  uint8_t* fb = reinterpret_cast<uint8_t*>(fst);
  // first sightings: first[v] = min over the meshlet's tokens of
  // (gen << 16 | t); gen falls from meshlet to meshlet, so a meshlet's
  // values undercut every stale one and the array is cleared only once
  for (uint32_t i = (uint32_t)lane; i < 256; i += 64) first[i] = 0xFFFFFFFFu;
  // each meshlet's first chain in the wave
  uint32_t ctotal = 0;
  const uint32_t cfirst = wave_excl_scan((uint32_t)lane < nmesh ? nchain : 0u, ctotal);
  // phase 2 is one wave: its LDS hand-offs need ordering, not a workgroup
  // barrier (which would also wait for the global stores in flight)
  wave_sync();
  resolve_chain_fifos(fb, nmesh, e.n_local, cfirst, nchain);
  for (uint32_t j = 0; j < nmesh; ++j) {
    const uint32_t nl = (uint32_t)__builtin_amdgcn_readlane((int)e.n_local, (int)j);
    const uint32_t nt = (uint32_t)__builtin_amdgcn_readlane((int)ntok, (int)j);
    const uint32_t rb = (uint32_t)__builtin_amdgcn_readlane((int)roff, (int)j);
    const uint32_t voff = (uint32_t)__builtin_amdgcn_readlane((int)e.v_off, (int)j);
    const uint32_t toff = (uint32_t)__builtin_amdgcn_readlane((int)e.t_off, (int)j);
    const uint32_t cb = (uint32_t)__builtin_amdgcn_readlane((int)cfirst, (int)j);
    const uint32_t gen = (63u - j) << 16;
    uint16_t* rj = recs + rb;
    // placeholders -> local ids (every record rewritten, no branch); first
    // sighting of every local id
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
    uint32_t* wout = walk_g + (((P.v_base + voff + 3) & ~3ull) + 4ull * (G.m0 + j));
    uint32_t* tout = T + 3ull * (P.t_base + toff);
    uint32_t wbase = 0, tbase = 0;
    // carried from the chunk before: the ids of its last two tokens, and
    // 1 + (t << 8 | v_{t-2}) of its last edge-bit-0 token t (0: none)
    int cv1 = 0, cv2 = 0;
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < nt; t0 += 64) {
      const uint32_t t = t0 + (uint32_t)lane;
      const bool act = t < nt;
      const uint32_t rr = rj[act ? t : nt - 1];
      const uint32_t r = act ? rr : kRecRoot;
      const uint32_t v = r & 0xFFu;
      const bool fresh = act && first[v] == (gen | t);
      // v_{t-1}, v_{t-2} by wave shifts; p0 = v_{s-2} of the last edge-bit-0
      // token s before t by a max scan over 1 + (s << 8 | v_{s-2})
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
        uint32_t* o = tout + (size_t)(3u * tidx);
        o[0] = voff + a;
        o[1] = voff + b;
        o[2] = voff + v;
      }
    }
    wave_sync();
  }
}

template <bool NARROW, bool CHECK>
__global__ __launch_bounds__(64) void stride_axes_kernel(
    const uint8_t* __restrict__ buf, const MeshletEntry* __restrict__ tab,
    const MeshParams* __restrict__ meshes, const GroupEntry* __restrict__ groups,
    const uint32_t* __restrict__ walk_g, float* __restrict__ V, uint32_t* __restrict__ overflow) {
  using CodeT = typename std::conditional<NARROW, uint16_t, int32_t>::type;
  extern __shared__ uint32_t lds32[];
  const int lane = threadIdx.x;
  const int slot = lane / 3, d = lane - 3 * slot;
  const uint32_t g = blockIdx.x;
  const uint32_t* words = reinterpret_cast<const uint32_t*>(buf);
  const GroupEntry G = groups[g];
  const uint32_t m = G.m0 + (uint32_t)slot;
  const bool active = slot < kAxisSlots && m < G.m1;
  MeshletEntry e;
  if (active) e = tab[m];
  const uint32_t nl = active ? e.n_local : 0u;
  const MeshParams* Pm = meshes + (active ? e.mesh : 0u);
  const uint64_t bb = active ? Pm->blob_byte * 8ull : 0ull;
  // LDS per meshlet: its three code arrays, then its axis words (the walk
  // stays in HBM); the slot's values are read from its axis-0 lane 3s
  const uint64_t alo = active ? axis_stage_lo(bb, e) : 0ull;
  const uint32_t an = active ? axis_stage_words(bb, e) : 0u;
  const uint32_t cw = split_code_words(nl, NARROW);
  uint32_t total = 0;
  const uint32_t pre0 = wave_excl_scan(d == 0 ? cw + an : 0u, total);
  const uint32_t pre = __shfl(pre0, 3 * (lane < 63 ? slot : 0), 64);
  const uint32_t nslots = G.m1 - G.m0;
  uint32_t amax = 0;
  for (uint32_t j = 0; j < nslots; ++j)
    amax = max(amax, (uint32_t)__builtin_amdgcn_readlane((int)an, (int)(3 * j)));
  // stage the axis words of every meshlet of the group: all the loads of a
  // 64-word row first, then the stores (slots past the group load word 0)
  for (uint32_t b0 = 0; b0 < amax; b0 += 64) {
    const uint32_t i = b0 + (uint32_t)lane;
    uint32_t t[kAxisSlots];
#pragma unroll
    for (int j = 0; j < kAxisSlots; ++j) {
      const uint32_t nj = (uint32_t)__builtin_amdgcn_readlane((int)an, 3 * j);
      const uint64_t lj =
          ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(alo >> 32), 3 * j) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)alo, 3 * j);
      t[j] = words[i < nj ? lj + i : lj];
    }
#pragma unroll
    for (int j = 0; j < kAxisSlots; ++j) {
      const uint32_t nj = (uint32_t)__builtin_amdgcn_readlane((int)an, 3 * j);
      const uint32_t oj = (uint32_t)__builtin_amdgcn_readlane((int)(pre0 + cw), 3 * j);
      if (i < nj) lds32[oj + i] = __builtin_bswap32(t[j]);
    }
  }
  __syncthreads();
  CodeT* codes0 = reinterpret_cast<CodeT*>(lds32 + pre);
  const uint32_t cstride = split_code_stride(nl, NARROW);
  if (active) {
    const MeshParams& P = *Pm;
    CodeT* codes = codes0 + cstride * (uint32_t)d;
    const int32_t n0 = pick3(d, P.pred_n[0], P.pred_n[3], P.pred_n[6]);
    const int32_t n1 = pick3(d, P.pred_n[1], P.pred_n[4], P.pred_n[7]);
    const int32_t n2 = pick3(d, P.pred_n[2], P.pred_n[5], P.pred_n[8]);
    const int32_t K = pick3(d, P.pred_K[0], P.pred_K[1], P.pred_K[2]);
    const Walk4* wv = reinterpret_cast<const Walk4*>(walk_g + walk_slot(P, e, m));
    const LdsStaged src{lds32 + pre + cw, alo};
    // one loop per wave: 32-bit IRLP when it is exact for every lane, the
    // Rice-only codeword path when every lane's substreams are Rice coded
    const bool i32 = __all(NARROW && irlp32_exact(n0, n1, n2, K));
    const bool rice = __all(axis_rice(src, bb, e, P, d));
    bool ok;
    if (i32 && rice) ok = split_axis_g<true, true, CHECK>(src, bb, e, P, d, wv, codes);
    else if (i32) ok = split_axis_g<true, false, CHECK>(src, bb, e, P, d, wv, codes);
    else if (rice) ok = split_axis_g<false, true, CHECK>(src, bb, e, P, d, wv, codes);
    else ok = split_axis_g<false, false, CHECK>(src, bb, e, P, d, wv, codes);
    if (CHECK && !ok) overflow[g] = 1u;
  }
  __syncthreads();
  // a group lies in one mesh; each meshlet's size and output offset come
  // from its slot lanes, not from memory
  SplitDequant dq;
  const uint32_t gmesh = (uint32_t)__builtin_amdgcn_readlane((int)(active ? e.mesh : 0u), 0);
  const MeshParams& Pg = meshes[gmesh];
  const uint32_t voff = active ? e.v_off : 0u;
  for (uint32_t j = 0; j < nslots; ++j) {
    const uint32_t pj = (uint32_t)__builtin_amdgcn_readlane((int)pre0, (int)(3 * j));
    dq.meshlet<NARROW>(reinterpret_cast<const CodeT*>(lds32 + pj),
                       (uint32_t)__builtin_amdgcn_readlane((int)nl, (int)(3 * j)),
                       (uint32_t)__builtin_amdgcn_readlane((int)voff, (int)(3 * j)), gmesh, Pg, V,
                       lane);
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
hipError_t launch_scan(const uint8_t* buf, const MeshletEntry* seeds, const MeshParams* meshes,
                       uint32_t n_m, MeshletScan* out, hipStream_t stream) {
  if (n_m == 0) return hipSuccess;
  const uint32_t tpb = 64;
  hipLaunchKernelGGL(stride_scan_kernel, dim3((n_m + tpb - 1) / tpb), dim3(tpb), 0, stream, buf,
                     seeds, meshes, n_m, out);
  return hipGetLastError();
}

hipError_t launch_chain_scan(const uint8_t* buf, const MeshletEntry* tab, const MeshParams* meshes,
                             const uint32_t* cbase, uint32_t n_m, uint32_t k0, ChainEntry* out,
                             hipStream_t stream) {
  if (n_m == 0) return hipSuccess;
  const uint32_t tpb = 64;
  hipLaunchKernelGGL(stride_chain_scan_kernel, dim3((n_m + tpb - 1) / tpb), dim3(tpb), 0, stream,
                     buf, tab, meshes, cbase, n_m, k0, out);
  return hipGetLastError();
}

template <int WAVES, int SW>
static hipError_t launch_decode_t(const DecodeLaunch& d, hipStream_t stream) {
  const uint32_t grid = d.grid ? d.grid : (d.n_m + WAVES - 1) / WAVES;
  hipLaunchKernelGGL((stride_decode_kernel<WAVES, SW>), dim3(grid), dim3(64 * WAVES), 0, stream,
                     d.buf, d.tab, d.meshes, d.idx, d.n_m, d.V, d.T, d.queue, d.use_queue,
                     d.overflow);
  return hipGetLastError();
}

hipError_t launch_group(const GroupLaunch& d, hipStream_t stream) {
  if (d.n_groups == 0) return hipSuccess;
  if (d.walk) {   // mode 4: token pass, then the axis pass
    if (d.n_tgroups && d.chains) {
      hipLaunchKernelGGL(stride_chain_tokens_kernel, dim3(d.n_tgroups), dim3(64), d.tlds_bytes,
                         stream, d.buf, d.tab, d.meshes, d.tgroups, d.chains, d.k0, d.walk, d.T);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    } else if (d.n_tgroups) {
      hipLaunchKernelGGL(stride_tokens_kernel, dim3(d.n_tgroups), dim3(64), d.tlds_bytes, stream,
                         d.buf, d.tab, d.meshes, d.tgroups, d.walk, d.T);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    // range checks only where a caller collects them (the layout's validation decode)
    if (d.narrow && d.overflow)
      hipLaunchKernelGGL((stride_axes_kernel<true, true>), dim3(d.n_groups), dim3(64), d.lds_bytes,
                         stream, d.buf, d.tab, d.meshes, d.groups, d.walk, d.V, d.overflow);
    else if (d.narrow)
      hipLaunchKernelGGL((stride_axes_kernel<true, false>), dim3(d.n_groups), dim3(64), d.lds_bytes,
                         stream, d.buf, d.tab, d.meshes, d.groups, d.walk, d.V, d.overflow);
    else if (d.overflow)
      hipLaunchKernelGGL((stride_axes_kernel<false, true>), dim3(d.n_groups), dim3(64), d.lds_bytes,
                         stream, d.buf, d.tab, d.meshes, d.groups, d.walk, d.V, d.overflow);
    else
      hipLaunchKernelGGL((stride_axes_kernel<false, false>), dim3(d.n_groups), dim3(64), d.lds_bytes,
                         stream, d.buf, d.tab, d.meshes, d.groups, d.walk, d.V, d.overflow);
    return hipGetLastError();
  }
  const uint32_t grid = d.grid ? std::min(d.grid, d.n_groups) : d.n_groups;
  if (d.split && d.narrow)
    hipLaunchKernelGGL(stride_decode_split_kernel<true>, dim3(grid), dim3(64), d.lds_bytes, stream,
                       d.buf, d.tab, d.meshes, d.groups, d.n_groups, d.V, d.T, d.overflow,
                       d.queue);
  else if (d.split)
    hipLaunchKernelGGL(stride_decode_split_kernel<false>, dim3(grid), dim3(64), d.lds_bytes,
                       stream, d.buf, d.tab, d.meshes, d.groups, d.n_groups, d.V, d.T, d.overflow,
                       d.queue);
  else
    hipLaunchKernelGGL(stride_decode_group_kernel, dim3(grid), dim3(64), d.lds_bytes, stream,
                       d.buf, d.tab, d.meshes, d.groups, d.n_groups, d.V, d.T, d.overflow);
  return hipGetLastError();
}

hipError_t launch_decode(const DecodeLaunch& d, const DecodeConfig& c, hipStream_t stream) {
  if (d.n_m == 0) return hipSuccess;
#define STRIDE_CASE(W, S) \
  if (c.waves == W && c.stage_words == S) return launch_decode_t<W, S>(d, stream);
  STRIDE_CASE(1, 256)
  STRIDE_CASE(1, 512)
  STRIDE_CASE(1, 1024)
  STRIDE_CASE(2, 256)
  STRIDE_CASE(2, 512)
  STRIDE_CASE(2, 1024)
  STRIDE_CASE(4, 256)
  STRIDE_CASE(4, 512)
  STRIDE_CASE(4, 1024)
#undef STRIDE_CASE
  return hipErrorInvalidValue;
}

size_t decode_lds_bytes(const DecodeConfig& c) {
  return (size_t)c.waves * (sizeof(WaveLds<1>) - 4 + 4 * (size_t)c.stage_words);
}

}  // namespace stride
