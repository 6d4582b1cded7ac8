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

#include "stride_bits.h"
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
  MeshletScan s = {};
  const uint32_t start = e.start_bit, end = e.end_bit;
  const uint32_t limit = end - start;
  GlobalWords src{reinterpret_cast<const uint32_t*>(buf + P.blob_byte), (end + 31u) / 32u + 2u};
  BitReader<GlobalWords> r;
  r.init(src, start);
  auto over = [&]() { return r.pos > end; };
  uint32_t nl = r.bits(16), nt = r.bits(16), ns = r.bits(16);
  s.n_local = nl;
  s.n_tris = nt;
  s.n_strips = ns;
  if (nl < 1 || nl > (uint32_t)kMaxLocal || nt < 1 || nt > (uint32_t)kMaxTris || ns < 1 ||
      ns > nt) {
    s.error = kScanErrSize;
    out[m] = s;
    return;
  }
  int ib = 1;
  while ((1u << ib) < nl + 1) ++ib;
  Fifo fifo;
  fifo.clear();
  Seen256 seen;
  seen.clear();
  uint32_t walk = 0, kind0 = 0, tri_total = 0;
  uint32_t err = kScanOk;
  auto token = [&](uint32_t& v) -> bool {
    uint32_t flag = r.bits(1);
    if (flag == 0) {
      uint32_t fi = r.bits(5);
      if (fi >= fifo.n) { err = kScanErrFifo; return false; }
      v = fifo.hit(fi);
    } else {
      v = r.bits(ib);
      if (v >= nl) { err = kScanErrIndex; return false; }
      fifo.raw(v);
    }
    return true;
  };
  for (uint32_t st = 0; st < ns && err == kScanOk; ++st) {
    uint32_t L = r.bits(16);
    tri_total += L;
    if (L < 1 || tri_total > nt) { err = kScanErrStrips; break; }
    uint32_t p[3] = {0, 0, 0};
    for (int j = 0; j < 3 && err == kScanOk; ++j) {
      if (!token(p[j])) break;
      if (!seen.test_set(p[j])) { walk++; kind0++; }
    }
    for (uint32_t t = 1; t < L && err == kScanOk; ++t) {
      uint32_t ec = r.bits(1);
      uint32_t v;
      if (!token(v)) break;
      uint32_t a = ec ? p[0] : p[1], b = p[2], c = ec ? p[1] : p[0];
      if (ec == 0) p[0] = p[1];
      p[1] = p[2];
      p[2] = v;
      if (!seen.test_set(v)) {
        walk++;
        // the decode kernel marks strip roots with refs (255, 255, 255)
        if (a == 255 && b == 255 && c == 255) err = kScanErrWalk;
      }
      if (over()) err = kScanErrOverrun;
    }
    if (over()) err = kScanErrOverrun;
  }
  if (err == kScanOk && tri_total != nt) err = kScanErrStrips;
  if (err == kScanOk && walk != nl) err = kScanErrWalk;
  if (err != kScanOk) {
    s.error = err;
    out[m] = s;
    return;
  }
  s.n_kind0 = kind0;
  s.resid_bit = r.pos;
  r.bits(P.g_bits[0]);
  r.bits(P.g_bits[1]);
  r.bits(P.g_bits[2]);
  uint32_t tag[6], kk[6];
  for (int j = 0; j < 6; ++j) {
    tag[j] = r.bits(8);
    kk[j] = r.bits(8);
  }
  const uint32_t n_delta = kind0 - 1, n_para = nl - kind0;
  for (int j = 0; j < 6 && err == kScanOk; ++j) {
    s.sub_bit[j] = r.pos;
    uint32_t cnt = j < 3 ? n_delta : n_para;
    if (cnt == 0) continue;
    if (kk[j] > 24) { err = kScanErrTag; break; }
    for (uint32_t i = 0; i < cnt; ++i) {
      uint64_t u = tag[j] == 1 ? r.rice((int)kk[j], limit) : r.exp_golomb((int)kk[j], limit);
      if (over() || (u >> 32) != 0) { err = kScanErrOverrun; break; }
    }
  }
  if (err == kScanOk && over()) err = kScanErrOverrun;
  s.end_bit = r.pos;
  s.error = err;
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

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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
  const uint32_t kRoot = 0xFFFFFF00u;  // a = b = c = 255 never occurs for a strip step
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
    if (!seen.test_set(p0)) L.walk[walk++ & (kMaxLocal - 1)] = kRoot | p0;
    if (!seen.test_set(p1)) L.walk[walk++ & (kMaxLocal - 1)] = kRoot | p1;
    if (!seen.test_set(p2)) L.walk[walk++ & (kMaxLocal - 1)] = kRoot | p2;
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
  BitReader<Src> r;
  r.init(src, e.sub_bit[s]);
  if (tag == 1) {
    for (uint32_t i = 0; i < cnt; ++i) dst[i] = (int32_t)unzigzag(r.rice(k, limit));
  } else {
    for (uint32_t i = 0; i < cnt; ++i) dst[i] = (int32_t)unzigzag(r.exp_golomb(k, limit));
  }
}

__device__ __forceinline__ int32_t irlp(int32_t a, int32_t b, int32_t c, int32_t n0, int32_t n1,
                                        int32_t n2, int32_t K) {
  int64_t s = (int64_t)n0 * a + (int64_t)n1 * b + (int64_t)n2 * c;
  if (K != 0) s = (s + ((int64_t)1 << (K - 1))) >> K;
  return (int32_t)s;
}

// Phase C (lane 0): walk-order reconstruction (paradelta_v5_dup.py:581-600).
template <int SW>
__device__ __forceinline__ void chain_phase(const MeshletEntry& e, const MeshParams& P,
                                            int32_t ax, int32_t ay, int32_t az,
                                            WaveLds<SW>& L) {
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
    if ((pk >> 8) == 0xFFFFFFu) {
      if (first) {
        cx = ax; cy = ay; cz = az;
        first = false;
      } else {
        cx = px + rx[cd]; cy = py + ry[cd]; cz = pz + rz[cd];
        ++cd;
      }
    } else {
      int4 A = L.codes[(pk >> 8) & 0xFFu];
      int4 B = L.codes[(pk >> 16) & 0xFFu];
      int4 C = L.codes[pk >> 24];
      uint32_t slot = (n_delta + cp) & (kMaxLocal - 1);
      cx = irlp(A.x, B.x, C.x, P.pred_n[0], P.pred_n[1], P.pred_n[2], P.pred_K[0]) + rx[slot];
      cy = irlp(A.y, B.y, C.y, P.pred_n[3], P.pred_n[4], P.pred_n[5], P.pred_K[1]) + ry[slot];
      cz = irlp(A.z, B.z, C.z, P.pred_n[6], P.pred_n[7], P.pred_n[8], P.pred_K[2]) + rz[slot];
      ++cp;
    }
    L.codes[v] = make_int4(cx, cy, cz, 0);
    px = cx; py = cy; pz = cz;
  }
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

template <int WAVES, int SW>
__global__ __launch_bounds__(64 * WAVES) void stride_decode_kernel(
    const uint8_t* __restrict__ buf, const MeshletEntry* __restrict__ tab,
    const MeshParams* __restrict__ meshes, uint32_t n_m, float* __restrict__ V,
    uint32_t* __restrict__ T, uint32_t* __restrict__ queue, int use_queue) {
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
    const MeshletEntry e = tab[m];
    const MeshParams P = meshes[e.mesh];
    const uint32_t* blob = reinterpret_cast<const uint32_t*>(buf + P.blob_byte);
    const uint32_t w0 = e.start_bit >> 5;
    const uint32_t nw = ((e.end_bit + 31u) >> 5) + 2u - w0;
    const bool staged = nw <= (uint32_t)SW;
    if (staged) {
      for (uint32_t i = lane; i < nw; i += 64) L.stage[i] = __builtin_bswap32(blob[w0 + i]);
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
      if (lane == 0) conn_phase(src, el.start_bit, e.n_local, e.n_strips, L);
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
      if (lane == 0) conn_phase(src, e.start_bit, e.n_local, e.n_strips, L);
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
      chain_phase(e, P, L.anchor[0], L.anchor[1], L.anchor[2], L);
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

template <int WAVES, int SW>
static hipError_t launch_decode_t(const DecodeLaunch& d, hipStream_t stream) {
  const uint32_t grid = d.grid ? d.grid : (d.n_m + WAVES - 1) / WAVES;
  hipLaunchKernelGGL((stride_decode_kernel<WAVES, SW>), dim3(grid), dim3(64 * WAVES), 0, stream,
                     d.buf, d.tab, d.meshes, d.n_m, d.V, d.T, d.queue, d.use_queue);
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
