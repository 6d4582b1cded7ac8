// DPRB v4 side-table blob (the format scripts/dump_stride_dup_blob.py:88-152
// writes for the reference's native decode bench): the DPR4 v3 meshlet
// body plus per-meshlet tables, so a consumer needs no index pre-pass.
//
// Layout, little-endian:
//   u32 magic 0x42525044 ('DPRB'), u32 version 4
//   u32 n_meshlets, u32 n_v_total, u32 n_t_total, u32 body_bytes, u32 0, u32 0
//   f32 center[3], scale, per_coord_err, g_min[3], g_range[3]
//   u32 g_bits[3], i16 pred_n[9], u8 pred_K[3], 3 zero bytes      (112 bytes)
//   body: the DPR4 bytes from the first meshlet on
//   u64 bit offset of each meshlet in the body
//   u32 n_local[], n_tris[], n_strips[], v_off[], t_off[]
//   u64 anchor bit offset of each meshlet in the body
//   u32 n_kind0[]
//   u16 starts of the dy, dz, px, py, pz substreams relative to the anchor,
//       5 per meshlet (stored modulo 2^16, as the reference's uint16 array)
//
// The tables come from the index pre-pass restated on the host
// (scan_meshlet), the same parse the decode kernels' pre-pass runs.
#include "stride_host.h"

#include <cstring>

#include "stride_lane.h"

namespace stride {

namespace {

uint32_t rd_be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
template <class T>
void put_le(std::vector<uint8_t>& o, T v) {
  uint8_t b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));   // the host is little-endian (x86-64)
  o.insert(o.end(), b, b + sizeof(T));
}
template <class T>
T get_le(const uint8_t* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

constexpr uint32_t kDprbMagic = 0x42525044u;
constexpr uint32_t kDprbVersion = 4u;
constexpr size_t kDprbHeader = 112;

}  // namespace

int dprb_from_dpr4(const uint8_t* d, size_t len, std::vector<uint8_t>& out, std::string& err,
                   int& code) {
  code = -13;
  if (len < 5) { code = -12; err = "stream shorter than the DPR4 header"; return code; }
  if (rd_be32(d) != kMagicDup) { code = -10; err = "bad magic"; return code; }
  if (d[4] != kVersionDup) { code = -11; err = "expected v3"; return code; }
  if (len < 86) { code = -12; err = "stream shorter than the DPR4 header"; return code; }
  const uint32_t n_m = rd_be32(d + 82);
  if ((uint64_t)len < 86ull + 4ull * n_m) {
    code = -12;
    err = "stream shorter than its meshlet offset table";
    return code;
  }
  int32_t g_bits[3];
  for (int i = 0; i < 3; ++i) {
    g_bits[i] = d[50 + i];
    if (g_bits[i] < 1 || g_bits[i] > 30) { code = -14; err = "g_bits outside [1, 30]"; return code; }
  }
  const uint64_t body_byte = 86ull + 4ull * n_m;
  const uint64_t body_bit = 8 * body_byte, end_all = 8ull * len;
  std::vector<uint32_t> words((len + 3) / 4 + 16, 0);
  std::memcpy(words.data(), d, len);
  GlobalWords src{words.data(), (uint32_t)words.size()};
  std::vector<uint64_t> off(n_m);
  std::vector<MeshletScan> sc(n_m);
  for (uint32_t m = 0; m < n_m; ++m) off[m] = rd_be32(d + 86 + 4 * m);
  for (uint32_t m = 0; m < n_m; ++m) {
    const uint64_t start = off[m], end = m + 1 < n_m ? off[m + 1] : end_all;
    if (start < body_bit || end > end_all || end < start || end > 0xFFFFFFFFull) {
      err = "meshlet offsets out of order or outside the stream";
      return code;
    }
    scan_meshlet(src, (uint32_t)start, (uint32_t)end, g_bits, sc[m]);
    if (sc[m].error || (m + 1 < n_m && sc[m].end_bit != end)) {
      err = "malformed meshlet " + std::to_string(m);
      return code;
    }
  }
  uint64_t n_v = 0, n_t = 0;
  for (uint32_t m = 0; m < n_m; ++m) {
    n_v += sc[m].n_local;
    n_t += sc[m].n_tris;
  }
  if (n_v != rd_be32(d + 74) || n_t != rd_be32(d + 78)) {
    err = "header vertex/triangle counts disagree with the meshlets";
    return code;
  }
  out.clear();
  const uint64_t body_bytes = len - body_byte;
  out.reserve(kDprbHeader + body_bytes + (size_t)n_m * 58);
  put_le<uint32_t>(out, kDprbMagic);
  put_le<uint32_t>(out, kDprbVersion);
  put_le<uint32_t>(out, n_m);
  put_le<uint32_t>(out, (uint32_t)n_v);
  put_le<uint32_t>(out, (uint32_t)n_t);
  put_le<uint32_t>(out, (uint32_t)body_bytes);
  put_le<uint32_t>(out, 0);
  put_le<uint32_t>(out, 0);
  // f32 fields: center[3], scale, pce, g_min[3], g_range[3] (BE in DPR4)
  for (int i = 0; i < 11; ++i) put_le<uint32_t>(out, rd_be32(d + 6 + 4 * i));
  for (int i = 0; i < 3; ++i) put_le<uint32_t>(out, (uint32_t)g_bits[i]);
  for (int i = 0; i < 9; ++i) put_le<uint16_t>(out, (uint16_t)((d[53 + 2 * i] << 8) | d[54 + 2 * i]));
  for (int i = 0; i < 3; ++i) out.push_back(d[71 + i]);
  for (int i = 0; i < 3; ++i) out.push_back(0);
  out.insert(out.end(), d + body_byte, d + len);
  for (uint32_t m = 0; m < n_m; ++m) put_le<uint64_t>(out, off[m] - body_bit);
  for (uint32_t m = 0; m < n_m; ++m) put_le<uint32_t>(out, sc[m].n_local);
  for (uint32_t m = 0; m < n_m; ++m) put_le<uint32_t>(out, sc[m].n_tris);
  for (uint32_t m = 0; m < n_m; ++m) put_le<uint32_t>(out, sc[m].n_strips);
  uint32_t vo = 0, to = 0;
  for (uint32_t m = 0; m < n_m; ++m) { put_le<uint32_t>(out, vo); vo += sc[m].n_local; }
  for (uint32_t m = 0; m < n_m; ++m) { put_le<uint32_t>(out, to); to += sc[m].n_tris; }
  for (uint32_t m = 0; m < n_m; ++m) put_le<uint64_t>(out, sc[m].resid_bit - body_bit);
  for (uint32_t m = 0; m < n_m; ++m) put_le<uint32_t>(out, sc[m].n_kind0);
  for (uint32_t m = 0; m < n_m; ++m)
    for (int j = 1; j < 6; ++j) put_le<uint16_t>(out, (uint16_t)(sc[m].sub_bit[j] - sc[m].resid_bit));
  code = 0;
  return 0;
}

int dprb_to_dpr4(const uint8_t* b, size_t len, std::vector<uint8_t>& dpr4, DprbTables& t,
                 std::string& err, int& code) {
  code = -13;
  if (len < kDprbHeader) { code = -12; err = "blob shorter than the DPRB header"; return code; }
  if (get_le<uint32_t>(b) != kDprbMagic) { code = -10; err = "bad DPRB magic"; return code; }
  if (get_le<uint32_t>(b + 4) != kDprbVersion) { code = -11; err = "expected DPRB v4"; return code; }
  t.n_m = get_le<uint32_t>(b + 8);
  t.n_v = get_le<uint32_t>(b + 12);
  t.n_t = get_le<uint32_t>(b + 16);
  const uint64_t body_bytes = get_le<uint32_t>(b + 20);
  const uint64_t n_m = t.n_m;
  const uint64_t need = kDprbHeader + body_bytes + n_m * (8 + 5 * 4 + 8 + 4 + 10);
  if ((uint64_t)len < need) { code = -12; err = "blob shorter than its tables"; return code; }
  if (len != need) { err = "blob longer than its tables"; return code; }
  // DPR4 header (the code-width byte is not carried by DPRB; it is written as 0)
  dpr4.assign(86 + 4 * n_m, 0);
  auto be = [&](size_t o, uint32_t v) {
    dpr4[o] = (uint8_t)(v >> 24); dpr4[o + 1] = (uint8_t)(v >> 16);
    dpr4[o + 2] = (uint8_t)(v >> 8); dpr4[o + 3] = (uint8_t)v;
  };
  be(0, kMagicDup);
  dpr4[4] = (uint8_t)kVersionDup;
  dpr4[5] = 0;
  for (int i = 0; i < 11; ++i) be(6 + 4 * i, get_le<uint32_t>(b + 32 + 4 * i));
  for (int i = 0; i < 3; ++i) {
    const uint32_t g = get_le<uint32_t>(b + 76 + 4 * i);
    if (g > 255) { err = "g_bits out of range"; return code; }
    dpr4[50 + i] = (uint8_t)g;
  }
  for (int i = 0; i < 9; ++i) {
    const uint16_t u = get_le<uint16_t>(b + 88 + 2 * i);
    dpr4[53 + 2 * i] = (uint8_t)(u >> 8);
    dpr4[54 + 2 * i] = (uint8_t)u;
  }
  for (int i = 0; i < 3; ++i) dpr4[71 + i] = b[106 + i];
  be(74, t.n_v);
  be(78, t.n_t);
  be(82, t.n_m);
  const uint64_t body_bit = 8ull * (86 + 4 * n_m);
  const uint8_t* p = b + kDprbHeader + body_bytes;
  t.rel_off.resize(n_m);
  t.resid_off.resize(n_m);
  for (auto* v : {&t.n_local, &t.n_tris, &t.n_strips, &t.v_off, &t.t_off, &t.n_kind0}) v->resize(n_m);
  t.subs.resize(5 * n_m);
  for (uint64_t m = 0; m < n_m; ++m, p += 8) t.rel_off[m] = get_le<uint64_t>(p);
  for (auto* v : {&t.n_local, &t.n_tris, &t.n_strips, &t.v_off, &t.t_off})
    for (uint64_t m = 0; m < n_m; ++m, p += 4) (*v)[m] = get_le<uint32_t>(p);
  for (uint64_t m = 0; m < n_m; ++m, p += 8) t.resid_off[m] = get_le<uint64_t>(p);
  for (uint64_t m = 0; m < n_m; ++m, p += 4) t.n_kind0[m] = get_le<uint32_t>(p);
  for (uint64_t i = 0; i < 5 * n_m; ++i, p += 2) t.subs[i] = get_le<uint16_t>(p);
  for (uint64_t m = 0; m < n_m; ++m) {
    const uint64_t a = body_bit + t.rel_off[m];
    if (t.rel_off[m] >= 8 * body_bytes || a > 0xFFFFFFFFull) { err = "meshlet offset outside the body"; return code; }
    be(86 + 4 * m, (uint32_t)a);
  }
  t.body_start_bit = body_bit;
  dpr4.insert(dpr4.end(), b + kDprbHeader, b + kDprbHeader + body_bytes);
  code = 0;
  return 0;
}

}  // namespace stride
