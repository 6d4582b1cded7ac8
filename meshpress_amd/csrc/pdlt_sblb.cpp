// SBLB v1 blob (the input of the reference's native PDLT v5 decode bench,
// written by scripts/dump_stride_blob.py:1-114 and read by
// bench_cpp/stride_decode_bench.cu:430-484): the PDLT v5 stream plus the
// tables that bench uploads, so it needs no pre-pass.
//
// Layout, little-endian:
//   u32 magic 0x424C4253 ('SBLB'), u32 version 1, u32 code_width,
//   u32 n_meshlets, u32 n_boundary, u32 n_v_total, u32 n_t_total,
//   u32 buf_size, f32 lin3[3], f32 delta, u32 per_warp_bytes,
//   u32 meshlets_per_block (1), u32 0, u32 0                   (64 bytes)
//   buf            u8  x buf_size        the PDLT v5 stream
//   ml_off_bits    u64 x n_meshlets      absolute meshlet bit offsets
//   ml_n_bnd, ml_n_int, ml_n_tris, ml_n_strips   i32 x n_meshlets each
//   ml_tris_off    i32 x (n_meshlets + 1)   prefix sum of n_tris
//   ml_int_cursor  i32 x n_meshlets         exclusive prefix sum of n_int
//   bnd_pos_norm   f32 x 3 x n_boundary     float32 of the dequantised table
//
// The tables come from this package's pre-pass (pdlt_parse), the same parse
// the decoder runs; per_warp_bytes repeats the reference decoder's shared
// memory sizing (utils/paradelta_v5_cuda.py:652-671), so a blob written here
// also drives the reference bench.
#include "stride_host.h"

#include <cstring>

#include "pdlt_host.h"

namespace stride {

namespace {

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

constexpr uint32_t kSblbMagic = 0x424C4253u;
constexpr uint32_t kSblbVersion = 1u;
constexpr size_t kSblbHeader = 64;

uint32_t per_warp_bytes(const PdltHost& H) {
  uint64_t nl = 0, nb = 0, nt = 0, ni = 0;
  for (const PdltEntry& e : H.tab) {
    nl = std::max<uint64_t>(nl, e.n_bnd + e.n_int);
    nb = std::max<uint64_t>(nb, e.n_bnd);
    nt = std::max<uint64_t>(nt, e.n_tris);
    ni = std::max<uint64_t>(ni, e.n_int);
  }
  const uint64_t code_bytes = H.code_width ? 4 : 2;
  const uint64_t b = nl * 16 + nb * 4 + 32 + ni * 3 * code_bytes + ni * 8 + nt * 3;
  return (uint32_t)((b + 15) & ~15ull);
}

}  // namespace

int sblb_from_pdlt(const uint8_t* d, size_t len, std::vector<uint8_t>& out, std::string& err) {
  PdltHost H;
  const int rc = pdlt_parse(d, len, H, err);
  if (rc) return rc;
  if (H.n_v > 0x7FFFFFFFull || H.n_t > 0x7FFFFFFFull || len > 0xFFFFFFFFull) {
    err = "mesh too large for the SBLB 32-bit fields";
    return STRIDE_E_UNSUPPORTED;
  }
  const uint32_t n_m = H.n_m;
  out.clear();
  out.reserve(kSblbHeader + len + 36ull * n_m + 12ull * H.n_b + 4);
  put_le<uint32_t>(out, kSblbMagic);
  put_le<uint32_t>(out, kSblbVersion);
  put_le<uint32_t>(out, H.code_width);
  put_le<uint32_t>(out, n_m);
  put_le<uint32_t>(out, H.n_b);
  put_le<uint32_t>(out, (uint32_t)H.n_v);
  put_le<uint32_t>(out, (uint32_t)H.n_t);
  put_le<uint32_t>(out, (uint32_t)len);
  for (int i = 0; i < 3; ++i) put_le<float>(out, H.lin3[i]);
  put_le<float>(out, (float)(2.0 * (double)H.pce));
  put_le<uint32_t>(out, per_warp_bytes(H));
  put_le<uint32_t>(out, 1);
  put_le<uint32_t>(out, 0);
  put_le<uint32_t>(out, 0);
  out.insert(out.end(), d, d + len);
  for (const PdltEntry& e : H.tab) put_le<uint64_t>(out, e.start_bit);
  for (const PdltEntry& e : H.tab) put_le<int32_t>(out, (int32_t)e.n_bnd);
  for (const PdltEntry& e : H.tab) put_le<int32_t>(out, (int32_t)e.n_int);
  for (const PdltEntry& e : H.tab) put_le<int32_t>(out, (int32_t)e.n_tris);
  for (const PdltEntry& e : H.tab) put_le<int32_t>(out, (int32_t)(e.n_strips & kStripMask));
  for (const PdltEntry& e : H.tab) put_le<int32_t>(out, (int32_t)e.t_off);
  put_le<int32_t>(out, (int32_t)H.n_t);
  for (const PdltEntry& e : H.tab) put_le<int32_t>(out, (int32_t)(e.int_base - H.n_b));
  for (double x : H.bnd) put_le<float>(out, (float)x);
  return STRIDE_OK;
}

int sblb_split(const uint8_t* b, size_t len, const uint8_t*& stream, size_t& stream_len,
               std::string& err) {
  if (len < kSblbHeader) { err = "blob shorter than the SBLB header"; return STRIDE_E_TRUNCATED; }
  if (get_le<uint32_t>(b) != kSblbMagic) { err = "bad SBLB magic"; return STRIDE_E_MAGIC; }
  if (get_le<uint32_t>(b + 4) != kSblbVersion) { err = "expected SBLB v1"; return STRIDE_E_VERSION; }
  const uint64_t n_m = get_le<uint32_t>(b + 12), n_b = get_le<uint32_t>(b + 16);
  const uint64_t buf = get_le<uint32_t>(b + 28);
  const uint64_t need = kSblbHeader + buf + 8 * n_m + 4 * 4 * n_m + 4 * (n_m + 1) + 4 * n_m + 12 * n_b;
  if ((uint64_t)len < need) { err = "blob shorter than its tables"; return STRIDE_E_TRUNCATED; }
  if ((uint64_t)len != need) { err = "blob longer than its tables"; return STRIDE_E_CORRUPT; }
  stream = b + kSblbHeader;
  stream_len = (size_t)buf;
  return STRIDE_OK;
}

int sblb_check(const uint8_t* b, size_t len, std::string& err) {
  const uint8_t* s = nullptr;
  size_t sl = 0;
  int rc = sblb_split(b, len, s, sl, err);
  if (rc) return rc;
  std::vector<uint8_t> mine;
  rc = sblb_from_pdlt(s, sl, mine, err);
  if (rc) return rc;
  if (mine.size() != len) { err = "SBLB tables disagree with the stream"; return STRIDE_E_CORRUPT; }
  // the reserved words and meshlets_per_block are not checked: they do not
  // affect the decode
  for (size_t i = 0; i < len; ++i) {
    if (i >= 52 && i < 64) continue;
    if (mine[i] != b[i]) {
      err = "SBLB table disagrees with the stream at byte " + std::to_string(i);
      return STRIDE_E_CORRUPT;
    }
  }
  return STRIDE_OK;
}

}  // namespace stride
