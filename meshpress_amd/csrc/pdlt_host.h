// Host-side PDLT v5 set-up shared by the decoder (pdlt_decode.hip) and the
// CPU harness of its kernel logic: global header, boundary table, meshlet
// pre-pass and the wave-group layout.  Host code only.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/stride_hip.h"
#include "pdlt_lane.h"

namespace stride {

struct PdltHost {
  // global header (encoder/paradelta_v5.py:290-311)
  uint32_t code_width = 0;
  float center[3], scale = 0.f, pce = 0.f, g_min[3], g_range[3], lin3[3];
  int32_t g_bits[3];
  uint32_t hdr_n_v = 0, hdr_n_t = 0;
  uint32_t n_b = 0, n_m = 0;
  uint64_t n_v = 0, n_t = 0;            // decoded counts
  uint64_t n_refs = 0;                  // boundary refs over all meshlets
  size_t stream_bytes = 0;
  std::vector<uint32_t> words;          // the stream as little-endian words, zero padded
  std::vector<double> bnd;              // boundary table, normalised f64 (n_b x 3)
  std::vector<PdltEntry> tab;
  PdltParams params{};
};

namespace pdlt_detail {

inline uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
inline float bef32(const uint8_t* p) {
  const uint32_t u = be32(p);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

inline int scan_status(uint32_t e, uint32_t m, std::string& err) {
  const std::string at = " (meshlet " + std::to_string(m) + ")";
  switch (e) {
    case kV5ErrSize: err = "meshlet with more than 256 vertices" + at; return STRIDE_E_UNSUPPORTED;
    case kV5ErrOverrun: err = "meshlet runs past the end of the stream" + at; return STRIDE_E_TRUNCATED;
    case kV5ErrFifo: err = "reuse FIFO index past the FIFO fill" + at; return STRIDE_E_CORRUPT;
    case kV5ErrIndex: err = "local vertex id >= n_local" + at; return STRIDE_E_CORRUPT;
    case kV5ErrStrips: err = "strip lengths do not add up to n_tris" + at; return STRIDE_E_CORRUPT;
    case kV5ErrTag: err = "bad residual tag" + at; return STRIDE_E_CORRUPT;
    case kV5ErrRef: err = "boundary ref outside the boundary table" + at; return STRIDE_E_UNSUPPORTED;
    case kV5ErrRange:
      err = "residual code or coder parameter outside 32-bit decoding" + at;
      return STRIDE_E_UNSUPPORTED;
    default: err = "malformed meshlet" + at; return STRIDE_E_CORRUPT;
  }
}

template <class F>
void par_for(int64_t n, F&& fn) {
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 256 || nt == 1) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&]() {
      for (;;) {
        const int64_t s = next.fetch_add(64);
        if (s >= n) break;
        for (int64_t i = s; i < std::min(n, s + 64); ++i) fn(i);
      }
    });
  for (auto& x : th) x.join();
}

}  // namespace pdlt_detail

// Parses and validates a PDLT v5 stream.  Returns STRIDE_OK or a STRIDE_E_*
// status with the message in err.  Accepted streams decode with
// decode_paradelta_v5 (encoder/paradelta_v5.py:286-462) to the same output
// the kernels produce; the meshlet offsets must match the sequential layout
// that decoder reads.
inline int pdlt_parse(const uint8_t* d, size_t len, PdltHost& H, std::string& err) {
  using namespace pdlt_detail;
  if (len < 5) { err = "stream shorter than the PDLT v5 header"; return STRIDE_E_TRUNCATED; }
  const uint32_t magic = be32(d);
  if (magic != kPdltMagic) {
    char b[64];
    std::snprintf(b, sizeof b, "bad magic 0x%08X", magic);
    err = b;
    return STRIDE_E_MAGIC;
  }
  if (d[4] != kPdltVersion) {
    char b[64];
    std::snprintf(b, sizeof b, "expected v5, got v%u", (unsigned)d[4]);
    err = b;
    return STRIDE_E_VERSION;
  }
  if (len < kPdltHeaderBytes) { err = "stream shorter than the PDLT v5 header"; return STRIDE_E_TRUNCATED; }
  if (len > 0x1FFFFF00ull) { err = "streams above 512 MiB are not supported"; return STRIDE_E_UNSUPPORTED; }
  H.stream_bytes = len;
  H.code_width = d[5];
  for (int i = 0; i < 3; ++i) H.center[i] = bef32(d + 6 + 4 * i);
  H.scale = bef32(d + 18);
  H.pce = bef32(d + 22);
  for (int i = 0; i < 3; ++i) H.g_min[i] = bef32(d + 26 + 4 * i);
  for (int i = 0; i < 3; ++i) H.g_range[i] = bef32(d + 38 + 4 * i);
  for (int i = 0; i < 3; ++i) H.g_bits[i] = d[50 + i];
  H.hdr_n_v = be32(d + 53);
  H.hdr_n_t = be32(d + 57);
  H.n_b = be32(d + 61);
  H.n_m = be32(d + 65);
  for (int i = 0; i < 3; ++i) H.lin3[i] = bef32(d + 69 + 4 * i);
  for (int i = 0; i < 3; ++i)
    if (H.g_bits[i] > 32) { err = "g_bits above 32 are not supported"; return STRIDE_E_UNSUPPORTED; }
  const uint64_t end_all = 8ull * len;
  const uint32_t n_b = H.n_b, n_m = H.n_m;
  // every boundary difference takes at least one bit, every offset 32
  if ((uint64_t)n_b > end_all + 1 || 32ull * n_m > end_all) {
    err = "stream shorter than its boundary and offset tables";
    return STRIDE_E_TRUNCATED;
  }
  const size_t nw = (len + 3) / 4 + 8;   // the staged ranges read up to 6 words past a meshlet
  H.words.assign(nw, 0);
  std::memcpy(H.words.data(), d, len);
  GlobalWords src{H.words.data(), (uint32_t)nw};

  // boundary table (encoder/paradelta_v5.py:313-332)
  H.bnd.assign(3ull * n_b, 0.0);
  BitReader<GlobalWords> r;
  r.init(src, 8 * kPdltHeaderBytes);
  {
    std::vector<int64_t> code(n_b);
    for (int a = 0; a < 3 && n_b > 0; ++a) {
      code[0] = r.bits(H.g_bits[a]);
      if (n_b > 1) {
        const uint32_t k = r.bits(8);
        if (k > 32) { err = "boundary coder parameter k > 32"; return STRIDE_E_UNSUPPORTED; }
        for (uint32_t i = 1; i < n_b; ++i) {
          uint64_t q;
          const uint32_t e = v5_scan_unary(r, end_all, k >= 32 ? 1ull : (1ull << (62 - k)), q);
          if (e == kV5ErrOverrun || r.pos > end_all) {
            err = "boundary table runs past the end of the stream";
            return STRIDE_E_TRUNCATED;
          }
          if (e) { err = "boundary table value outside 64-bit decoding"; return STRIDE_E_UNSUPPORTED; }
          const uint64_t u = (q << k) | r.bits((int)k);
          code[i] = code[i - 1] + (int64_t)((u >> 1) ^ (0ull - (u & 1)));
        }
      }
      if (r.pos > end_all) { err = "boundary table runs past the end of the stream"; return STRIDE_E_TRUNCATED; }
      const uint64_t mx = (H.g_bits[a] ? (1ull << H.g_bits[a]) : 1ull) - 1;
      // g_min + code / mx * g_range in f64 (paradelta_v5.py:322-332)
      for (uint32_t i = 0; i < n_b; ++i)
        H.bnd[3ull * i + a] = mx == 0 ? (double)H.g_min[a]
                                      : (double)H.g_min[a] + (double)code[i] / (double)mx * (double)H.g_range[a];
    }
  }
  const uint64_t pos = (r.pos + 7) & ~7ull;
  if (pos + 32ull * n_m > end_all) {
    err = "stream shorter than its meshlet offset table";
    return STRIDE_E_TRUNCATED;
  }
  std::vector<uint32_t> off(n_m);
  for (uint32_t m = 0; m < n_m; ++m) off[m] = be32(d + pos / 8 + 4ull * m);
  const uint64_t region = pos + 32ull * n_m;
  if (n_m > 0 && off[0] != region) {
    err = "meshlet offsets disagree with the meshlet layout";
    return STRIDE_E_CORRUPT;
  }

  // pre-pass (the parse of paradelta_v5.py:345-446)
  H.tab.assign(n_m, PdltEntry{});
  std::vector<uint32_t> status(n_m, 0);
  par_for(n_m, [&](int64_t m) {
    if (off[m] >= end_all) {
      status[m] = kV5ErrOverrun;
      return;
    }
    status[m] = pdlt_scan(src, off[m], end_all, n_b, H.tab[m]);
  });
  uint64_t iv = n_b, tv = 0, rv = 0;
  for (uint32_t m = 0; m < n_m; ++m) {
    if (status[m]) return scan_status(status[m], m, err);
    if (m + 1 < n_m && H.tab[m].end_bit != off[m + 1]) {
      err = "meshlet offsets disagree with the meshlet layout";
      return STRIDE_E_CORRUPT;
    }
    H.tab[m].int_base = (uint32_t)iv;
    H.tab[m].t_off = (uint32_t)tv;
    H.tab[m].ref_base = (uint32_t)rv;
    iv += H.tab[m].n_int;
    tv += H.tab[m].n_tris;
    rv += H.tab[m].n_bnd;
    if (iv > 0xFFFFFFFFull || 3 * tv > 0xFFFFFFFFull || rv > 0xFFFFFFFFull) {
      err = "more than 2^32 output values";
      return STRIDE_E_UNSUPPORTED;
    }
  }
  H.n_v = iv;
  H.n_t = tv;
  H.n_refs = rv;
  PdltParams& P = H.params;
  for (int i = 0; i < 3; ++i) P.w[i] = (double)H.lin3[i];
  P.delta = 2.0 * (double)H.pce;
  for (int i = 0; i < 3; ++i) P.center[i] = (double)H.center[i];
  P.scale = (double)H.scale;
  P.world = 1;
  return STRIDE_OK;
}

// Greedy wave groups: up to gs consecutive meshlets whose staged words and
// LDS blocks fit `budget` bytes (a meshlet alone always forms a group).
// Sets each entry's blk_off; lds = the largest group's LDS bytes.  refs:
// the blocks hold a slot of n_bnd words for the boundary refs (the
// one-kernel decode decodes them there; the chain pass of the two-pass
// decode reads them from HBM and keeps the slot only where
// pdlt_chain_block_refs says so).
inline int pdlt_layout(std::vector<PdltEntry>& tab, int gs, int budget, std::vector<PdltGroup>& groups,
                       uint32_t& lds, std::string& err, uint32_t rw = 6, bool refs = true) {
  if (gs < 1 || gs > 16) { err = "group_size must be in [1, 16]"; return STRIDE_E_ARG; }
  if (budget < 4096 || budget > 65536) { err = "lds_bytes must be in [4096, 65536]"; return STRIDE_E_ARG; }
  groups.clear();
  uint32_t max_lds = 0;
  const uint32_t n_m = (uint32_t)tab.size();
  uint32_t m = 0;
  while (m < n_m) {
    PdltGroup g{};
    g.m0 = m;
    g.w_lo = tab[m].start_bit >> 5;
    uint32_t blk = 0, words = 0, j = m;
    while (j < n_m && (int)(j - m) < gs) {
      // the words holding [start, end) plus the axis lanes' read-ahead (the
      // stream window reads up to 3 words past its position), even count
      const uint32_t w = ((((tab[j].end_bit - 1) >> 5) + 6 - g.w_lo) + 1) & ~1u;
      const uint32_t b = pdlt_block_words(tab[j].n_bnd, tab[j].n_int, rw, refs);
      if (j > m && 4ull * (w + blk + b) > (uint64_t)budget) break;
      tab[j].blk_off = blk;
      blk += b;
      words = w;
      ++j;
    }
    g.m1 = j;
    g.n_words = words;
    g.blk_base = words;
    max_lds = std::max(max_lds, 4 * (words + blk));
    groups.push_back(g);
    m = j;
  }
  if (max_lds > 65536) { err = "a meshlet needs more than 64 KiB of LDS"; return STRIDE_E_UNSUPPORTED; }
  lds = max_lds;
  return STRIDE_OK;
}

// Token groups of the two-pass decode: up to gmax (<= 64) consecutive
// meshlets per wave (one token lane each) whose staged words, refs and
// interleaved seen sets (7 kTokenSeenStride + lanes words) fit `budget` bytes.  Sets each entry's tok_ref_off; lds = the largest group's
// LDS bytes.
inline int pdlt_layout_tokens(std::vector<PdltEntry>& tab, int budget, int gmax,
                              std::vector<PdltGroup>& groups, uint32_t& lds, std::string& err) {
  if (budget < 4096 || budget > 65536) { err = "lds_bytes must be in [4096, 65536]"; return STRIDE_E_ARG; }
  if (gmax < 1 || gmax > 64) { err = "tok_group must be in [1, 64]"; return STRIDE_E_ARG; }
  groups.clear();
  uint32_t max_lds = 0;
  const uint32_t n_m = (uint32_t)tab.size();
  uint32_t m = 0;
  while (m < n_m) {
    PdltGroup g{};
    g.m0 = m;
    g.w_lo = tab[m].start_bit >> 5;
    uint32_t refs = 0, words = 0, j = m;
    while (j < n_m && j - m < (uint32_t)gmax) {
      const uint32_t w = ((((tab[j].end_bit - 1) >> 5) + 2 - g.w_lo) + 1) & ~1u;
      // per meshlet: refs (n_bnd words); then the group's seen sets
      const uint64_t need = w + refs + tab[j].n_bnd + 7 * kTokenSeenStride + (j - m + 1);
      if (j > m && 4 * need > (uint64_t)budget) break;
      tab[j].tok_ref_off = refs;
      refs += tab[j].n_bnd;
      words = w;
      ++j;
    }
    g.m1 = j;
    g.n_words = words;
    g.blk_base = words;
    g.seen_base = words + refs;
    max_lds = std::max(max_lds, 4 * (words + refs + 7 * kTokenSeenStride + (j - m)));
    groups.push_back(g);
    m = j;
  }
  if (max_lds > 65536) { err = "a meshlet needs more than 64 KiB of LDS"; return STRIDE_E_UNSUPPORTED; }
  lds = max_lds;
  return STRIDE_OK;
}

// Chain token waves of the two-pass decode (pdlt_chain_tokens_kernel):
// consecutive meshlets, up to gmax meshlets and 64 token chains (about k0 a
// meshlet, pdlt_chain_tokens) a wave, whose LDS fits kPdltChainLds bytes:
// the staged bits from the first meshlet header to the end of the last
// token section (+ 3 words: the refs' read-ahead window), the refs (n_bnd
// words a meshlet), the 16-bit token records (each meshlet's padded to an
// even count), 16 bytes of final FIFO per chain and 256 first-sighting
// words; at most 64 boundary-ref chunks a wave.  Sets each entry's
// tok_ref_off and chunk_base, each group's refs base (blk_base), records
// base (seen_base) and chain range [pad[0], pad[1]), and writes the parser
// state at every cut to `chains` and the ref chunks to `rchunks` (stream =
// the stream's words).  STRIDE_E_UNSUPPORTED when a meshlet alone does not
// fit (its tokens then need the token lanes).
constexpr uint32_t kPdltChainLds = 49152;
inline int pdlt_layout_chains(std::vector<PdltEntry>& tab, const std::vector<uint32_t>& stream,
                              int k0, int gmax, std::vector<PdltGroup>& groups,
                              std::vector<ChainEntry>& chains, std::vector<RefChunk>& rchunks,
                              uint32_t& lds, std::string& err) {
  if (k0 < 1 || k0 > kMaxChains) { err = "tok_chains must be in [0, 16]"; return STRIDE_E_ARG; }
  if (gmax < 1 || gmax > 64) { err = "tok_group must be in [1, 64]"; return STRIDE_E_ARG; }
  groups.clear();
  const uint32_t n_m = (uint32_t)tab.size();
  std::vector<uint32_t> cbase(n_m, 0);
  auto need = [](uint64_t w, uint64_t refs, uint64_t recw) {
    return 4 * (w + refs + recw) + 16 * 64 + 4 * 256;
  };
  uint32_t max_lds = 0, nc = 0, nr = 0, m = 0;
  while (m < n_m) {
    PdltGroup g{};
    g.m0 = m;
    g.w_lo = tab[m].start_bit >> 5;
    uint32_t refs = 0, recw = 0, words = 0, cn = 0, rn = 0, j = m;
    while (j < n_m && j - m < (uint32_t)gmax) {
      const PdltEntry& e = tab[j];
      const uint32_t nt = pdlt_tokens_of(e);
      const uint32_t cnt = chain_count(nt, pdlt_chain_tokens(e, (uint32_t)k0));
      const uint32_t rcn = pdlt_ref_chunks(e.n_bnd);
      const uint32_t w = std::max(words, ((pdlt_tokens_end(e) + 31) >> 5) + 3 - g.w_lo);
      const uint32_t rw = ((nt + 1) & ~1u) / 2;
      // a chunk's code offset from the header must fit 16 bits
      if (e.strips_bit - e.start_bit > 0xFFFFu ||
          (j == m && need(w, refs + e.n_bnd, recw + rw) > kPdltChainLds)) {
        err = "meshlet " + std::to_string(j) + " too large for the chain token pass";
        return STRIDE_E_UNSUPPORTED;
      }
      if (need(w, refs + e.n_bnd, recw + rw) > kPdltChainLds || cn + cnt > 64 || rn + rcn > 64)
        break;
      tab[j].tok_ref_off = refs;
      tab[j].chunk_base = nr + rn;
      cbase[j] = nc + cn;
      refs += e.n_bnd;
      recw += rw;
      words = w;
      cn += cnt;
      rn += rcn;
      ++j;
    }
    g.m1 = j;
    g.n_words = words;
    g.blk_base = words;
    g.seen_base = words + refs;
    g.pad[0] = nc;
    g.pad[1] = nc + cn;
    nc += cn;
    nr += rn;
    max_lds = std::max(max_lds, (uint32_t)need(words, refs, recw));
    groups.push_back(g);
    m = j;
  }
  chains.assign(nc, ChainEntry{0u, 0u});
  rchunks.assign(nr, RefChunk{0u, 0u});
  const GlobalWords src{stream.data(), (uint32_t)stream.size()};
  pdlt_detail::par_for(n_m, [&](int64_t i) {
    const PdltEntry& e = tab[i];
    chain_starts_at(src, e.strips_bit, e.strips_bit, pdlt_tokens_of(e), e.n_bnd + e.n_int,
                    pdlt_chain_tokens(e, (uint32_t)k0), chains.data() + cbase[i]);
    // ref chunks: the code offset and the ref before every kRefChunk-th ref
    RefChunk* rc = rchunks.data() + e.chunk_base;
    rc[0] = RefChunk{64u, 0u};
    if (e.n_bnd > 1) {
      BitReader<GlobalWords> r;
      r.init(src, e.start_bit + 64);
      uint32_t prev = r.bits(32);
      const uint32_t k = r.bits(8);
      for (uint32_t q = 1; q < e.n_bnd; ++q) {
        if (q % kRefChunk == 0)
          rc[q / kRefChunk] = RefChunk{(r.pos - e.start_bit) | ((q / kRefChunk) << 16), prev};
        uint32_t z = 0;
        while (r.bits(1) == 0) ++z;
        prev = prev + ((z << k) | r.bits((int)k)) + 1u;
      }
    }
  });
  lds = max_lds;
  return STRIDE_OK;
}

}  // namespace stride
