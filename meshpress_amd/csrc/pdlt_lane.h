// PDLT v5 (ParaDelta v5) decode logic shared by the gfx950 kernels
// (pdlt_decode.hip) and the CPU harness that checks it against the oracle
// (compiled as host C++ there).
//
// Stream layout (encoder/paradelta_v5.py:144-283, paradelta_v5_nb.py:328-539):
//   global header (81 bytes, big endian), boundary table (per axis: first
//   code, then Rice-coded zigzag differences), pad to a byte, n_meshlets
//   absolute bit offsets (u32), then the meshlets, bit packed back to back:
//     n_bnd, n_int, n_tris, n_strips     4 x u16
//     boundary refs                       u32 first, then u8 k and
//                                         Rice(k) of (ref[i+1] - ref[i] - 1)
//     strips                              u16 length, 3 root tokens, then
//                                         (length - 1) x (edge bit, token)
//     residual codes (if n_int > 0)       per axis: u8 tag; 0: i16 min, u8
//                                         width, n_int fixed-width values;
//                                         1 / 2: u8 k, n_int Rice / Exp-Golomb
//                                         zigzag values
// Tokens are those of the DPR4 stream: '0' + 5-bit reuse-FIFO index (oldest
// first) or '1' + raw local id of idx_bits(n_local) bits.
//
// Decode (encoder/paradelta_v5.py:286-462): boundary locals take the
// dequantised boundary-table position of their ref; each interior local is
// predicted, in first-use order along the strips, from the boundary
// centroid (strip-root vertices) or from LIN3 weights on the parallelogram
// (a, b, c) of the strip step, and corrected by code * 2 * per_coord_err,
// all in float64.
#pragma once

#include <cstdint>

#include "stride_bits.h"
#include "stride_lane.h"

namespace stride {

constexpr uint32_t kPdltMagic = 0x504C4454u;   // 'PLDT' as written (MAGIC)
constexpr uint32_t kPdltVersion = 5u;
constexpr uint32_t kPdltHeaderBytes = 81;

// Per-meshlet side table from the pre-pass.  Bit offsets are absolute in
// the stream.
struct PdltEntry {
  uint32_t start_bit;    // meshlet header
  uint32_t strips_bit;   // first strip length
  uint32_t axis_bit[3];  // residual tag byte of each axis (0 when n_int == 0)
  uint32_t end_bit;      // one past the last bit
  uint32_t n_bnd, n_int, n_tris;
  uint32_t n_strips;     // | kFlagRawInFifo
  uint32_t int_base;     // output index of the first interior vertex
  uint32_t t_off;        // first output triangle
  uint32_t blk_off;      // LDS block, 32-bit words from the group's block base
  uint32_t tok_ref_off;  // two-pass layout: refs in the token group's LDS, words
  uint32_t ref_base;     // two-pass layout: first word of its refs in the refs buffer
  uint32_t chunk_base;   // chain token pass: its first boundary-ref chunk
};
static_assert(sizeof(PdltEntry) == 64, "PdltEntry is one 64-byte line");

// Consecutive meshlets decoded by one wave.
struct PdltGroup {
  uint32_t m0, m1;       // meshlet range [m0, m1)
  uint32_t w_lo;         // first staged stream word
  uint32_t n_words;      // staged words (even)
  uint32_t blk_base;     // LDS word offset of the meshlet blocks (even)
  uint32_t seen_base;    // token groups: LDS word offset of the interleaved seen sets
  uint32_t pad[2];
};
static_assert(sizeof(PdltGroup) == 32, "PdltGroup is 32 bytes");

// Two-pass layout: the walk of meshlet m (n_int words and one spare word,
// which the token lane's unconditional entry stores may reach) goes to the
// walk buffer at pdlt_walk_off; its seen set (8 words) and walk length
// (1 word) to the seen buffer at kSeenWords * m; its boundary refs (n_bnd
// words) to the refs buffer at ref_base (the meshlets' refs back to back).
constexpr uint32_t kSeenWords = 9;

// Walk buffer words: n_int + n_meshlets.
STRIDE_DEV uint64_t pdlt_walk_off(uint32_t int_base, uint32_t m, uint32_t n_boundary) {
  return (uint64_t)(int_base - n_boundary) + m;
}

// Strip tokens of a meshlet (a strip of length L has max(L, 1) triangles
// and two more tokens).
STRIDE_DEV uint32_t pdlt_tokens_of(const PdltEntry& e) {
  return e.n_tris + 2 * (e.n_strips & kStripMask);
}

// Chain token pass: meshlets of up to kPdltChainMaxTokens tokens are cut
// into about k0 chains (chain_tokens), so a cut's strip token index (9 bits)
// and strip token count (10 bits) fit a ChainEntry; longer ones stay one
// chain.  Returns the tokens per chain.
constexpr uint32_t kPdltChainMaxTokens = 511;
STRIDE_DEV uint32_t pdlt_chain_tokens(const PdltEntry& e, uint32_t k0) {
  const uint32_t n = pdlt_tokens_of(e);
  if (n > kPdltChainMaxTokens) return n;
  return chain_tokens(n, e.n_bnd + e.n_int, e.n_strips, k0);
}

// Boundary-ref chunks of the chain token pass: a meshlet's refs are cut
// every kRefChunk refs so that several lanes decode them at once (every
// meshlet has at least one chunk).  pos: bit offset of the chunk's first
// Rice code from the meshlet header (chunk 0: 64, the first ref's u32), in
// the low 16 bits | chunk index << 16; prev: the ref before the chunk's
// first.
constexpr uint32_t kRefChunk = 16;
struct RefChunk {
  uint32_t pos;
  uint32_t prev;
};
STRIDE_DEV uint32_t pdlt_ref_chunks(uint32_t n_bnd) {
  return n_bnd > kRefChunk ? (n_bnd + kRefChunk - 1) / kRefChunk : 1u;
}

// One past the last strip-token bit of a meshlet: its residual section, or
// its end when it has no interior vertex.
STRIDE_DEV uint32_t pdlt_tokens_end(const PdltEntry& e) {
  return e.n_int ? e.axis_bit[0] : e.end_bit;
}

// Decode constants, widened to float64 as the reference decoder does.
struct PdltParams {
  double w[3];           // LIN3 weights
  double delta;          // 2 * per_coord_err
  double center[3];
  double scale;
  int32_t world;         // 1: world coordinates, 0: normalised
  int32_t chain32;       // 1: the encoder's float32 chain, 0: float64 (decode_paradelta_v5)
  int32_t pad[2];
};

enum : uint32_t {
  kV5Ok = 0,
  kV5ErrSize = 1,      // n_bnd + n_int > 256
  kV5ErrOverrun = 2,   // a field runs past the end of the stream
  kV5ErrFifo = 3,      // reuse-FIFO index past the FIFO fill
  kV5ErrIndex = 4,     // raw local id >= n_local
  kV5ErrStrips = 5,    // strip lengths do not add up to n_tris
  kV5ErrTag = 6,       // residual tag not 0, 1 or 2
  kV5ErrRef = 7,       // boundary ref outside the boundary table
  kV5ErrRange = 8,     // a code or coder parameter outside 32-bit decoding
};

// LDS block of one meshlet, in 32-bit words: reconstruction (3 axes x
// n_local float64, axis major), refs (n_bnd words when refs = true: the
// one-kernel decode's lanes decode them there, the float64 chain pass keeps
// the slot unused, pdlt_chain_block_refs), walk (n_int), the seen set (8
// words) and the walk length (1 word); even, so blocks stay 8-byte aligned.
// rw = words of reconstruction per local: 6 (three float64, the float64
// chain) or 3 (three float32, the float32 chain)
STRIDE_DEV uint32_t pdlt_block_words(uint32_t n_bnd, uint32_t n_int, uint32_t rw = 6,
                                     bool refs = true) {
  const uint32_t n_local = n_bnd + n_int;
  return (rw * n_local + (refs ? n_bnd : 0u) + n_int + 9 + 1) & ~1u;
}
// Whether the chain pass's blocks keep the refs slot (the chain pass reads
// the refs from HBM and never uses it).  The float32 chain drops it: 0.220 ->
// 0.209 ms on the 1900 x 1900 heightfield.  The float64 chain keeps it: without
// it every LDS budget measured about 2% slower (0.2525 -> 0.258 ms in groups of
// 16 in 20 KB), and so did the same words at the block's end (the same group
// packing) or a slot twice as long: the walk's position matters there.
constexpr bool pdlt_chain_block_refs(bool chain32) { return !chain32; }

STRIDE_DEV uint32_t pdlt_idx_bits(uint32_t n_local) {
  uint32_t b = 1;
  while ((1u << b) < n_local + 1) ++b;
  return b;
}

// Unary run at r of any length, the terminating 1 consumed: kV5ErrOverrun
// when the stream ends first, kV5ErrRange when the run reaches `bound`.
template <class Src>
STRIDE_DEV uint32_t v5_scan_unary(BitReader<Src>& r, uint64_t end_all, uint64_t bound, uint64_t& q) {
  q = 0;
  for (;;) {
    r.refill();
    const int lz = r.win ? __builtin_clzll(r.win) : 64;
    if (lz < r.valid) {
      r.win = (r.win << lz) << 1;
      r.valid -= lz + 1;
      r.pos += (uint32_t)lz + 1u;
      q += (uint64_t)lz;
      break;
    }
    q += (uint64_t)r.valid;
    r.pos += (uint32_t)r.valid;
    r.win = 0;
    r.valid = 0;
    if (r.pos > end_all) return kV5ErrOverrun;
    if (q >= bound) return kV5ErrRange;
  }
  if (r.pos > end_all) return kV5ErrOverrun;
  return q >= bound ? kV5ErrRange : kV5Ok;
}

// ---------------------------------------------------------------------------
// Pre-pass: one meshlet from start_bit, every field bounds checked.  Fills
// counts and section offsets; returns kV5Ok or an error code.  Accepts
// exactly the meshlets decode_paradelta_v5 decodes, minus the documented
// limits (n_local <= 256, refs inside the boundary table, codes and coder
// parameters within 32-bit decoding).
// ---------------------------------------------------------------------------
template <class Src>
STRIDE_DEV uint32_t pdlt_scan(const Src& src, uint32_t start, uint64_t end_all,
                              uint64_t n_boundary, PdltEntry& e) {
  BitReader<Src> r;
  r.init(src, start);
  e.start_bit = start;
  e.n_bnd = r.bits(16);
  e.n_int = r.bits(16);
  e.n_tris = r.bits(16);
  const uint32_t n_strips = r.bits(16);
  e.n_strips = n_strips;
  e.axis_bit[0] = e.axis_bit[1] = e.axis_bit[2] = 0;
  if (r.pos > end_all) return kV5ErrOverrun;
  const uint32_t n_local = e.n_bnd + e.n_int;
  if (n_local > 256) return kV5ErrSize;
  if (e.n_bnd > 0) {
    uint64_t prev = r.bits(32);
    if (prev >= n_boundary) return kV5ErrRef;
    if (e.n_bnd > 1) {
      const uint32_t k = r.bits(8);
      if (k > 31) return kV5ErrRange;
      for (uint32_t i = 1; i < e.n_bnd; ++i) {
        uint64_t q;
        const uint32_t err = v5_scan_unary(r, end_all, 1ull << (32 - k), q);
        if (err) return err;
        const uint64_t u = (q << k) | r.bits((int)k);
        prev = prev + u + 1;
        if (prev >= n_boundary) return kV5ErrRef;
      }
    }
  }
  if (r.pos > end_all) return kV5ErrOverrun;
  e.strips_bit = r.pos;
  const uint32_t ib = pdlt_idx_bits(n_local);
  MtfLite fifo;
  fifo.clear();
  bool raw_in_fifo = false;
  uint64_t tris = 0;
  for (uint32_t s = 0; s < n_strips; ++s) {
    const uint32_t L = r.bits(16);
    const uint32_t steps = L ? L - 1 : 0;
    tris += steps + 1;
    for (uint32_t t = 0; t < 3 + steps; ++t) {
      if (t >= 3) r.bits(1);
      uint32_t j, v;
      if (r.bits(1) == 0) {
        const uint32_t idx = r.bits(5);
        if (idx >= fifo.n) return kV5ErrFifo;
        j = fifo.n - 1 - idx;
        v = fifo.rank(j);
      } else {
        v = r.bits((int)ib);
        if (v >= n_local) return kV5ErrIndex;
        const int pf = fifo.find(v);
        raw_in_fifo |= pf >= 0;
        j = pf >= 0 ? (uint32_t)pf : fifo.n;
        if (j == fifo.n && fifo.n < 16) fifo.n++;
      }
      fifo.to_front(j, v);
    }
    if (r.pos > end_all) return kV5ErrOverrun;
  }
  if (tris != e.n_tris) return kV5ErrStrips;
  if (raw_in_fifo) e.n_strips |= kFlagRawInFifo;
  if (e.n_int > 0) {
    for (int d = 0; d < 3; ++d) {
      e.axis_bit[d] = r.pos;
      const uint32_t tag = r.bits(8);
      if (tag == 0) {
        const uint32_t mu = r.bits(16);
        const int64_t mn = (mu & 0x8000u) ? (int64_t)mu - 0x10000 : (int64_t)mu;
        const uint32_t bw = r.bits(8);
        if (bw > 32) return kV5ErrRange;
        if (mn + (int64_t)((1ull << bw) - 1) > 0x7FFFFFFFll) {
          for (uint32_t i = 0; i < e.n_int; ++i)
            if (mn + (int64_t)r.bits((int)bw) > 0x7FFFFFFFll) return kV5ErrRange;
        } else {
          const uint64_t skip = r.pos + (uint64_t)bw * e.n_int;
          if (skip > end_all) return kV5ErrOverrun;
          r.init(r.src, (uint32_t)skip);
        }
      } else if (tag == 1 || tag == 2) {
        const uint32_t k = r.bits(8);
        if (k > 31) return kV5ErrRange;
        for (uint32_t i = 0; i < e.n_int; ++i) {
          uint64_t q;
          const uint32_t err = v5_scan_unary(r, end_all, tag == 1 ? (1ull << (32 - k)) : 32ull, q);
          if (err) return err;
          if (tag == 1) {
            r.bits((int)k);
          } else {
            const uint64_t shifted = ((1ull << q) - 1) + r.bits((int)q);
            const uint64_t u = (shifted << k) | r.bits((int)k);
            if (u >> 32) return kV5ErrRange;
          }
        }
      } else {
        return kV5ErrTag;
      }
      if (r.pos > end_all) return kV5ErrOverrun;
    }
  }
  e.end_bit = r.pos;
  if (r.pos > end_all) return kV5ErrOverrun;
  return kV5Ok;
}

// ---------------------------------------------------------------------------
// Lane code over the staged stream (big-endian words in LDS; `rel` maps a
// stream bit to a staged bit: staged = stream - 32 * w_lo).
// ---------------------------------------------------------------------------

// Unsigned Rice(k) value from a stream: one window step for codewords of at
// most 32 bits, bit-serial beyond (the pre-pass bounds the value below 2^32).
template <class Stream>
STRIDE_DEV uint32_t rice_raw(Stream& s, uint32_t k) {
  uint32_t u;
  const uint32_t len = codeword_win(s.win, 0u, (int)k, u);
  if (len) {
    s.drop(len);
    return u;
  }
  const uint32_t q = unary_s(s);
  return (q << k) | (k ? s.bits(k) : 0u);
}

STRIDE_DEV void pdlt_refs(const uint32_t* __restrict__ st, int64_t rel, const PdltEntry& e,
                          uint32_t* __restrict__ refs) {
  const uint32_t n_bnd = e.n_bnd;
  if (n_bnd == 0) return;
  // the refs stream through a read-ahead window (the strip tokens follow)
  BitStream<LdsStaged> bs;
  bs.init(LdsStaged{st, 0}, (uint64_t)((int64_t)e.start_bit + rel) + 64u);
  uint32_t prev = bs.bits(32);
  refs[0] = prev;
  if (n_bnd > 1) {
    const uint32_t k = bs.bits(8);
    for (uint32_t i = 1; i < n_bnd; ++i) {
      prev = prev + rice_raw(bs, k) + 1u;
      refs[i] = prev;
    }
  }
}

// The refs of one chunk of a meshlet (header at stream bit `start`, n_bnd
// refs) into refs[] (the meshlet's whole ref array), from the staged stream
// as pdlt_refs.
STRIDE_DEV void pdlt_refs_chunk(const uint32_t* __restrict__ st, int64_t rel, uint32_t start,
                                uint32_t n_bnd, const RefChunk& ch,
                                uint32_t* __restrict__ refs) {
  if (n_bnd == 0) return;
  const uint32_t c = ch.pos >> 16;
  const uint32_t end = (c + 1) * kRefChunk < n_bnd ? (c + 1) * kRefChunk : n_bnd;
  const uint32_t h = (uint32_t)((int64_t)start + rel);
  const uint32_t k = n_bnd > 1 ? win32(st, h + 96) >> 24 : 0u;
  BitStream<LdsStaged> bs;
  bs.init(LdsStaged{st, 0}, h + (ch.pos & 0xFFFFu));
  uint32_t prev = ch.prev, i = c * kRefChunk;
  if (c == 0) {
    prev = bs.bits(32);
    refs[0] = prev;
    bs.drop(n_bnd > 1 ? 8u : 0u);
    i = 1;
  }
  for (; i < end; ++i) {
    prev = prev + rice_raw(bs, k) + 1u;
    refs[i] = prev;
  }
}

// Output ids of a PDLT meshlet's locals: boundary locals through their
// global refs, interior locals after the meshlet's interior base; strip
// roots enter the walk as v in every byte (the chain's a == v test).
// The lookup is branch-free: an interior local reads refs[0], which lies in
// the LDS block in both layouts (the seen set follows the refs), and drops it.
struct PdltIds {
  static constexpr bool kCarry = true;
  const uint32_t* __restrict__ refs;
  uint32_t n_bnd, ibase;
  STRIDE_DEV uint32_t operator()(uint32_t x) const {
    const bool b = x < n_bnd;
    const uint32_t r = refs[b ? x : 0u];
    return b ? r : ibase + x;
  }
};

// Token lane: boundary refs into refs[], strips into the walk (interior
// locals in first-use order: v * 0x01010101 for a strip-root vertex,
// v | a << 8 | b << 16 | c << 24 for a parallelogram step, a != v) and the
// triangles into tout (global ids).  `seen` (LDS in the kernels) is the live
// seen set; the walk length goes to *n_emit.  Sink = WalkSlot stores every
// token's entry (the walk needs one spare word after n_int).
template <class Sink = WalkArray>
STRIDE_DEV void pdlt_tokens(const uint32_t* __restrict__ st, int64_t rel, const PdltEntry& e,
                            uint32_t* __restrict__ refs, uint32_t* __restrict__ walk,
                            BitSeen seen, uint32_t* __restrict__ n_emit,
                            uint32_t* __restrict__ tout) {
  const uint32_t n_bnd = e.n_bnd, n_local = e.n_bnd + e.n_int;
  pdlt_refs(st, rel, e, refs);
  // boundary locals preset: only interior locals enter the walk
  seen.preset(n_bnd);
  LdsCursor cur{st, (uint32_t)((int64_t)e.strips_bit + rel)};
  TokenLane<PdltIds> t;
  t.init(cur, n_local, e.n_tris, e.n_strips, tout, PdltIds{refs, n_bnd, e.int_base - n_bnd});
  Sink sink{walk};
  tokens_run(t, cur, seen, sink);
  *n_emit = t.nw;
}

// Axis lane d: boundary centroid of axis d (the boundary rows summed in
// order, then divided by n_bnd), then the residual codes of axis d along
// the walk.  rc = the axis-d reconstruction array of the meshlet, whose
// boundary entries are filled.
//
// T selects the arithmetic (the decoder runs T = float exactly when
// P.chain32 = 1): double = float64, as decode_paradelta_v5
// (encoder/paradelta_v5.py:422-446); float = the float32 chain the encoder
// computed its codes against (encoder/paradelta_v5_nb.py:44-59,113-176:
// float32 boundary positions, float32 centroid, float32 LIN3 terms summed
// left to right, code * float32 delta), so every position is the encoder's
// own reconstruction.
template <class T>
STRIDE_DEV void pdlt_axis(const uint32_t* __restrict__ st, int64_t rel, const PdltEntry& e, int d,
                          const PdltParams& P, T* __restrict__ rc,
                          const uint32_t* __restrict__ walk, uint32_t n_emit) {
  if (n_emit == 0) return;   // no residual section is read (n_int may be 0)
  const uint32_t n_bnd = e.n_bnd;
  constexpr bool c32 = sizeof(T) == 4;
  double fb;
  if (c32) {
    float s = 0.f;
    for (uint32_t i = 0; i < n_bnd; ++i) s = s + (float)rc[i];
    fb = n_bnd ? (double)(s / (float)n_bnd) : 0.0;
  } else {
    double s = 0.0;
    for (uint32_t i = 0; i < n_bnd; ++i) s = s + (double)rc[i];
    fb = n_bnd ? s / (double)n_bnd : 0.0;
  }
  // a mask blend, not a select of three loads: that select folds into an
  // indexed load, which keeps the caller's entry out of registers
  const uint32_t m1 = 0u - (uint32_t)(d == 1), m2 = 0u - (uint32_t)(d == 2);
  const uint32_t ab = (e.axis_bit[0] & ~(m1 | m2)) | (e.axis_bit[1] & m1) | (e.axis_bit[2] & m2);
  uint32_t p = (uint32_t)((int64_t)ab + rel);
  const uint64_t H = win64(st, p);
  const uint32_t tag = (uint32_t)(H >> 56);
  int32_t mn = 0;
  uint32_t bw = 0, k = 0;
  if (tag == 0) {
    mn = (int32_t)(int16_t)(uint16_t)(H >> 40);
    bw = (uint32_t)(H >> 32) & 0xFFu;
    p += 32;
  } else {
    k = (uint32_t)(H >> 48) & 0xFFu;
    p += 16;
  }
  const double w0 = P.w[0], w1 = P.w[1], w2 = P.w[2], delta = P.delta;
  const float w0f = (float)w0, w1f = (float)w1, w2f = (float)w2, deltaf = (float)delta;
  // the codes stream through a window with the next word read ahead, so no
  // step waits on an LDS read for its bits (the staged range holds 6 words
  // past the meshlet's end for the read-ahead); a codeword of at most 32
  // bits decodes from the window in one step (tag 1 Rice, 2 Exp-Golomb: the
  // coders of DPR4's residuals), a longer one bit-serially
  BitStream<LdsStaged> bs;
  bs.init(LdsStaged{st, 0}, p);
  const bool eg = tag != 1;
  // A step reads its three reconstructions first, then the next walk entry
  // (walk[n_emit] lies in the block: the seen set follows the walk), then
  // decodes its code, so the reads' latency passes under the decode; a
  // strip root's entry names v three times, whose reads are not used (a
  // select, not a branch around them).  The window refills without a
  // branch.
  uint32_t ent = walk[0];
  for (uint32_t i = 0; i < n_emit; ++i) {
    const uint32_t v = ent & 0xFFu, a = (ent >> 8) & 0xFFu;
    const uint32_t b = (ent >> 16) & 0xFFu, c = ent >> 24;
    const T xa = rc[a], xb = rc[b], xc = rc[c];
    ent = walk[i + 1];
    int32_t code;
    if (tag == 0) {
      code = (int32_t)((int64_t)mn + (int64_t)(bw ? (uint32_t)(bs.win >> (64 - bw)) : 0u));
      bs.drop_nb(bw);
    } else {
      uint32_t u;
      const uint32_t len = codeword_win(bs.win, eg ? ~0u : 0u, (int)k, u);
      if (len) {
        bs.drop_nb(len);
        code = (int32_t)((u >> 1) ^ (0u - (u & 1u)));
      } else {
        code = (int32_t)codeword_s(bs, eg, (int)k);
      }
    }
    if constexpr (c32) {
      const float t0 = w0f * (float)xa;
      const float t1 = w1f * (float)xb;
      const float t2 = w2f * (float)xc;
      const float lin = (t0 + t1) + t2;
      const float pred = a != v ? lin : (float)fb;
      rc[v] = (T)(pred + (float)code * deltaf);
    } else {
      const double t0 = w0 * (double)xa;
      const double t1 = w1 * (double)xb;
      const double t2 = w2 * (double)xc;
      const double lin = (t0 + t1) + t2;
      const double pred = a != v ? lin : fb;
      rc[v] = (T)(pred + (double)code * delta);
    }
  }
}

// Boundary-table position as the chain starts from it.
STRIDE_DEV double pdlt_bnd(double x, const PdltParams& P) {
  return P.chain32 ? (double)(float)x : x;
}

// Output value of local v on axis d (interior locals never reached by a
// strip keep position 0, as in the reference decoder).
STRIDE_DEV float pdlt_out(double x, const PdltParams& P, int d) {
  const double c = d == 0 ? P.center[0] : (d == 1 ? P.center[1] : P.center[2]);
  return P.world ? (float)(x * P.scale + c) : (float)x;
}

}  // namespace stride
