// Per-lane STRIDE decode logic shared by the gfx950 kernels and the CPU
// harness that checks it against the oracle (compiled as host C++ there).
#pragma once

#include <cmath>
#include <cstdint>

#include "stride_bits.h"
#include "stride_device.h"

namespace stride {

// Integer-rational predictor (paradelta_v5_dup.py:91-104): floor shift.
STRIDE_DEV int64_t irlp64(int64_t a, int64_t b, int64_t c, int32_t n0, int32_t n1,
                                          int32_t n2, int32_t K) {
  int64_t s = (int64_t)n0 * a + (int64_t)n1 * b + (int64_t)n2 * c;
  if (K != 0) s = (s + ((int64_t)1 << (K - 1))) >> K;
  return s;
}

STRIDE_DEV bool fits32(int64_t v) { return v == (int64_t)(int32_t)v; }

// ---------------------------------------------------------------------------
// Group decode: one wave per group of consecutive meshlets, one lane per
// meshlet.  Each lane runs the whole serial decode of its meshlet (strip
// tokens, reuse FIFO, residual codewords, IRLP chain) in a single token
// loop, so one wave instruction advances up to 64 meshlets.  Codes live in
// LDS as three 21-bit fields per vertex (bit 63 marks "not decoded yet"),
// so the seen test and the code lookup are one 8-byte LDS read.
// ---------------------------------------------------------------------------
constexpr uint64_t kUnseen = 1ull << 63;
constexpr int64_t kCodeLim = 1 << 20;   // packed field range [-2^20, 2^20)

STRIDE_DEV uint64_t pack21(int64_t x, int64_t y, int64_t z) {
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)x & m) | (((uint64_t)y & m) << 21) | (((uint64_t)z & m) << 42);
}
STRIDE_DEV int32_t field21(uint64_t c, int f) {
  return (int32_t)((int64_t)(c << (43 - 21 * f)) >> 43);
}
STRIDE_DEV bool fits21(int64_t v) { return v >= -kCodeLim && v < kCodeLim; }

// 64 bits of the staged stream starting at bit p (at least 33 valid).
STRIDE_DEV uint64_t win64(const uint32_t* __restrict__ st, uint32_t p) {
  const uint32_t w = p >> 5;
  const uint64_t x = ((uint64_t)st[w] << 32) | st[w + 1];
  return x << (p & 31);
}

// Codeword at bit p: Rice(k) (eg == false) or Exp-Golomb(k), zigzag
// decoded; p advances past it.  Codewords of at most 33 bits come from one
// window; longer ones (only for outliers) take the bit-serial path.
STRIDE_DEV int64_t codeword(const uint32_t* __restrict__ st, uint32_t n_words,
                                            uint32_t& p, bool eg, int k, bool needed) {
  const uint64_t W = win64(st, p);
  const uint32_t q = W ? (uint32_t)__builtin_clzll(W) : 64u;
  const uint32_t extra = eg ? q : 0u;
  const uint32_t nb = extra + (uint32_t)k;
  const uint32_t len = q + 1 + nb;
  uint64_t u;
  if (len <= 33) {
    const uint64_t body = nb ? ((W << (q + 1)) >> (64 - nb)) : 0ull;
    const uint64_t off = body >> k;
    const uint64_t low = body & ((1ull << k) - 1);
    const uint64_t shifted = eg ? (((1ull << q) - 1) + off) : (uint64_t)q;
    u = (shifted << k) | low;
    p += len;
  } else if (needed) {
    BitReader<LdsWords> r;
    r.init(LdsWords{st, n_words}, p);
    u = eg ? r.exp_golomb(k, 0xFFFFFFu) : r.rice(k, 0xFFFFFFu);
    p = r.pos;
  } else {
    u = 0;
  }
  return unzigzag(u);
}

// Serial decode of one meshlet by one lane (decode_dup's walk fused with
// the token stream).  Returns false when a code leaves the packed 21-bit
// range (the group is then decoded by the wave-per-meshlet kernel).
STRIDE_DEV bool decode_lane(const uint32_t* __restrict__ st, uint32_t n_words,
                                            uint32_t rel, const MeshletEntry& e,
                                            const MeshParams& P, uint64_t* __restrict__ codes,
                                            uint32_t* __restrict__ tout) {
  const uint32_t n_local = e.n_local, n_tris = e.n_tris;
  const uint32_t n_strips = e.n_strips & kStripMask;
  const bool raw_in_fifo = (e.n_strips & kFlagRawInFifo) != 0;
  // anchor and the six (tag, k) bytes
  uint32_t hp = rel + (e.resid_bit - e.start_bit);
  int32_t anc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int gb = P.g_bits[d];
    anc[d] = (int32_t)(win64(st, hp) >> (64 - gb));
    hp += gb;
  }
  const uint64_t TK = win64(st, hp);            // 32 bits: dx dy
  const uint64_t TK2 = win64(st, hp + 32);      // 32 bits: dz px
  const uint64_t TK3 = win64(st, hp + 64);      // 32 bits: py pz
  auto tagk = [&](int j, bool& eg, int& k) {
    const uint64_t src = j < 2 ? TK : (j < 4 ? TK2 : TK3);
    const int sh = 48 - 16 * (j & 1);
    eg = ((src >> (sh + 8)) & 0xFF) != 1;
    k = (int)((src >> sh) & 31);
  };
  bool deg[3], peg[3];
  int dk[3], pk[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    tagk(d, deg[d], dk[d]);
    tagk(3 + d, peg[d], pk[d]);
  }
  uint32_t pd[3], pp[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pd[d] = rel + (e.sub_bit[d] - e.start_bit);
    pp[d] = rel + (e.sub_bit[3 + d] - e.start_bit);
  }
  uint32_t pc = rel + 48;
  int ib = 1;
  while ((1u << ib) < n_local + 1) ++ib;
  Mtf fifo;
  fifo.clear();
  bool ok = true, first = true;
  int32_t prv[3] = {0, 0, 0};                // previous walk entry
  uint32_t q0 = 0, q1 = 0, q2 = 0;            // previous triangle ids
  int32_t c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0};
  uint32_t root_j = 3, steps_left = 0, tc = 0;
  const uint32_t v_off = e.v_off;
  const uint32_t n_tok = n_tris + 2 * n_strips;
  for (uint32_t it = 0; it < n_tok; ++it) {
    // strip header (when a strip starts), edge bit, token: one window
    uint64_t W = win64(st, pc);
    const bool start = root_j == 3 && steps_left == 0;
    const uint32_t L = (uint32_t)(W >> 48);
    W = start ? (W << 16) : W;
    steps_left = start ? L - 1 : steps_left;
    root_j = start ? 0u : root_j;
    const bool is_root = root_j < 3;
    const uint32_t ec = is_root ? 0u : (uint32_t)(W >> 63);
    W = is_root ? W : (W << 1);
    const uint32_t flag = (uint32_t)(W >> 63);
    W <<= 1;
    const uint32_t val = flag ? (uint32_t)(W >> (64 - ib)) : (uint32_t)(W >> 59);
    pc += (start ? 16u : 0u) + (is_root ? 0u : 1u) + 1u + (flag ? (uint32_t)ib : 5u);
    // reuse FIFO: a hit names FIFO index val (oldest first), a raw id is
    // appended (the oldest entry leaving a full FIFO); either way the value
    // becomes the most recent entry
    const bool hit = flag == 0;
    uint32_t j = hit ? ((fifo.n - 1 - (val & 15u)) & 15u) : fifo.n;
    const uint32_t v = hit ? fifo.rank(j) : (val & 0xFFu);
    if (raw_in_fifo && !hit) {
      const int pf = fifo.find(v);
      j = pf >= 0 ? (uint32_t)pf : j;
    }
    fifo.n += (!hit && j == fifo.n && fifo.n < 16) ? 1u : 0u;
    fifo.to_front(j, v);
    // code of v: stored, or decoded now from its residual substreams
    const uint64_t cv = codes[v];
    const bool fresh = (cv & kUnseen) != 0;
    const bool para = !is_root;
    const bool anchor = is_root && first;
    const bool need = fresh && !anchor;
    int32_t code[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      uint32_t pos = para ? pp[d] : pd[d];
      const int64_t r = codeword(st, n_words, pos, para ? peg[d] : deg[d], para ? pk[d] : dk[d],
                                 need);
      pp[d] = (need && para) ? pos : pp[d];
      pd[d] = (need && !para) ? pos : pd[d];
      // refs (a, b, c): ec 0 -> (q1, q2, q0), ec 1 -> (q0, q2, q1)
      const int32_t ca = ec ? c0[d] : c1[d], cc = ec ? c1[d] : c0[d];
      const int64_t pred = irlp64(ca, c2[d], cc, P.pred_n[3 * d], P.pred_n[3 * d + 1],
                                  P.pred_n[3 * d + 2], P.pred_K[d]);
      const int64_t base = para ? pred : (anchor ? (int64_t)anc[d] : (int64_t)prv[d]);
      const int64_t cnew = base + (anchor ? 0 : r);
      ok &= !fresh || fits21(cnew);
      code[d] = fresh ? (int32_t)cnew : field21(cv, d);
    }
    if (fresh) {
      codes[v] = pack21(code[0], code[1], code[2]);
#pragma unroll
      for (int d = 0; d < 3; ++d) prv[d] = code[d];
      first = false;
    }
    // triangle (a, b, v) and the next previous triangle
    uint32_t ta, tb;
    if (is_root) {
      ta = q0;
      tb = q1;
      if (root_j == 0) {
        q0 = v;
#pragma unroll
        for (int d = 0; d < 3; ++d) c0[d] = code[d];
      } else if (root_j == 1) {
        q1 = v;
#pragma unroll
        for (int d = 0; d < 3; ++d) c1[d] = code[d];
        tb = v;
      } else {
        q2 = v;
#pragma unroll
        for (int d = 0; d < 3; ++d) c2[d] = code[d];
      }
      ++root_j;
    } else {
      ta = ec ? q0 : q1;
      tb = q2;
      if (ec == 0) {
        q0 = q1;
#pragma unroll
        for (int d = 0; d < 3; ++d) c0[d] = c1[d];
      }
      q1 = q2;
      q2 = v;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        c1[d] = c2[d];
        c2[d] = code[d];
      }
      --steps_left;
    }
    if (root_j == 3 || !is_root) {
      uint32_t* o = tout + 3 * (tc & (kMaxTris - 1));
      o[0] = v_off + ta;
      o[1] = v_off + tb;
      o[2] = v_off + v;
      ++tc;
    }
  }
  return ok;
}


// ---------------------------------------------------------------------------
// Split decode: four lanes per meshlet.  A token lane decodes the strip
// tokens (reuse FIFO, triangles) and records the walk -- the local ids in
// first-seen order with their parallelogram references -- in LDS; then one
// lane per axis decodes that axis's two residual substreams along the walk
// and runs its prediction chain.  The serial path of a wave is the token
// loop plus one axis chain instead of the token loop with all three axes
// folded in.
//
// Per meshlet the LDS block holds the walk (n_local words) and the three
// code arrays: int32 (3 n_local words) or uint16 (split_code_words(n)
// words, each axis array padded to an even count).
//
// Walk entry i: v * 0x01010101 for a strip-root vertex (anchor or delta),
// v | a << 8 | b << 16 | c << 24 for a parallelogram vertex.  A walk entry
// is a first sighting of v while a is a vertex of the previous triangle,
// seen before, so a != v and the byte test of walk_root never confuses the
// two forms.  A root's prediction reads codes[v] (in range, discarded).
// ---------------------------------------------------------------------------
STRIDE_DEV bool walk_root(uint32_t w) { return ((w >> 8) & 0xFFu) == (w & 0xFFu); }


// 32-bit words of one meshlet's split block (walk + three code arrays).
STRIDE_DEV uint32_t split_block_words(uint32_t n_local, bool narrow) {
  return narrow ? n_local + 3 * ((n_local + 1) / 2) : 4 * n_local;
}
// elements between the starts of two consecutive axis code arrays
STRIDE_DEV uint32_t split_code_stride(uint32_t n_local, bool narrow) {
  return narrow ? 2 * ((n_local + 1) / 2) : n_local;
}

// 32 bits of the staged stream starting at bit p.
STRIDE_DEV uint32_t win32(const uint32_t* __restrict__ st, uint32_t p) {
  const uint32_t w = p >> 5, sh = p & 31u;
  return (st[w] << sh) | ((st[w + 1] >> 1) >> (31u - sh));
}

// Seen sets of the token lane: fresh(v) is true on the first sighting of
// local id v and marks it.
// In LDS: n_local zeroed entries of scratch (the x-axis code array, which the
// axis lanes overwrite afterwards).
template <class T>
struct ScratchSeen {
  T* __restrict__ p;
  STRIDE_DEV bool fresh(uint32_t v) {
    const bool f = p[v] == 0;
    p[v] = 1;
    return f;
  }
};
// In memory owned by one lane (LDS in the kernels): 8 words, updated with an
// atomic or whose returned old word tells whether v was seen -- one LDS
// instruction per token instead of a select over registers.
// Lanes of a token wave interleave their seen sets at this stride (words).
constexpr uint32_t kTokenSeenStride = 64;
struct BitSeen {
  uint32_t* __restrict__ w;
  // words between consecutive 32-id blocks: a wave's lanes interleave their
  // sets (lane i at w + i, stride = lanes) so one access touches distinct banks
  uint32_t stride = 1;
  // ids [0, n) start as seen
  STRIDE_DEV void preset(uint32_t n) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t lo = 32u * (uint32_t)q;
      w[q * stride] = n >= lo + 32 ? ~0u : (n <= lo ? 0u : ((1u << (n - lo)) - 1u));
    }
  }
  STRIDE_DEV bool fresh(uint32_t v) {
    const uint32_t bit = 1u << (v & 31u);
    const uint32_t old = __atomic_fetch_or(w + ((v >> 5) & 7u) * stride, bit, __ATOMIC_RELAXED);
    return (old & bit) == 0;
  }
};

// Bit cursors of the token lane: peek32() = the next 32 bits, advance(n)
// consumes n <= 27 of them.
struct LdsCursor {            // staged words in LDS
  const uint32_t* __restrict__ st;
  uint32_t pc;
  STRIDE_DEV uint32_t peek32() const { return win32(st, pc); }
  STRIDE_DEV void advance(uint32_t n) { pc += n; }
  // the next 32 bits as the two words holding them: read now, combined
  // (window()) where the bits are needed, so the reads' latency passes
  // under the work in between
  struct Raw {
    uint32_t a, b, sh;
  };
  STRIDE_DEV Raw peek_raw() const {
    const uint32_t w = pc >> 5;
    return Raw{st[w], st[w + 1], pc & 31u};
  }
  STRIDE_DEV static uint32_t window(const Raw& r) {
    return (r.a << r.sh) | ((r.b >> 1) >> (31u - r.sh));
  }
};

// Token lane: strip tokens -> triangles (tout) and the walk.  `cur` starts
// at the first strip length.
// Output ids of a DPR4 meshlet's locals (v_off + local).
struct MeshletIds {
  static constexpr bool kCarry = false;   // ids recomputed per token (an add)
  uint32_t v_off;
  STRIDE_DEV uint32_t operator()(uint32_t x) const { return v_off + x; }
};

// The token lane as a stepper: one strip token per step().  A first
// sighting's walk entry (local id + parallelogram refs, first-seen order;
// v * 0x01010101 for strip roots, walk_root) goes to sink(index, entry); a
// sink with kEveryToken takes every token's entry at the next index, which
// the next first sighting overwrites.  The triangles go to `o` as
// ids(local).  The split, two-pass and PDLT token kernels run steps back to
// back (split_tokens_c, pdlt_tokens).
template <class Ids>
struct TokenLane {
  uint32_t n_tok, it;      // tokens in the meshlet, tokens done
  uint32_t ib;
  uint32_t nw;             // walk entries emitted
  uint32_t q0, q1, q2, root_j, steps_left;
  uint32_t g0, g1, g2;     // ids(q0..q2) when Ids::kCarry (an id lookup reads LDS)
  uint32_t Wn;             // the next token's 32-bit window, read one token ahead
  bool raw_in_fifo;
  MtfLite fifo;
  Ids ids;
  uint32_t* __restrict__ o;

  // strips_flags: the strip count | kFlagRawInFifo
  template <class Cursor>
  STRIDE_DEV void init(Cursor& cur, uint32_t n_local, uint32_t n_tris, uint32_t strips_flags,
                       uint32_t* __restrict__ tout, const Ids& id_map) {
    n_tok = n_tris + 2 * (strips_flags & kStripMask);
    raw_in_fifo = (strips_flags & kFlagRawInFifo) != 0;
    ib = 1;
    while ((1u << ib) < n_local + 1) ++ib;
    fifo.clear();
    it = 0;
    nw = 0;
    q0 = q1 = q2 = 0;
    g0 = g1 = g2 = 0;
    root_j = 3;
    steps_left = 0;
    ids = id_map;
    o = tout;
    // strip length (16 bits) + edge bit + flag + id (<= 9 bits): one 32-bit
    // window per token.  The next token's window is read as soon as this
    // token's length is known, so its latency hides under the FIFO work; the
    // read after the last token stays inside the meshlet's anchor and
    // substream bits or the buffer space after them.
    Wn = cur.peek32();
  }
  STRIDE_DEV bool done() const { return it >= n_tok; }

  // kFind = false leaves out the FIFO search of raw ids, for waves none of
  // whose meshlets sets kFlagRawInFifo (the loops choose with wave_any)
  template <bool kFind = true, class Cursor, class Seen, class Sink>
  STRIDE_DEV void step(Cursor& cur, Seen& seen, Sink& sink) {
    uint32_t W = Wn;
    const bool start = root_j == 3 && steps_left == 0;
    // a strip of length L has max(L, 1) triangles (PDLT allows L = 0)
    const uint32_t L = W >> 16;
    steps_left = start ? (L ? L - 1 : 0u) : steps_left;
    W = start ? (W << 16) : W;
    root_j = start ? 0u : root_j;
    const bool is_root = root_j < 3;
    const uint32_t ec = is_root ? 0u : (W >> 31);
    W = is_root ? W : (W << 1);
    const uint32_t flag = W >> 31;
    const uint32_t val = flag ? ((W << 1) >> (32 - ib)) : ((W >> 26) & 31u);
    cur.advance((start ? 16u : 0u) + (is_root ? 1u : 2u) + (flag ? ib : 5u));
    Wn = cur.peek32();
    // reuse FIFO: a hit names FIFO index val (oldest first), a raw id is
    // appended (the oldest entry leaving a full FIFO); either way the value
    // becomes the most recent entry
    const bool hit = flag == 0;
    uint32_t j = hit ? ((fifo.n - 1 - val) & 15u) : fifo.n;
    const uint32_t v = hit ? fifo.rank(j) : (val & 0xFFu);
    if constexpr (kFind) {
      if (raw_in_fifo && !hit) {
        const int pf = fifo.find(v);
        j = pf >= 0 ? (uint32_t)pf : j;
      }
    }
    // triangle (a, b, v): its output ids are looked up as soon as v is known
    // (for PDLT v5 a boundary local's id is an LDS read, so a and b carry
    // theirs from the tokens that named them), and the lookup overlaps the
    // FIFO and seen-set work before the store
    const uint32_t ra = ec ? q0 : q1, rc = ec ? q1 : q0;
    const uint32_t iv = ids(v);
    uint32_t ia, ibb, ga = 0;
    if constexpr (Ids::kCarry) {
      ga = ec ? g0 : g1;
      ia = is_root ? g0 : ga;
      ibb = is_root ? g1 : g2;
    } else {
      ia = ids(is_root ? q0 : ra);
      ibb = ids(is_root ? q1 : q2);
    }
    fifo.n += (!hit && j == fifo.n && fifo.n < 16) ? 1u : 0u;
    fifo.to_front(j, v);
    // first sighting -> walk entry; parallelogram refs (a, b, c):
    // ec 0 -> (q1, q2, q0), ec 1 -> (q0, q2, q1)
    const bool fresh = seen.fresh(v);
    // the entry's three ref bytes are v for a root (v * 0x01010101),
    // selected byte by byte: a select of two whole entries compiles to a
    // branch
    const uint32_t ea = is_root ? v : ra, eb = is_root ? v : q2, ec2 = is_root ? v : rc;
    const uint32_t entry = v | (ea << 8) | (eb << 16) | (ec2 << 24);
    if constexpr (Sink::kEveryToken) {
      sink(nw, entry);   // a stale entry at nw is overwritten by the next first sighting
    } else {
      if (fresh) sink(nw, entry);
    }
    nw += fresh ? 1u : 0u;
    // the triangle, and the next previous triangle.  Every token stores one:
    // the first two roots of a strip store into the slot the third root
    // then fills, so only emitting tokens advance o.
    const bool emit = !is_root || root_j == 2;
    o[0] = ia;
    o[1] = ibb;
    o[2] = iv;
    o += emit ? 3 : 0;
    q0 = is_root ? (root_j == 0 ? v : q0) : ra;
    q1 = is_root ? (root_j == 1 ? v : q1) : q2;
    q2 = is_root ? (root_j == 2 ? v : q2) : v;
    if constexpr (Ids::kCarry) {
      g0 = is_root ? (root_j == 0 ? iv : g0) : ga;
      g1 = is_root ? (root_j == 1 ? iv : g1) : g2;
      g2 = is_root ? (root_j == 2 ? iv : g2) : iv;
    }
    root_j += is_root ? 1u : 0u;
    steps_left -= is_root ? 0u : 1u;
    ++it;
  }
};

// True when p holds on any active lane of the wave (on the host: this lane).
STRIDE_DEV bool wave_any(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __any(p ? 1 : 0) != 0;
#else
  return p;
#endif
}

// All steps of a token lane; the FIFO search is compiled in only when some
// lane of the wave needs it (kFlagRawInFifo).
template <class Lane, class Cursor, class Seen, class Sink>
STRIDE_DEV void tokens_run(Lane& t, Cursor& cur, Seen& seen, Sink& sink) {
  if (wave_any(t.raw_in_fifo)) {
    for (uint32_t it = 0; it < t.n_tok; ++it) t.template step<true>(cur, seen, sink);
  } else {
    for (uint32_t it = 0; it < t.n_tok; ++it) t.template step<false>(cur, seen, sink);
  }
}

// ---------------------------------------------------------------------------
// Chain token pass.  A meshlet's strip tokens are cut into chains of about
// equal length (chain_tokens); the chain-start scan (chain_starts) records the
// parser state at every cut.  Phase 1: one lane per chain parses its tokens
// (strip headers, edge bits, FIFO hits and raw ids) and stores one 16-bit
// record per token.  A chain after the meshlet's first one starts with a FIFO
// of placeholders: rank r holds n_local + r, so a hit on an entry the chain
// did not see itself records that placeholder, and the lane's final FIFO (its
// newest-first ranks) tells the next chain what its placeholders stand for.
// Phase 2 (a whole wave per meshlet) resolves the placeholders, finds first
// sightings, and forms triangles and walk entries from the records alone:
//
//   body token t, edge bit ec:  triangle (a, b, v_t), walk refs (a, b, c) with
//     b = v_{t-1},  ec 0: a = v_{t-2}, c = p0;  ec 1: a = p0, c = v_{t-2},
//     p0 = v_{s-2}, s the last token before t with edge bit 0 (strip roots
//     count as edge bit 0; a strip's third root precedes its body tokens).
//   third root: triangle (v_{t-2}, v_{t-1}, v_t), the ec-0 form.
//
// This is decode_dup's strip state (encoder/paradelta_v5_dup.py:512-542):
// prev_tri = (p0, p1, p2) with p2 = v_{t-1} and p1 = v_{t-2} at every body
// token, and p0 replaced by p1 exactly on edge bit 0.
// ---------------------------------------------------------------------------
constexpr uint32_t kRecEc = 1u << 8;       // record: edge bit
constexpr uint32_t kRecRoot = 1u << 9;     // strip root token
constexpr uint32_t kRecRoot2 = 1u << 10;   // third strip root (emits the first triangle)
constexpr uint32_t kMinChainTokens = 16;   // no chain shorter than this is cut off

// Tokens per chain of a meshlet with n_tok tokens cut into about k0 chains
// (k0 <= kMaxChains).  Meshlets with raw ids already in the FIFO, or too
// many locals for the placeholders, stay one chain.
STRIDE_DEV uint32_t chain_tokens(uint32_t n_tok, uint32_t n_local, uint32_t strips_flags,
                                 uint32_t k0) {
  const bool splittable = k0 > 1 && n_local <= (uint32_t)kChainMaxLocal &&
                          (strips_flags & kFlagRawInFifo) == 0;
  if (!splittable || n_tok <= kMinChainTokens) return n_tok > 0 ? n_tok : 1u;
  const uint32_t lm = (n_tok + k0 - 1) / k0;
  return lm < kMinChainTokens ? kMinChainTokens : lm;
}
STRIDE_DEV uint32_t chain_count(uint32_t n_tok, uint32_t lm) {
  return n_tok > 0 ? (n_tok + lm - 1) / lm : 1u;
}
STRIDE_DEV uint32_t meshlet_tokens(const MeshletEntry& e) {
  return e.n_tris + 2 * (e.n_strips & kStripMask);
}
// One past the last token of chain c of a meshlet of n_tok tokens cut every
// lm tokens.
STRIDE_DEV uint32_t chain_end(const ChainEntry& c, uint32_t n_tok, uint32_t lm) {
  const uint32_t t0 = c.pos >> 16;
  return t0 + lm < n_tok ? t0 + lm : n_tok;
}

// The parser state at every cut of a meshlet's n_tok strip tokens, cut
// every lm tokens (the first at bit tok_bit of src; ChainEntry.pos counts
// from bit `base`), written to out[0 .. chain_count).  A strip of length 0
// has three root tokens, as one of length 1 (PDLT v5 allows it; DPR4's
// pre-pass rejects it).  The stream was validated by the pre-pass.
template <class Src>
STRIDE_DEV void chain_starts_at(const Src& src, uint32_t base, uint32_t tok_bit, uint32_t n_tok,
                                uint32_t n_local, uint32_t lm, ChainEntry* __restrict__ out) {
  BitReader<Src> r;
  r.init(src, tok_bit);
  int ib = 1;
  while ((1u << ib) < n_local + 1) ++ib;
  Fifo fifo;
  fifo.clear();
  uint32_t j = 0, lp2 = 0, c = 0, next = 0;
  for (uint32_t t = 0; t < n_tok; ++t) {
    if (t == next) {
      ChainEntry ce;
      ce.pos = (r.pos - base) | (t << 16);
      ce.state = j | (lp2 << 9) | (fifo.n << 19) | (c << 24);
      next = chain_end(ce, n_tok, lm);
      out[c++] = ce;
    }
    if (j == lp2) {
      const uint32_t L = r.bits(16);
      lp2 = (L ? L : 1u) + 2u;
      j = 0;
    }
    if (j >= 3) r.bits(1);
    if (r.bits(1) == 0) {
      fifo.hit(r.bits(5));
    } else {
      fifo.raw(r.bits(ib));
    }
    ++j;
  }
  if (n_tok == 0) out[0] = ChainEntry{0u, 0u};   // one empty chain
}

// The chain cuts of DPR4 meshlet e (its header at bit `start` of src; the
// tokens follow the 48-bit header).
template <class Src>
STRIDE_DEV void chain_starts(const Src& src, uint32_t start, const MeshletEntry& e,
                             uint32_t k0, ChainEntry* __restrict__ out) {
  const uint32_t n_tok = meshlet_tokens(e);
  chain_starts_at(src, start, start + 48, n_tok, e.n_local,
                  chain_tokens(n_tok, e.n_local, e.n_strips, k0), out);
}

// Phase 1: the lane of one chain, over a Cursor (LdsCursor in the kernel).
// kZeroLen: strips of length 0 occur (PDLT v5) and take three tokens.
template <class Cursor, bool kZeroLen = false>
struct ChainLane {
  uint32_t t, t1;       // next token, one past the chain's last
  uint32_t j, lp2;      // token index in the strip, tokens in the strip (length + 2)
  uint32_t ib;
  uint32_t tag;         // chain index in the meshlet << 12
  typename Cursor::Raw Wn;   // the next token's window, read one token ahead
  bool raw_in_fifo;
  MtfLite fifo;

  // `cur` is at the chain's first token: the meshlet header + (c.pos & 0xFFFF);
  // `end` is one past its last token (chain_end).
  STRIDE_DEV void init(Cursor& cur, const ChainEntry& c, uint32_t end, uint32_t n_local,
                       uint32_t strips_flags) {
    t = c.pos >> 16;
    t1 = end;
    j = c.state & 511u;
    lp2 = (c.state >> 9) & 1023u;
    const uint32_t n = (c.state >> 19) & 31u;
    const uint32_t ci = (c.state >> 24) & 15u;
    tag = ci << 12;
    raw_in_fifo = (strips_flags & kFlagRawInFifo) != 0;
    ib = 1;
    while ((1u << ib) < n_local + 1) ++ib;
    fifo.clear();
    fifo.n = n;
    // placeholders: rank r holds n_local + r (chains after the first only;
    // the first chain starts with an empty FIFO)
    uint64_t lo = 0, hi = 0;
    for (uint32_t r = 0; r < 16; ++r) {
      const uint64_t b = (uint64_t)((n_local + r) & 0xFFu);
      if (r < 8) lo |= b << (8 * r);
      else hi |= b << (8 * (r - 8));
    }
    fifo.lo = ci ? lo : 0ull;
    fifo.hi = ci ? hi : 0ull;
    Wn = cur.peek_raw();
  }
  STRIDE_DEV bool done() const { return t >= t1; }

  template <bool kFind = true, class Sink>
  STRIDE_DEV void step(Cursor& cur, Sink& sink) {
    uint32_t W = Cursor::window(Wn);
    const bool start = j == lp2;
    const uint32_t L = W >> 16;
    lp2 = start ? (kZeroLen && L == 0 ? 1u : L) + 2u : lp2;
    W = start ? (W << 16) : W;
    j = start ? 0u : j;
    const bool root = j < 3;
    const uint32_t ec = root ? 0u : (W >> 31);
    W = root ? W : (W << 1);
    const uint32_t flag = W >> 31;
    const uint32_t val = flag ? ((W << 1) >> (32 - ib)) : ((W >> 26) & 31u);
    cur.advance((start ? 16u : 0u) + (root ? 1u : 2u) + (flag ? ib : 5u));
    Wn = cur.peek_raw();
    const bool hit = flag == 0;
    uint32_t jf = hit ? ((fifo.n - 1 - val) & 15u) : fifo.n;
    // the hit's entry, read on every token and blended in: a select here
    // compiles to a branch around the read
    const uint32_t hm = 0u - (uint32_t)hit;
    const uint32_t v = (fifo.rank(jf) & hm) | (val & 0xFFu & ~hm);
    if constexpr (kFind) {
      if (raw_in_fifo && !hit) {
        const int pf = fifo.find(v);
        jf = pf >= 0 ? (uint32_t)pf : jf;
      }
    }
    fifo.n += (!hit && jf == fifo.n && fifo.n < 16) ? 1u : 0u;
    fifo.to_front(jf, v);
    const uint32_t rec = v | (ec << 8) | (root ? kRecRoot : 0u) | (j == 2 ? kRecRoot2 : 0u) | tag;
    sink(t, rec);
    ++j;
    ++t;
  }
};

// Phase 2 of one token from its record and the ids of tokens t-1, t-2 and
// s-2 (see above): the triangle's first two ids (a, b) and the walk entry
// (v * 0x01010101 for a strip root, v | a << 8 | b << 16 | c << 24 else).
STRIDE_DEV uint32_t chain_entry(uint32_t rec, uint32_t v1, uint32_t v2, uint32_t p0, uint32_t& a,
                                uint32_t& b) {
  const uint32_t v = rec & 0xFFu;
  const bool ec = (rec & kRecEc) != 0;
  a = ec ? p0 : v2;
  b = v1;
  const uint32_t c = ec ? v2 : p0;
  // a root's entry blended in by mask (a select of the two entries compiles
  // to a branch)
  const uint32_t rm = 0u - ((rec >> 9) & 1u);
  return (byte4(v) & rm) | ((v | (a << 8) | (b << 16) | (c << 24)) & ~rm);
}
STRIDE_DEV bool chain_emits(uint32_t rec) {
  return (rec & kRecRoot) == 0 || (rec & kRecRoot2) != 0;
}

// Walk sinks: the whole walk in one array.  WalkArray stores first
// sightings only (the split kernel's LDS block has no room past n_local);
// WalkSlot stores every token's entry at the next index, which needs one
// word past the walk (the two-pass slots have it: walk_slot).
struct WalkArray {
  static constexpr bool kEveryToken = false;
  uint32_t* __restrict__ walk;
  STRIDE_DEV void operator()(uint32_t i, uint32_t entry) const { walk[i & 255u] = entry; }
};
struct WalkSlot {
  static constexpr bool kEveryToken = true;
  uint32_t* __restrict__ walk;
  STRIDE_DEV void operator()(uint32_t i, uint32_t entry) const { walk[i] = entry; }
};

template <class Sink = WalkArray, class Cursor, class Seen>
STRIDE_DEV void split_tokens_c(Cursor& cur, const MeshletEntry& e, uint32_t* __restrict__ walk,
                               Seen& seen, uint32_t* __restrict__ tout) {
  TokenLane<MeshletIds> t;
  t.init(cur, e.n_local, e.n_tris, e.n_strips, tout, MeshletIds{e.v_off});
  Sink sink{walk};
  tokens_run(t, cur, seen, sink);
}

template <class Sink = WalkArray, class Seen>
STRIDE_DEV void split_tokens(const uint32_t* __restrict__ st, uint32_t rel, const MeshletEntry& e,
                             uint32_t* __restrict__ walk, Seen& seen,
                             uint32_t* __restrict__ tout) {
  LdsCursor cur{st, rel + 48};
  split_tokens_c<Sink>(cur, e, walk, seen, tout);
}

STRIDE_DEV uint32_t pick3(int d, uint32_t x0, uint32_t x1, uint32_t x2) {
  return d == 0 ? x0 : (d == 1 ? x1 : x2);
}

// Code storage of the split decoder: int32, or uint16 when every axis of
// the mesh has at most 16 bits (codes of a well-formed stream lie in
// [0, 2^g_bits - 1]).
template <class CodeT>
STRIDE_DEV bool code_fits(int64_t c) {
  return sizeof(CodeT) == 2 ? (c >= 0 && c <= 0xFFFF) : fits32(c);
}

// Axis lane d: anchor, the delta and parallelogram substreams of axis d,
// and the chain along the walk; codes[v] receives the code of local id v.
// Returns false when a code leaves the range of CodeT.
template <class CodeT>
STRIDE_DEV bool split_axis(const uint32_t* __restrict__ st, uint32_t n_words, uint32_t rel,
                           const MeshletEntry& e, const MeshParams& P, int d,
                           const uint32_t* __restrict__ walk, CodeT* __restrict__ codes) {
  const uint32_t n_local = e.n_local;
  const uint32_t gb0 = (uint32_t)P.g_bits[0], gb1 = (uint32_t)P.g_bits[1],
                 gb2 = (uint32_t)P.g_bits[2];
  const uint32_t gb = pick3(d, gb0, gb1, gb2);
  const uint32_t r0 = rel + (e.resid_bit - e.start_bit);
  const uint32_t ap = r0 + (d > 0 ? gb0 : 0u) + (d > 1 ? gb1 : 0u);
  const int64_t anchor = (int64_t)(win64(st, ap) >> (64 - gb));
  // (tag, k) of delta-d at tag byte pair d, of para-d at pair 3 + d
  const uint32_t tb = r0 + gb0 + gb1 + gb2;
  const uint64_t TD = win64(st, tb + 16u * (uint32_t)d);
  const uint64_t TP = win64(st, tb + 16u * (uint32_t)(3 + d));
  const bool deg = (uint32_t)(TD >> 56) != 1u, peg = (uint32_t)(TP >> 56) != 1u;
  const int dk = (int)((TD >> 48) & 31u), pk = (int)((TP >> 48) & 31u);
  uint32_t pd = rel + (pick3(d, e.sub_bit[0], e.sub_bit[1], e.sub_bit[2]) - e.start_bit);
  uint32_t pp = rel + (pick3(d, e.sub_bit[3], e.sub_bit[4], e.sub_bit[5]) - e.start_bit);
  const int32_t n0 = (int32_t)pick3(d, (uint32_t)P.pred_n[0], (uint32_t)P.pred_n[3], (uint32_t)P.pred_n[6]);
  const int32_t n1 = (int32_t)pick3(d, (uint32_t)P.pred_n[1], (uint32_t)P.pred_n[4], (uint32_t)P.pred_n[7]);
  const int32_t n2 = (int32_t)pick3(d, (uint32_t)P.pred_n[2], (uint32_t)P.pred_n[5], (uint32_t)P.pred_n[8]);
  const int32_t K = (int32_t)pick3(d, (uint32_t)P.pred_K[0], (uint32_t)P.pred_K[1], (uint32_t)P.pred_K[2]);
  bool ok = true;
  int64_t prev = anchor;
  codes[walk[0] & 0xFFu] = (CodeT)anchor;
  ok &= code_fits<CodeT>(anchor);
  for (uint32_t i = 1; i < n_local; ++i) {
    const uint32_t w = walk[i];
    const bool root = walk_root(w);
    uint32_t pos = root ? pd : pp;
    const int64_t r = codeword(st, n_words, pos, root ? deg : peg, root ? dk : pk, true);
    pd = root ? pos : pd;
    pp = root ? pp : pos;
    const uint32_t ia = (w >> 8) & 0xFFu, ib = (w >> 16) & 0xFFu, ic = w >> 24;
    const int64_t pred = irlp64((int64_t)codes[ia], (int64_t)codes[ib], (int64_t)codes[ic], n0, n1,
                                n2, K);
    const int64_t c = (root ? prev : pred) + r;
    ok &= code_fits<CodeT>(c);
    codes[w & 0xFFu] = (CodeT)c;
    prev = c;
  }
  return ok;
}

// ---------------------------------------------------------------------------
// Streaming reads from the stream in global memory (the two-pass axis
// lanes): no LDS staging.  A lane reads its substream sequentially through a
// 64-bit window holding at least 33 valid bits, with the next word already
// requested, so the load latency overlaps the decoding of the bits before
// it.  The stream buffer is zero padded past its end, so the look-ahead load
// stays inside the allocation.
// ---------------------------------------------------------------------------
STRIDE_DEV uint32_t load_be(const uint32_t* __restrict__ g, uint64_t i) {
  return __builtin_bswap32(g[i]);
}

// Word sources of a bit stream, indexed by absolute word of the device
// buffer: the little-endian buffer itself (byte-swapped per read), or words
// staged MSB-first in LDS from word `lo` on.
struct BufferWords {
  const uint32_t* __restrict__ g;
  STRIDE_DEV uint32_t operator()(uint64_t i) const { return load_be(g, i); }
};
struct LdsStaged {
  const uint32_t* __restrict__ s;
  uint64_t lo;
  STRIDE_DEV uint32_t operator()(uint64_t i) const { return s[i - lo]; }
};

// n <= 32 bits at absolute bit position p of a word source.
template <class Src>
STRIDE_DEV uint32_t bits_at_src(const Src& src, uint64_t p, uint32_t n) {
  const uint64_t w = p >> 5;
  const uint64_t x = (((uint64_t)src(w) << 32) | src(w + 1)) << (p & 31);
  return n ? (uint32_t)(x >> (64 - n)) : 0u;
}
// n <= 32 bits at absolute bit position p of the little-endian words g.
STRIDE_DEV uint32_t bits_at(const uint32_t* __restrict__ g, uint64_t p, uint32_t n) {
  return bits_at_src(BufferWords{g}, p, n);
}

// A sequential reader: a 64-bit window with >= 33 valid bits and the next
// word read ahead, so a refill never waits on the read it needs.
template <class Src>
struct BitStream {
  Src src;
  uint64_t win;      // valid bits, top aligned; zeros below
  uint64_t next;     // index of the next word to request
  uint32_t valid;    // 33..64 between reads
  uint32_t ahead;    // word next - 1, not yet in win

  STRIDE_DEV void init(const Src& s, uint64_t bit) {
    src = s;
    const uint64_t w0 = bit >> 5;
    const uint32_t sh = (uint32_t)(bit & 31);
    win = (((uint64_t)src(w0) << 32) | src(w0 + 1)) << sh;
    valid = 64 - sh;
    ahead = src(w0 + 2);
    next = w0 + 3;
    fill();
  }
  // one word at most: a drop of n <= 32 bits leaves valid >= 1
  STRIDE_DEV void fill() {
    if (valid <= 32) {
      win |= (uint64_t)ahead << (32 - valid);
      valid += 32;
      ahead = src(next++);
    }
  }
  // drops n <= 32 bits
  STRIDE_DEV void drop(uint32_t n) {
    win = n < 64 ? (win << n) : 0ull;
    valid -= n;
    fill();
  }
  // fill() and drop() without a branch: the word after the read-ahead one
  // is read on every call (cheap from LDS; it lies no further than the next
  // fill() would read) and taken in only when the window has room
  STRIDE_DEV void fill_nb() {
    const bool f = valid <= 32;
    const uint32_t nx = src(next);
    win |= f ? (uint64_t)ahead << (32 - valid) : 0ull;
    valid += f ? 32u : 0u;
    ahead = f ? nx : ahead;
    next += f ? 1u : 0u;
  }
  STRIDE_DEV void drop_nb(uint32_t n) {
    win = n < 64 ? (win << n) : 0ull;
    valid -= n;
    fill_nb();
  }
  STRIDE_DEV uint32_t bits(uint32_t n) {   // n <= 32
    const uint32_t v = n ? (uint32_t)(win >> (64 - n)) : 0u;
    drop(n);
    return v;
  }
  // token-lane cursor interface
  STRIDE_DEV uint32_t peek32() const { return (uint32_t)(win >> 32); }
  STRIDE_DEV void advance(uint32_t n) { drop(n); }
  using Raw = uint32_t;
  STRIDE_DEV Raw peek_raw() const { return peek32(); }
  STRIDE_DEV static uint32_t window(Raw r) { return r; }
};

// The stream over the little-endian device buffer.
struct StreamBits : BitStream<BufferWords> {
  STRIDE_DEV void init(const uint32_t* __restrict__ words, uint64_t bit) {
    BitStream<BufferWords>::init(BufferWords{words}, bit);
  }
};

// codeword() on a stream.
// A unary prefix (zeros, then a one) of any length from a stream: the
// window's top 32 bits (always valid) at a time.  Returns the zero count.
template <class Stream>
STRIDE_DEV uint32_t unary_s(Stream& s) {
  uint32_t q = 0;
  for (;;) {
    const uint32_t top = (uint32_t)(s.win >> 32);
    if (top) {
      const uint32_t z = (uint32_t)__builtin_clz(top);
      s.drop(z + 1);
      return q + z;
    }
    q += 32;
    s.drop(32);
  }
}

template <class Stream>
STRIDE_DEV int64_t codeword_s(Stream& s, bool eg, int k) {
  const uint64_t W = s.win;
  const uint32_t q = W ? (uint32_t)__builtin_clzll(W) : 64u;
  const uint32_t nb = (eg ? q : 0u) + (uint32_t)k;
  const uint32_t len = q + 1 + nb;
  uint64_t u;
  if (len <= 32) {
    const uint64_t body = nb ? ((W << (q + 1)) >> (64 - nb)) : 0ull;
    const uint64_t off = body >> k;
    const uint64_t low = body & ((1ull << k) - 1);
    const uint64_t shifted = eg ? (((1ull << q) - 1) + off) : (uint64_t)q;
    u = (shifted << k) | low;
    s.drop(len);
  } else {
    // long codeword (outliers): the unary part 32 bits at a time, then the rest
    const uint32_t z = unary_s(s);
    uint64_t shifted = z;
    if (eg) shifted = z ? (((1ull << z) - 1) + s.bits(z)) : 0ull;
    const uint64_t low = k ? s.bits((uint32_t)k) : 0ull;
    u = (shifted << k) | low;
  }
  return unzigzag(u);
}

// The short form of codeword_s on a window W (>= 33 valid bits): the
// codeword's length with u = its zigzagged value, or 0 for a codeword longer
// than 32 bits (decoded by codeword_s instead; a stream then refills one
// word at most per drop).
// Within 32 bits everything fits 32-bit arithmetic on the window's top word
// T: q <= 31, the body (nb <= 31 bits) is the low nb bits of the top len
// bits of T, and u < 2^32 (a Rice code has (q + 1) << k <= 2^31, an EG code
// u < 2^(q + 1 + k) <= 2^32).  The shift counts are masked to 5 bits: they
// are exact whenever len <= 32, and the value is dropped otherwise.
// egm: all ones for Exp-Golomb(k), 0 for Rice(k) (a mask, not a bool: the
// coder choice then costs an and and a bit-field insert, no compares).
// Leading zeros of t, 32 for t == 0.  v_ffbh_u32 returns all ones for 0, so
// one min replaces the zero test, which the compiler otherwise emits as a
// 64-bit compare on the window, a select and a hazard wait.
STRIDE_DEV uint32_t clz_or_32(uint32_t t) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t q;
  asm("v_ffbh_u32 %0, %1" : "=v"(q) : "v"(t));
  return q < 32u ? q : 32u;
#else
  return t ? (uint32_t)__builtin_clz(t) : 32u;
#endif
}

STRIDE_DEV uint32_t codeword_win(uint64_t W, uint32_t egm, int k, uint32_t& u) {
  const uint32_t T = (uint32_t)(W >> 32);
  const uint32_t q = clz_or_32(T);
  const uint32_t nb = (q & egm) + (uint32_t)k;
  const uint32_t len = q + 1 + nb;
  const uint32_t x = T >> ((32u - len) & 31u);
  const uint32_t body = x & ((1u << (nb & 31u)) - 1u);
  const uint32_t off = body >> k;
  const uint32_t low = body & ((1u << k) - 1u);
  const uint32_t eg_val = ((1u << (q & 31u)) - 1u) + off;
  const uint32_t shifted = (eg_val & egm) | (q & ~egm);
  u = (shifted << k) | low;
  return len <= 32 ? len : 0u;
}
// codeword_win for Rice(k) alone: q = the zero count, the k low bits right
// after the terminating one.  Shift counts are masked to 5 bits: exact
// whenever len <= 32, and the value is dropped otherwise.
STRIDE_DEV uint32_t codeword_rice(uint64_t W, int k, uint32_t& u) {
  const uint32_t T = (uint32_t)(W >> 32);
  const uint32_t q = clz_or_32(T);
  const uint32_t len = q + 1 + (uint32_t)k;
  const uint32_t low = (T >> ((32u - len) & 31u)) & ((1u << (k & 31)) - 1u);
  u = (q << (k & 31)) | low;
  return len <= 32 ? len : 0u;
}
STRIDE_DEV int64_t unzigzag32(uint32_t u) {
  return (int64_t)(int32_t)((u >> 1) ^ (0u - (u & 1u)));
}

// IRLP in 32-bit arithmetic (24-bit multiplies): equal to irlp64 when the
// codes are 16-bit unsigned and irlp32_exact(n0, n1, n2, K) holds.
// The int16 numerators and 16-bit codes make every product a 24-bit multiply.
STRIDE_DEV int32_t irlp24(int32_t a, int32_t b, int32_t c, int32_t n0, int32_t n1, int32_t n2,
                          int32_t K, int32_t rnd) {
  return ((int32_t)(int16_t)n0 * a + (int32_t)(int16_t)n1 * b + (int32_t)(int16_t)n2 * c + rnd) >>
         K;
}
// int16 numerators, K <= 30 and sum |n_i| * 65535 + 2^(K-1) < 2^31: no
// product or partial sum of irlp24 leaves int32.
STRIDE_DEV bool irlp32_exact(int32_t n0, int32_t n1, int32_t n2, int32_t K) {
  auto mag = [](int32_t x) { return (int64_t)(x < 0 ? -(int64_t)x : (int64_t)x); };
  const bool small = n0 == (int16_t)n0 && n1 == (int16_t)n1 && n2 == (int16_t)n2;
  const int64_t top = (mag(n0) + mag(n1) + mag(n2)) * 65535 + (K > 0 ? (1ll << (K - 1)) : 0);
  return small && K >= 0 && K <= 30 && top <= 0x7FFFFFFFll;
}

// True when both substreams of axis d of meshlet e are Rice coded (tag 1).
template <class Src>
STRIDE_DEV bool axis_rice(const Src& words, uint64_t base, const MeshletEntry& e,
                          const MeshParams& P, int d) {
  const uint64_t tb = base + e.resid_bit + (uint32_t)P.g_bits[0] + (uint32_t)P.g_bits[1] +
                      (uint32_t)P.g_bits[2];
  const uint32_t TD = bits_at_src(words, tb + 16u * (uint32_t)d, 8);
  const uint32_t TP = bits_at_src(words, tb + 16u * (uint32_t)(3 + d), 8);
  return TD == 1u && TP == 1u;
}

// The words an axis lane pass reads of meshlet e: from the word holding its
// residual header (resid_bit) through 5 words past the one holding end_bit
// (the three read-ahead words of a stream that ends there, and one more).
STRIDE_DEV uint64_t axis_stage_lo(uint64_t blob_bit, const MeshletEntry& e) {
  return (blob_bit + e.resid_bit) >> 5;
}
STRIDE_DEV uint32_t axis_stage_words(uint64_t blob_bit, const MeshletEntry& e) {
  return (uint32_t)(((blob_bit + e.end_bit) >> 5) + 5 - axis_stage_lo(blob_bit, e));
}

// code / mx rounded to nearest in float64, from r = RN(1 / mx): q0 = RN(code
// r) lies within an ulp of the quotient, the remainder code - q0 mx is exact
// (an fma), and one correction step rounds it correctly.  Checked against
// the division for every mx = 2^b - 1 with b <= 16 and every code < 2^16
// (tests/codeword_check.cpp), the range of the uint16 code kernels.
STRIDE_DEV double quot_mx(double code, double mx, double r) {
  const double q0 = code * r;
  const double e = std::fma(-q0, mx, code);
  return std::fma(e, r, q0);
}

#if defined(__HIPCC__)
// Copies words [lo, lo + n) of the little-endian buffer into dst (LDS),
// byte-swapped to MSB-first, by the block's nthr threads with kStageDepth
// loads in flight per thread before the stores (a load-then-store loop waits
// a full memory round trip per row of words).  Threads past the end re-read
// word lo, in bounds, and store nothing.
constexpr uint32_t kStageDepth = 8;
__device__ __forceinline__ void stage_words(uint32_t* __restrict__ dst,
                                            const uint32_t* __restrict__ words, uint64_t lo,
                                            uint32_t n, uint32_t tid, uint32_t nthr) {
  for (uint32_t i0 = 0; i0 < n; i0 += kStageDepth * nthr) {
    uint32_t t[kStageDepth];
#pragma unroll
    for (uint32_t u = 0; u < kStageDepth; ++u) {
      const uint32_t i = i0 + u * nthr + tid;
      t[u] = words[lo + (i < n ? i : 0u)];
    }
#pragma unroll
    for (uint32_t u = 0; u < kStageDepth; ++u) {
      const uint32_t i = i0 + u * nthr + tid;
      if (i < n) dst[i] = __builtin_bswap32(t[u]);
    }
  }
}
#endif

#if defined(__HIPCC__)
// dst[i] = src[i] for i = first, first + step, ... < n, kStageDepth loads in
// flight before the stores (as stage_words; a load-then-store loop waits a
// memory round trip per element).  Slots past n re-read element i0, in
// bounds, and store nothing.
__device__ __forceinline__ void copy_strided(uint32_t* __restrict__ dst,
                                             const uint32_t* __restrict__ src, uint32_t n,
                                             uint32_t first, uint32_t step) {
  for (uint32_t i0 = first; i0 < n; i0 += kStageDepth * step) {
    uint32_t t[kStageDepth];
#pragma unroll
    for (uint32_t u = 0; u < kStageDepth; ++u) {
      const uint32_t i = i0 + u * step;
      t[u] = src[i < n ? i : i0];
    }
#pragma unroll
    for (uint32_t u = 0; u < kStageDepth; ++u) {
      const uint32_t i = i0 + u * step;
      if (i < n) dst[i] = t[u];
    }
  }
}
#endif

// Walk slot of global meshlet m in the two-pass walk buffer: its first
// vertex rounded up to 4 words, + 4 words per preceding meshlet, so every
// slot is 16-byte aligned and a 4-entry read past a slot's end stays inside
// the next slot or the 8 words after the last.
STRIDE_DEV uint64_t walk_slot(const MeshParams& P, const MeshletEntry& e, uint32_t m) {
  return ((P.v_base + e.v_off + 3) & ~3ull) + 4ull * m;
}
struct alignas(16) Walk4 {
  uint32_t x, y, z, w;
};

// Code storage words of one meshlet in the two-pass axis pass (the three
// code arrays only; the walk stays in HBM).
STRIDE_DEV uint32_t split_code_words(uint32_t n_local, bool narrow) {
  return narrow ? 3 * ((n_local + 1) / 2) : 3 * n_local;
}

// One axis lane's chain as a stepper over the walk: init() reads the
// anchor and the (tag, k) pairs of axis d and opens its delta and
// parallelogram substreams; first(w0) stores the anchor at the walk's first
// vertex, step(w) decodes the code of each later walk entry.  `words` is a
// word source indexed by absolute buffer word (the device buffer, or the
// meshlet's axis words staged in LDS: axis_stage_words), `base` the bit
// offset of the mesh's blob.  I32: IRLP in 32-bit arithmetic (the caller
// checks irlp32_exact; uint16 codes only).  ok turns false when a code
// leaves the range of CodeT.
// RICE: both substreams of the lane are Rice coded (the kernels choose it
// per wave when every lane's are), so codewords take codeword_rice.
// CHECK: track whether every code fits CodeT (good()).  The layout's
// validation decode needs it; later decodes of the same stream produce the
// same codes and skip the two instructions a step it costs.
template <bool I32, class Src, class CodeT, bool RICE = false, bool CHECK = true>
struct AxisChain {
  BitStream<Src> sd, sp;
  bool deg, peg, ok;
  // each stream's coder in one word: k | (Exp-Golomb ? 0xFFFFFFE0 : 0), so a
  // step selects one word and takes the mask as kp >> 5 (arithmetic)
  uint32_t dkp, pkp;
  int dk, pk;
  int32_t n0, n1, n2, K, rnd;
  int64_t anchor, prev;
  int32_t prev32;   // the 32-bit chain's previous code
  uint32_t bad;     // the 32-bit chain's range flags (bits 16..31)

  STRIDE_DEV void init(const Src& words, uint64_t base, const MeshletEntry& e, const MeshParams& P,
                       int d) {
    const uint32_t gb0 = (uint32_t)P.g_bits[0], gb1 = (uint32_t)P.g_bits[1],
                   gb2 = (uint32_t)P.g_bits[2];
    const uint32_t gb = pick3(d, gb0, gb1, gb2);
    const uint64_t r0 = base + e.resid_bit;
    anchor = (int64_t)bits_at_src(words, r0 + (d > 0 ? gb0 : 0u) + (d > 1 ? gb1 : 0u), gb);
    // (tag, k) of delta-d at tag byte pair d, of para-d at pair 3 + d
    const uint64_t tb = r0 + gb0 + gb1 + gb2;
    const uint32_t TD = bits_at_src(words, tb + 16u * (uint32_t)d, 16);
    const uint32_t TP = bits_at_src(words, tb + 16u * (uint32_t)(3 + d), 16);
    deg = (TD >> 8) != 1u;
    peg = (TP >> 8) != 1u;
    dkp = (TD & 31u) | (deg ? 0xFFFFFFE0u : 0u);
    pkp = (TP & 31u) | (peg ? 0xFFFFFFE0u : 0u);
    dk = (int)(TD & 31u);
    pk = (int)(TP & 31u);
    sd.init(words, base + pick3(d, e.sub_bit[0], e.sub_bit[1], e.sub_bit[2]));
    sp.init(words, base + pick3(d, e.sub_bit[3], e.sub_bit[4], e.sub_bit[5]));
    n0 = (int32_t)pick3(d, (uint32_t)P.pred_n[0], (uint32_t)P.pred_n[3], (uint32_t)P.pred_n[6]);
    n1 = (int32_t)pick3(d, (uint32_t)P.pred_n[1], (uint32_t)P.pred_n[4], (uint32_t)P.pred_n[7]);
    n2 = (int32_t)pick3(d, (uint32_t)P.pred_n[2], (uint32_t)P.pred_n[5], (uint32_t)P.pred_n[8]);
    K = (int32_t)pick3(d, (uint32_t)P.pred_K[0], (uint32_t)P.pred_K[1], (uint32_t)P.pred_K[2]);
    rnd = (I32 && K > 0) ? (1 << (K - 1)) : 0;
    ok = true;
    prev = anchor;
    prev32 = (int32_t)anchor;
    bad = 0;
  }
  STRIDE_DEV void first(uint32_t w0, CodeT* __restrict__ codes) {
    codes[w0 & 0xFFu] = (CodeT)anchor;
    if constexpr (CHECK) ok &= code_fits<CodeT>(anchor);
  }
  STRIDE_DEV void step(uint32_t w, CodeT* __restrict__ codes) {
    const bool root = walk_root(w);
    const uint32_t ia = (w >> 8) & 0xFFu, ib = (w >> 16) & 0xFFu, ic = w >> 24;
    // the reference codes are read before the codeword, so that the LDS
    // reads overlap its decode (the previous step's store is already issued)
    const CodeT ca = codes[ia], cb = codes[ib], cc = codes[ic];
    // the code comes from the delta stream (strip roots) or the
    // parallelogram stream; only the chosen one advances
    uint32_t u;
    const uint32_t kp = root ? dkp : pkp;
    const uint32_t len = RICE ? codeword_rice(root ? sd.win : sp.win, (int)(kp & 31u), u)
                              : codeword_win(root ? sd.win : sp.win,
                                             (uint32_t)((int32_t)kp >> 5), (int)(kp & 31u), u);
    int64_t r;
    if (len) {
      sd.drop(root ? len : 0u);
      sp.drop(root ? 0u : len);
      r = unzigzag32(u);
    } else {
      r = root ? codeword_s(sd, deg, dk) : codeword_s(sp, peg, pk);
      if constexpr (CHECK && I32 && sizeof(CodeT) == 2) {
        if (r != (int64_t)(int32_t)r) bad |= 0x80000000u;   // no 16-bit code then
      }
    }
    if constexpr (I32 && sizeof(CodeT) == 2) {
      // 32-bit chain.  The stored codes are uint16 and irlp32_exact holds,
      // so pred and prev lie in int32; c = base + r wraps only on a signed
      // overflow, which sets the sign bit of (base ^ c) & (r ^ c), and a
      // code outside [0, 65535] has c's top 16 bits set.  bad collects both
      // (and a general-path residual outside int32); the chain is exact
      // while bad's top 16 bits stay clear.
      const int32_t pred = irlp24((int32_t)ca, (int32_t)cb, (int32_t)cc, n0, n1, n2, K, rnd);
      const uint32_t r32 = (uint32_t)r;
      const uint32_t base = (uint32_t)(root ? prev32 : pred);
      const uint32_t c = base + r32;
      if constexpr (CHECK) bad |= c | ((base ^ c) & (r32 ^ c) & 0x80000000u);
      codes[w & 0xFFu] = (CodeT)c;
      prev32 = (int32_t)c;
    } else {
      const int64_t pred =
          irlp64((int64_t)ca, (int64_t)cb, (int64_t)cc, n0, n1, n2, K);
      const int64_t c = (root ? prev : pred) + r;
      if constexpr (CHECK) ok &= code_fits<CodeT>(c);
      codes[w & 0xFFu] = (CodeT)c;
      prev = c;
    }
  }
  // every code of the chain lies in CodeT
  STRIDE_DEV bool good() const { return !CHECK || (ok && (bad >> 16) == 0u); }
};

// The two-pass axis lane: the chain along the meshlet's walk slot in HBM
// (read four entries ahead).  Returns false when a code leaves CodeT.
template <bool I32, bool RICE = false, bool CHECK = true, class Src, class CodeT>
STRIDE_DEV bool split_axis_g(const Src& words, uint64_t base,
                             const MeshletEntry& e, const MeshParams& P, int d,
                             const Walk4* __restrict__ walk, CodeT* __restrict__ codes) {
  const uint32_t n_local = e.n_local;
  AxisChain<I32, Src, CodeT, RICE, CHECK> ax;
  ax.init(words, base, e, P, d);
  // the walk slot is 16-byte aligned, with readable words past its end
  // (quad (n_local >> 2) + 1 at most is read).  Entries 1..3, then whole
  // quads with no per-entry guard (each guard costs a compare and three
  // exec-mask instructions), then the last partial quad.
  Walk4 wq = walk[0];
  ax.first(wq.x, codes);
  Walk4 nq = walk[1];
  if (1 < n_local) ax.step(wq.y, codes);
  if (2 < n_local) ax.step(wq.z, codes);
  if (3 < n_local) ax.step(wq.w, codes);
  wq = nq;
  uint32_t i0 = 4;
  for (; i0 + 4 <= n_local; i0 += 4) {
    nq = walk[(i0 >> 2) + 1];
    ax.step(wq.x, codes);
    ax.step(wq.y, codes);
    ax.step(wq.z, codes);
    ax.step(wq.w, codes);
    wq = nq;
  }
  if (i0 < n_local) ax.step(wq.x, codes);
  if (i0 + 1 < n_local) ax.step(wq.y, codes);
  if (i0 + 2 < n_local) ax.step(wq.z, codes);
  return ax.good();
}

// Index pre-pass for one meshlet (one lane): parses the whole meshlet once,
// derives the side-table fields the stream does not store, and validates
// every field against the meshlet's bit range [start, end).
template <class Src>
STRIDE_DEV void scan_meshlet(const Src& src, uint32_t start, uint32_t end, const int32_t* g_bits,
                             MeshletScan& s) {
  s = MeshletScan{};
  const uint32_t limit = end - start;
  BitReader<Src> r;
  r.init(src, start);
  auto over = [&]() { return r.pos > end; };
  uint32_t nl = r.bits(16), nt = r.bits(16), ns = r.bits(16);
  s.n_local = nl;
  s.n_tris = nt;
  s.n_strips = ns;
  if (nl < 1 || nl > (uint32_t)kMaxLocal || nt < 1 || nt > (uint32_t)kMaxTris || ns < 1 ||
      ns > nt) {
    s.error = kScanErrSize;
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
  uint32_t flags = 0;
  auto token = [&](uint32_t& v) -> bool {
    uint32_t flag = r.bits(1);
    if (flag == 0) {
      uint32_t fi = r.bits(5);
      if (fi >= fifo.n) { err = kScanErrFifo; return false; }
      v = fifo.hit(fi);
    } else {
      v = r.bits(ib);
      if (v >= nl) { err = kScanErrIndex; return false; }
      if (fifo.find(v) >= 0) flags |= kFlagRawInFifo;
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
      if (ec == 0) p[0] = p[1];
      p[1] = p[2];
      p[2] = v;
      if (!seen.test_set(v)) walk++;
      if (over()) err = kScanErrOverrun;
    }
    if (over()) err = kScanErrOverrun;
  }
  if (err == kScanOk && tri_total != nt) err = kScanErrStrips;
  if (err == kScanOk && walk != nl) err = kScanErrWalk;
  if (err != kScanOk) {
    s.error = err;
    return;
  }
  s.n_kind0 = kind0;
  s.resid_bit = r.pos;
  r.bits(g_bits[0]);
  r.bits(g_bits[1]);
  r.bits(g_bits[2]);
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
  s.flags = flags;
}


}  // namespace stride
