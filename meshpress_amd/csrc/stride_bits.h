// Device-side bit-level primitives for the DPR4 stream: an MSB-first bit
// reader over big-endian 32-bit words, the 16-entry move-to-back reuse FIFO
// packed into two 64-bit registers, and a 256-bit seen set.
//
// Field semantics follow utils/bit_codec.py:192-257 (MSB-first; Rice = q
// zeros, '1', k low bits; Exp-Golomb = lb zeros, '1', lb-bit offset, k low
// bits) and encoder/paradelta_codec.py:330-340 (FIFO).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define STRIDE_DEV __host__ __device__ __forceinline__
#else
#define STRIDE_DEV inline
#endif

namespace stride {

// Word sources: return big-endian-order word i (already byte swapped), or 0
// past the end.  Reads never leave [0, n).
struct LdsWords {
  const uint32_t* w;
  uint32_t n;
  STRIDE_DEV uint32_t operator()(uint32_t i) const { return i < n ? w[i] : 0u; }
};

struct GlobalWords {
  const uint32_t* w;   // little-endian words of the blob
  uint32_t n;
  STRIDE_DEV uint32_t operator()(uint32_t i) const {
    return i < n ? __builtin_bswap32(w[i]) : 0u;
  }
};

template <class Src>
struct BitReader {
  uint64_t win;    // top-aligned pending bits
  int valid;       // number of pending bits in win
  uint32_t widx;   // next word to load
  uint32_t pos;    // absolute bit position of the next unread bit
  Src src;

  STRIDE_DEV void init(const Src& s, uint32_t bit) {
    src = s;
    uint32_t w0 = bit >> 5;
    int skip = (int)(bit & 31u);
    win = ((uint64_t)src(w0) << 32) | (uint64_t)src(w0 + 1);
    win <<= skip;
    valid = 64 - skip;
    widx = w0 + 2;
    pos = bit;
  }
  STRIDE_DEV void refill() {
    if (valid <= 32) {
      win |= (uint64_t)src(widx++) << (32 - valid);
      valid += 32;
    }
  }
  // 0 <= n <= 32
  STRIDE_DEV uint32_t bits(int n) {
    refill();
    uint32_t out = n ? (uint32_t)(win >> (64 - n)) : 0u;
    win = n ? (win << n) : win;
    valid -= n;
    pos += (uint32_t)n;
    return out;
  }
  // Count zeros before the next '1' and consume the '1'.  Stops counting
  // once the run passes `limit` bits (malformed input); the caller checks
  // pos against the meshlet end.
  STRIDE_DEV uint32_t unary(uint32_t limit) {
    uint32_t q = 0;
    for (;;) {
      refill();
      int lz = win ? __builtin_clzll(win) : 64;
      if (lz < valid) {
        win = (win << lz) << 1;
        valid -= lz + 1;
        pos += (uint32_t)lz + 1u;
        return q + (uint32_t)lz;
      }
      q += (uint32_t)valid;
      pos += (uint32_t)valid;
      win = 0;
      valid = 0;
      if (q > limit) return q;
    }
  }
  // Rice(k): u = (q << k) | low   (utils/bit_codec.py:226-235)
  STRIDE_DEV uint64_t rice(int k, uint32_t limit) {
    uint64_t q = unary(limit);
    uint32_t low = bits(k);
    return (q << k) | low;
  }
  // Exp-Golomb(k) (utils/bit_codec.py:237-250)
  STRIDE_DEV uint64_t exp_golomb(int k, uint32_t limit) {
    uint32_t lb = unary(limit);
    uint64_t shifted = 0;
    if (lb > 0) {
      uint32_t off = lb <= 32 ? bits((int)lb) : 0u;
      shifted = ((uint64_t)1 << (lb & 63)) - 1 + off;
    }
    uint32_t low = bits(k);
    return (shifted << k) | low;
  }
};

STRIDE_DEV int64_t unzigzag(uint64_t u) {
  return (int64_t)(u >> 1) ^ -(int64_t)(u & 1);
}

// Reuse FIFO: byte i of the 128-bit value (lo, hi) is entry i, entry 0 the
// oldest.  Bytes at positions >= n are kept zero.
struct Fifo {
  uint64_t lo, hi;
  uint32_t n;
  STRIDE_DEV void clear() { lo = 0; hi = 0; n = 0; }
  STRIDE_DEV uint32_t get(uint32_t i) const {
    uint64_t w = i < 8 ? lo : hi;
    return (uint32_t)(w >> (8 * (i & 7))) & 0xFFu;
  }
  // Remove entry i (< n); entries above it move down by one.
  STRIDE_DEV void remove(uint32_t i) {
    uint64_t slo = (lo >> 8) | (hi << 56);
    uint64_t shi = hi >> 8;
    if (i < 8) {
      uint64_t m = i ? (~0ull >> (64 - 8 * i)) : 0ull;
      lo = (lo & m) | (slo & ~m);
      hi = shi;
    } else {
      uint32_t j = i - 8;
      uint64_t m = j ? (~0ull >> (64 - 8 * j)) : 0ull;
      hi = (hi & m) | (shi & ~m);
    }
    n--;
  }
  STRIDE_DEV void append(uint32_t v) {
    if (n < 8) lo |= (uint64_t)(v & 0xFFu) << (8 * n);
    else hi |= (uint64_t)(v & 0xFFu) << (8 * (n - 8));
    n++;
  }
  // Index of entry equal to v among the first n entries, or -1.
  STRIDE_DEV int find(uint32_t v) const {
    const uint64_t ones = 0x0101010101010101ull, low7 = 0x7F7F7F7F7F7F7F7Full;
    const uint64_t pat = (uint64_t)(v & 0xFFu) * ones;
    uint64_t t0 = lo ^ pat, t1 = hi ^ pat;
    uint64_t z0 = ~(((t0 & low7) + low7) | t0 | low7);
    uint64_t z1 = ~(((t1 & low7) + low7) | t1 | low7);
    uint64_t v0 = n >= 8 ? ~0ull : (n ? (~0ull >> (64 - 8 * n)) : 0ull);
    uint64_t v1 = n <= 8 ? 0ull : (n >= 16 ? ~0ull : (~0ull >> (64 - 8 * (n - 8))));
    z0 &= v0;
    z1 &= v1;
    if (z0) return __builtin_ctzll(z0) >> 3;
    if (z1) return 8 + (__builtin_ctzll(z1) >> 3);
    return -1;
  }
  // Token resolution (encoder/paradelta_codec.py:330-340): a FIFO hit
  // takes entry fi; a raw id leaves the FIFO if present; the value is
  // moved to the back, the oldest entry dropping out of a full FIFO.
  STRIDE_DEV uint32_t hit(uint32_t fi) {
    uint32_t v = get(fi);
    remove(fi);
    append(v);
    return v;
  }
  STRIDE_DEV void raw(uint32_t v) {
    int p = find(v);
    if (p >= 0) remove((uint32_t)p);
    else if (n == 16) remove(0);
    append(v);
  }
};

// The same reuse FIFO kept newest-first, for branch-free updates: byte j of
// (lo, hi) is the entry of recency rank j (0 = most recently used); FIFO
// index fi (oldest first) is rank n - 1 - fi.  Every token is one
// move-to-front: ranks below j move up by one and the value lands at rank 0;
// rank j is dropped (j = n for a value not in the FIFO, which with n == 16
// drops the oldest entry).
struct Mtf {
  uint64_t lo, hi;
  uint32_t n;
  STRIDE_DEV void clear() { lo = 0; hi = 0; n = 0; }
  // entry j < 16.  A mask blend, not a select of the two halves: a select of
  // two loads folds into one load through a computed address, which keeps
  // a stepper's state out of registers.
  STRIDE_DEV uint32_t rank(uint32_t j) const {
    const uint32_t sh = 8 * (j & 7);
    const uint64_t m = 0ull - (uint64_t)(j >> 3);   // ones for j in [8, 16)
    return (uint32_t)(((lo >> sh) & ~m) | ((hi >> sh) & m)) & 0xFFu;
  }
  // 128-bit mask of bytes [0, j), j in [0, 17]
  STRIDE_DEV static void mask(uint32_t j, uint64_t& mlo, uint64_t& mhi) {
    const uint32_t s = 8 * j;
    mlo = s >= 64 ? ~0ull : ((1ull << s) - 1);
    mhi = s <= 64 ? 0ull : (s >= 128 ? ~0ull : ((1ull << (s - 64)) - 1));
  }
  STRIDE_DEV void to_front(uint32_t j, uint32_t v) {
    uint64_t mlo, mhi, klo, khi;
    mask(j, mlo, mhi);
    mask(j + 1, klo, khi);
    const uint64_t llo = lo & mlo, lhi = hi & mhi;
    const uint64_t nlo = (lo & ~klo) | (llo << 8) | (uint64_t)(v & 0xFFu);
    const uint64_t nhi = (hi & ~khi) | (lhi << 8) | (llo >> 56);
    lo = nlo;
    hi = nhi;
  }
  // rank of v among the n entries, or -1
  STRIDE_DEV int find(uint32_t v) const {
    const uint64_t ones = 0x0101010101010101ull, low7 = 0x7F7F7F7F7F7F7F7Full;
    const uint64_t pat = (uint64_t)(v & 0xFFu) * ones;
    const uint64_t t0 = lo ^ pat, t1 = hi ^ pat;
    uint64_t z0 = ~(((t0 & low7) + low7) | t0 | low7);
    uint64_t z1 = ~(((t1 & low7) + low7) | t1 | low7);
    uint64_t v0, v1;
    mask(n, v0, v1);
    z0 &= v0;
    z1 &= v1;
    if (z0) return __builtin_ctzll(z0) >> 3;
    if (z1) return 8 + (__builtin_ctzll(z1) >> 3);
    return -1;
  }
};

// The newest-first reuse FIFO as two 64-bit halves with cheap masks: the
// move-to-front of rank j keeps the bytes above j and shifts bytes 0..j up
// by one, new = (old & ~M) | ((old << 8) & M) | v with M = bytes [0, j].
// v (< 256) in every byte: one byte permute on the device (a multiply by
// 0x01010101 compiles to a quarter-rate v_mul_lo_u32).
STRIDE_DEV uint32_t byte4(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm(0u, v, 0u);
#else
  return (v & 0xFFu) * 0x01010101u;
#endif
}

struct MtfLite {
  uint64_t lo, hi;
  uint32_t n;
  STRIDE_DEV void clear() { lo = 0; hi = 0; n = 0; }
  // entry j < 16.  A mask blend, not a select of the two halves: a select of
  // two loads folds into one load through a computed address, which keeps
  // a stepper's state out of registers.
  STRIDE_DEV uint32_t rank(uint32_t j) const {
    const uint32_t sh = 8 * (j & 7);
    const uint64_t m = 0ull - (uint64_t)(j >> 3);   // ones for j in [8, 16)
    return (uint32_t)(((lo >> sh) & ~m) | ((hi >> sh) & m)) & 0xFFu;
  }
  STRIDE_DEV void to_front(uint32_t j, uint32_t v) {
    // bytes [0, j] of the low half: all ones >> (56 - 8 min(j, 7))
    const uint32_t jl = j < 7 ? j : 7u;
    const uint64_t ml = ~0ull >> (56u - 8u * jl);
    // bytes [8, j] of the high half (none when j < 8)
    const uint32_t jh = j < 8 ? 0u : (j - 8 < 7 ? j - 8 : 7u);
    const uint64_t mh = j < 8 ? 0ull : (~0ull >> (56u - 8u * jh));
    const uint64_t sl = (lo << 8) | (uint64_t)(v & 0xFFu);
    const uint64_t sh = (hi << 8) | (lo >> 56);
    lo = (lo & ~ml) | (sl & ml);
    hi = (hi & ~mh) | (sh & mh);
  }
  // 0x80 in each zero byte of x
  STRIDE_DEV static uint32_t zero_bytes(uint32_t x) {
    const uint32_t low7 = 0x7F7F7F7Fu;
    return ~(((x & low7) + low7) | x | low7);
  }
  // rank of v among the n entries, or -1.  Branch-free: the first byte equal
  // to v over all 16 bytes.  Bytes [n, 16) sit above every entry, so the
  // first match is an entry whenever any entry matches.
  STRIDE_DEV int find(uint32_t v) const {
    const uint32_t pat = byte4(v);
    const uint64_t z0 = (uint64_t)zero_bytes((uint32_t)lo ^ pat) |
                        ((uint64_t)zero_bytes((uint32_t)(lo >> 32) ^ pat) << 32);
    const uint64_t z1 = (uint64_t)zero_bytes((uint32_t)hi ^ pat) |
                        ((uint64_t)zero_bytes((uint32_t)(hi >> 32) ^ pat) << 32);
    uint32_t i = z1 ? 8u + ((uint32_t)__builtin_ctzll(z1) >> 3) : 16u;
    i = z0 ? ((uint32_t)__builtin_ctzll(z0) >> 3) : i;
    return i < n ? (int)i : -1;
  }
};

struct Seen256 {
  uint64_t s[4];
  STRIDE_DEV void clear() { s[0] = s[1] = s[2] = s[3] = 0; }
  STRIDE_DEV bool test_set(uint32_t v) {
    uint32_t q = (v >> 6) & 3u;
    uint64_t b = 1ull << (v & 63);
    uint64_t w = q == 0 ? s[0] : q == 1 ? s[1] : q == 2 ? s[2] : s[3];
    bool was = (w & b) != 0;
    w |= b;
    if (q == 0) s[0] = w;
    else if (q == 1) s[1] = w;
    else if (q == 2) s[2] = w;
    else s[3] = w;
    return was;
  }
};

}  // namespace stride
