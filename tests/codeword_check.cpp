// Host build of the axis lanes' codeword fast path: codeword_win (32-bit,
// codewords of at most 32 bits) and its Rice-only form codeword_rice, and
// irlp24 against the general stream
// decoder codeword_s and irlp64 (meshpress_amd/csrc/stride_lane.h), over
// every unary length q, coder parameter k and coder, with random bodies and
// trailing bits, and over the edge windows (all zeros, all ones); the longer
// codewords (codeword_s's general path) against the bit-serial BitReader
// decode of the same words (meshpress_amd/csrc/stride_bits.h); and the
// dequantisation quotient quot_mx against the float64 division for every
// g_bits <= 16 and every 16-bit code; and the token lanes' reuse FIFO
// (MtfLite: rank, to_front, find) against a plain list over random hit /
// raw sequences, raw ids present or not, ids 0 and 255 included; and the
// walk-entry root test: walk_root holds for every byte4(v) and for no
// parallelogram entry (a != v), refs (255, 255, 255) and (0, 0, 0) included.
//
//   codeword_check      exit 0 = all equal, 1 = a mismatch (printed)
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <vector>

#include "../meshpress_amd/csrc/stride_lane.h"

using namespace stride;

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint64_t next64() {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return rng;
}

int main() {
  long checked = 0, fast = 0, longer = 0;
  for (int eg = 0; eg < 2; ++eg)
    for (int k = 0; k < 32; ++k)
      for (uint32_t q = 0; q <= 40; ++q)
        for (int rep = 0; rep < 64; ++rep) {
          // window: q zeros, a one, then random bits (or the edge patterns)
          uint64_t W = next64();
          if (rep == 0) W = 0;
          if (rep == 1) W = ~0ull;
          if (q < 64 && rep > 1) W = (W >> (q + 1)) | (1ull << (63 - q));
          if (q >= 64) W = 0;
          // the same bits as a big-endian word stream, then more random words
          std::vector<uint32_t> le(8);
          const uint32_t hi = (uint32_t)(W >> 32), lo = (uint32_t)W;
          le[0] = __builtin_bswap32(hi);
          le[1] = __builtin_bswap32(lo);
          for (int i = 2; i < 8; ++i) le[i] = (uint32_t)next64();
          StreamBits s;
          s.init(le.data(), 0);
          uint32_t u = 0;
          const uint32_t len = codeword_win(s.win, eg ? ~0u : 0u, k, u);
          if (!eg) {   // the Rice-only form gives the same length and value
            uint32_t ur = 0;
            const uint32_t lr = codeword_rice(s.win, k, ur);
            if (lr != len || (len && ur != u)) {
              std::printf("codeword_rice mismatch: k %d W %016llx (%u, %u) vs (%u, %u)\n", k,
                          (unsigned long long)W, lr, ur, len, u);
              return 1;
            }
          }
          const uint32_t nb = (eg ? (W ? (uint32_t)__builtin_clzll(W) : 64u) : 0u) + (uint32_t)k;
          const uint32_t qq = W ? (uint32_t)__builtin_clzll(W) : 64u;
          if (qq + 1 + nb > 32) {
            if (len != 0) {
              std::printf("long codeword not rejected: eg %d k %d W %016llx len %u\n", eg, k,
                          (unsigned long long)W, len);
              return 1;
            }
            // the general path against the bit-serial reader (W != 0: the
            // unary run ends inside the first 64 bits; an Exp-Golomb prefix of
            // more than 32 zeros is a value of 2^33 or more, outside every
            // stream, which the reader does not decode)
            if (W != 0 && !(eg && qq > 32)) {
              uint32_t be[8];
              for (int i = 0; i < 8; ++i) be[i] = __builtin_bswap32(le[i]);
              BitReader<LdsWords> br;
              br.init(LdsWords{be, 8}, 0);
              const uint64_t want_u = eg ? br.exp_golomb(k, 0xFFFFFFu) : br.rice(k, 0xFFFFFFu);
              const int64_t got = codeword_s(s, eg != 0, k);
              const uint64_t used = 64 - s.valid + 32 * (s.next - 3);
              if (got != unzigzag(want_u) || used != br.pos) {
                std::printf("long codeword mismatch: eg %d k %d W %016llx: (%llu, %lld) vs "
                            "(%u, %lld)\n", eg, k, (unsigned long long)W,
                            (unsigned long long)used, (long long)got, br.pos,
                            (long long)unzigzag(want_u));
                return 1;
              }
              ++longer;
            }
            ++checked;
            continue;
          }
          if (W == 0) continue;   // no terminating one inside the window: long path only
          const int64_t r = codeword_s(s, eg != 0, k);
          const uint64_t used = 64 - s.valid + 32 * (s.next - 3);   // bits consumed
          if (len == 0 || unzigzag32(u) != r || used != len) {
            std::printf("mismatch: eg %d k %d W %016llx fast (%u, %lld) general (%llu, %lld)\n", eg,
                        k, (unsigned long long)W, len, (long long)unzigzag32(u),
                        (unsigned long long)used, (long long)r);
            return 1;
          }
          ++checked;
          ++fast;
        }
  // irlp24 against irlp64 wherever irlp32_exact holds
  long irl = 0;
  for (int rep = 0; rep < 2000000; ++rep) {
    const int32_t K = (int32_t)(next64() % 32);
    int32_t n[3];
    for (int j = 0; j < 3; ++j) {
      const uint64_t t = next64();
      n[j] = (t & 1) ? (int32_t)(int16_t)(t >> 8) : (int32_t)((int64_t)((t >> 8) % 2049) - 1024);
    }
    if (!irlp32_exact(n[0], n[1], n[2], K)) continue;
    const int32_t a = (int32_t)(next64() & 0xFFFF), b = (int32_t)(next64() & 0xFFFF),
                  c = (int32_t)(next64() & 0xFFFF);
    const int32_t rnd = K > 0 ? (1 << (K - 1)) : 0;
    const int64_t want = irlp64(a, b, c, n[0], n[1], n[2], K);
    const int64_t got = irlp24(a, b, c, n[0], n[1], n[2], K, rnd);
    if (want != got) {
      std::printf("irlp mismatch: n %d %d %d K %d codes %d %d %d: %lld vs %lld\n", n[0], n[1],
                  n[2], K, a, b, c, (long long)want, (long long)got);
      return 1;
    }
    ++irl;
  }
  // the bound itself: the largest int16 numerators at K = 0 overflow int32
  if (irlp32_exact(-32768, -32768, -32768, 0) || !irlp32_exact(512, 512, -512, 9)) {
    std::printf("irlp32_exact bound wrong\n");
    return 1;
  }
  // quot_mx against the division: every g_bits <= 16, every code < 2^16
  long quot = 0;
  for (int b = 1; b <= 16; ++b) {
    const double mx = (double)((1u << b) - 1u), r = 1.0 / mx;
    for (uint32_t c = 0; c < 65536u; ++c) {
      const double want = (double)c / mx, got = quot_mx((double)c, mx, r);
      if (want != got) {
        std::printf("quot_mx mismatch: b %d code %u: %.17g vs %.17g\n", b, c, got, want);
        return 1;
      }
      ++quot;
    }
  }
  // reuse FIFO: front = index 0 (the most recent entry), at most 16 entries
  long fifo_ops = 0;
  for (int trial = 0; trial < 4000; ++trial) {
    std::deque<uint32_t> ref;
    MtfLite f;
    f.clear();
    const uint32_t n_ids = trial % 4 == 0 ? 256u : 1u + (uint32_t)(next64() % 256);
    for (int step = 0; step < 400; ++step) {
      const uint64_t r = next64();
      uint32_t j, v;
      if (!ref.empty() && r % 3 != 0) {   // hit: entry j
        j = (uint32_t)((r >> 8) % ref.size());
        v = ref[j];
        if (f.rank(j) != v) {
          std::printf("fifo rank mismatch: trial %d step %d\n", trial, step);
          return 1;
        }
      } else {                            // raw id: in the FIFO or not
        v = (r >> 2) % 4 == 0 && !ref.empty() ? ref[(r >> 16) % ref.size()]
                                             : (uint32_t)((r >> 24) % n_ids);
        int want = -1;
        for (size_t i = 0; i < ref.size(); ++i)
          if (ref[i] == v) { want = (int)i; break; }
        const int got = f.find(v);
        if (got != want) {
          std::printf("fifo find mismatch: trial %d step %d v %u: %d vs %d\n", trial, step, v,
                      got, want);
          return 1;
        }
        j = want >= 0 ? (uint32_t)want : f.n;
        if (j == f.n && f.n < 16) f.n++;
      }
      f.to_front(j, v);
      if (j < ref.size()) ref.erase(ref.begin() + j);
      else if (ref.size() == 16) ref.pop_back();
      ref.push_front(v);
      if (f.n != ref.size()) {
        std::printf("fifo size mismatch: trial %d step %d\n", trial, step);
        return 1;
      }
      ++fifo_ops;
    }
  }
  // walk entries: roots v * 0x01010101, parallelograms v | a << 8 | b << 16 | c << 24
  long entries = 0;
  for (uint32_t v = 0; v < 256; ++v) {
    if (!walk_root(byte4(v)) || byte4(v) != v * 0x01010101u) {
      std::printf("root entry wrong: v %u\n", v);
      return 1;
    }
    for (uint32_t a = 0; a < 256; ++a) {
      if (a == v) continue;
      for (int rep = 0; rep < 16; ++rep) {
        const uint64_t r = next64();
        uint32_t b = (uint32_t)(r & 0xFFu), c = (uint32_t)((r >> 8) & 0xFFu);
        if (rep == 0) b = c = 255;
        if (rep == 1) b = c = 0;
        if (rep == 2) b = c = v;
        const uint32_t w = v | (a << 8) | (b << 16) | (c << 24);
        if (walk_root(w)) {
          std::printf("parallelogram entry taken for a root: %08x\n", w);
          return 1;
        }
        ++entries;
      }
    }
  }
  std::printf("codewords %ld (fast path %ld, longer %ld), irlp %ld, quotients %ld, fifo ops %ld, "
              "walk entries %ld\n",
              checked, fast, longer, irl, quot, fifo_ops, entries);
  return 0;
}
