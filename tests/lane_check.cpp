// Host build of the per-lane decode logic of the gfx950 kernels
// (meshpress_amd/csrc/stride_lane.h): the index pre-pass (scan_meshlet) and
// the lane decoder (decode_lane) run here on the CPU, under the sanitizers
// when built with them, and their output is compared with the oracle.
//
//   lane_check <stream.dpr4> <verts.f32 out> <tris.u32 out>
//              [split|split16|twopass|twopass16|chains<K>|chains<K>n16]
//
// With "split" the split decoder runs instead: every meshlet's token lane
// (split_tokens), then its three axis lanes (split_axis), as the kernel's
// phases do.  "twopass" / "twopass16" run the two-pass decoder: all token
// lanes first with register seen sets, then the axis lanes on blocks that
// hold only the copied walk.  "chains<K>" runs the chain token pass instead
// of the token lanes (chain-start scan with K chains per meshlet, one
// ChainLane per chain, then the per-meshlet record resolution the kernel's
// wave does, here one token after another), then the two-pass axis lanes
// ("chains<K>n16": with uint16 codes).
// Exit status: 0 ok, 2 stream rejected by the pre-pass, 3 code overflow
// (the kernels route such groups to the wave-per-meshlet decoder), 4 the
// 32-bit and 64-bit axis loops disagree.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../meshpress_amd/csrc/stride_lane.h"

using namespace stride;

static uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

int main(int argc, char** argv) {
  if (argc < 4) return 1;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 1;
  std::vector<uint8_t> data;
  uint8_t buf[65536];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + r);
  std::fclose(f);
  if (data.size() < 86) return 2;
  MeshParams P;
  std::memset(&P, 0, sizeof P);
  const uint8_t* d = data.data();
  auto f32 = [&](int o) {
    uint32_t u = be32(d + o);
    float x;
    std::memcpy(&x, &u, 4);
    return x;
  };
  for (int i = 0; i < 3; ++i) {
    P.center[i] = f32(6 + 4 * i);
    P.g_min[i] = f32(26 + 4 * i);
    P.g_range[i] = f32(38 + 4 * i);
    P.g_bits[i] = d[50 + i];
    P.pred_K[i] = d[71 + i];
  }
  P.scale = f32(18);
  for (int i = 0; i < 9; ++i) {
    uint32_t u = ((uint32_t)d[53 + 2 * i] << 8) | d[54 + 2 * i];
    P.pred_n[i] = (u & 0x8000u) ? (int32_t)u - 0x10000 : (int32_t)u;
  }
  const uint32_t n_m = be32(d + 82);
  const size_t n_words = (data.size() + 3) / 4 + 16;
  std::vector<uint32_t> le(n_words, 0), st(n_words, 0);
  std::memcpy(le.data(), d, data.size());
  for (size_t i = 0; i < n_words; ++i) st[i] = __builtin_bswap32(le[i]);
  std::vector<MeshletEntry> tab(n_m);
  uint64_t nv = 0, nt = 0;
  for (uint32_t m = 0; m < n_m; ++m) {
    const uint32_t start = be32(d + 86 + 4 * m);
    const uint32_t end = m + 1 < n_m ? be32(d + 86 + 4 * (m + 1)) : (uint32_t)(8 * data.size());
    MeshletScan s;
    GlobalWords src{le.data(), (end + 31u) / 32u + 2u};
    scan_meshlet(src, start, end, P.g_bits, s);
    if (s.error || (m + 1 < n_m && s.end_bit != end)) {   // as the host validation
      std::fprintf(stderr, "meshlet %u rejected (%u)\n", m, s.error);
      return 2;
    }
    MeshletEntry& e = tab[m];
    e.start_bit = start;
    e.v_off = (uint32_t)nv;
    e.t_off = (uint32_t)nt;
    e.n_local = s.n_local;
    e.n_tris = s.n_tris;
    e.n_strips = s.n_strips | (s.flags & kFlagRawInFifo);
    e.n_kind0 = s.n_kind0;
    e.resid_bit = s.resid_bit;
    for (int j = 0; j < 6; ++j) e.sub_bit[j] = s.sub_bit[j];
    e.end_bit = s.end_bit;
    nv += s.n_local;
    nt += s.n_tris;
  }
  const bool chains = argc > 4 && std::strncmp(argv[4], "chains", 6) == 0;
  const uint32_t k0 = chains ? (uint32_t)std::atoi(argv[4] + 6) : 0u;
  const bool two = chains || (argc > 4 && std::strncmp(argv[4], "twopass", 7) == 0);
  const bool split = two || (argc > 4 && std::strncmp(argv[4], "split", 5) == 0);
  const bool narrow = argc > 4 && (std::strcmp(argv[4], "split16") == 0 ||
                                   std::strcmp(argv[4], "twopass16") == 0 ||
                                   (chains && std::strstr(argv[4], "n16") != nullptr));
  std::vector<uint64_t> codes(nv, kUnseen);
  std::vector<uint32_t> tris(3 * nt, 0xFFFFFFFFu);
  std::vector<int32_t> codes3(3 * nv, 0);
  int rc = 0;
  // two-pass: every token lane first (register seen set, walks to one
  // exactly sized array), then the axis lanes over blocks holding stale data
  // except for the copied walk
  // the walk buffer of the device: 16-byte aligned slots (walk_slot)
  std::vector<Walk4> walk4(two ? (nv + 4ull * n_m + 8) / 4 + 1 : 0);
  uint32_t* walk_g = reinterpret_cast<uint32_t*>(walk4.data());
  if (chains)
    for (uint32_t m = 0; m < n_m; ++m) {
      const MeshletEntry& e = tab[m];
      const uint32_t nl = e.n_local, nt = meshlet_tokens(e);
      const uint32_t lm = chain_tokens(nt, nl, e.n_strips, k0);
      const uint32_t cnt = chain_count(nt, lm);
      std::vector<ChainEntry> ch(cnt);
      GlobalWords src{le.data(), (uint32_t)n_words};
      chain_starts(src, e.start_bit, e, k0, ch.data());
      // phase 1: every chain on its own, exactly sized record array
      std::vector<uint16_t> rec(nt);
      std::vector<uint8_t> fin(16 * cnt);
      struct Sink {
        std::vector<uint16_t>* r;
        void operator()(uint32_t t, uint32_t x) const { r->at(t) = (uint16_t)x; }
      } sink{&rec};
      for (uint32_t c = 0; c < cnt; ++c) {
        if (((ch[c].state >> 24) & 15u) != c) return 5;
        StreamBits cur;
        cur.init(le.data(), e.start_bit + (ch[c].pos & 0xFFFFu));
        ChainLane<StreamBits> cl;
        cl.init(cur, ch[c], chain_end(ch[c], nt, lm), nl, e.n_strips);
        while (!cl.done()) cl.step<true>(cur, sink);
        if (cl.t != (c + 1 < cnt ? (ch[c + 1].pos >> 16) : nt)) return 5;
        for (int i = 0; i < 16; ++i)
          fin[16 * c + i] = (uint8_t)((i < 8 ? cl.fifo.lo >> (8 * i) : cl.fifo.hi >> (8 * (i - 8))) & 0xFFu);
      }
      // phase 2: placeholders, first sightings, triangles and walk entries
      for (uint32_t c = 1; c + 1 < cnt; ++c)
        for (int i = 0; i < 16; ++i) {
          const uint32_t x = fin[16 * c + i];
          if (x >= nl) fin[16 * c + i] = fin[16 * (c - 1) + ((x - nl) & 15u)];
        }
      std::vector<uint32_t> first(nl, 0xFFFFFFFFu);
      for (uint32_t t = 0; t < nt; ++t) {
        uint32_t v = rec[t] & 0xFFu;
        if (v >= nl) {
          const uint32_t c = rec[t] >> 12;
          if (c == 0) return 5;
          v = fin[16 * (c - 1) + ((v - nl) & 15u)];
          rec[t] = (uint16_t)((rec[t] & 0xFF00u) | v);
        }
        if (v >= nl) return 5;
        first[v] = std::min(first[v], t);
      }
      uint32_t* wout = walk_g + walk_slot(P, e, m);
      uint32_t* tout = tris.data() + 3ull * e.t_off;
      int last = -1;
      uint32_t wi = 0, ti = 0;
      for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t r = rec[t], v = r & 0xFFu;
        const int s = last;
        const uint32_t v1 = rec[t >= 1 ? t - 1 : 0] & 0xFFu, v2 = rec[t >= 2 ? t - 2 : 0] & 0xFFu;
        const uint32_t p0 = rec[s >= 2 ? (uint32_t)s - 2 : 0] & 0xFFu;
        uint32_t a, b;
        const uint32_t entry = chain_entry(r, v1, v2, p0, a, b);
        if (first[v] == t) wout[wi++] = entry;
        if (chain_emits(r)) {
          tout[3 * ti] = e.v_off + a;
          tout[3 * ti + 1] = e.v_off + b;
          tout[3 * ti + 2] = e.v_off + v;
          ++ti;
        }
        if ((r & kRecEc) == 0) last = (int)t;
      }
      if (wi != nl || ti != e.n_tris) return 5;
    }
  else if (two)
    for (uint32_t m = 0; m < n_m; ++m) {
      const MeshletEntry& e = tab[m];
      uint32_t sw[8];
      BitSeen seen{sw};
      seen.preset(0);
      StreamBits cur;
      cur.init(le.data(), e.start_bit + 48);
      split_tokens_c<WalkSlot>(cur, e, walk_g + walk_slot(P, e, m), seen,
                               tris.data() + 3ull * e.t_off);
    }
  for (uint32_t m = 0; m < n_m; ++m) {
    const MeshletEntry& e = tab[m];
    // one group = the whole stream staged: rel is the absolute start bit
    if (two) {
      // the two-pass axis lanes: bits and walk from "HBM", codes-only block
      std::vector<uint32_t> blk(split_code_words(e.n_local, narrow), 0xA5A5A5A5u);
      const uint32_t cs = split_code_stride(e.n_local, narrow);
      const Walk4* wv = reinterpret_cast<const Walk4*>(walk_g + walk_slot(P, e, m));
      // the kernel's LDS copy of the meshlet's axis words, exactly sized
      const uint64_t alo = axis_stage_lo(0, e);
      std::vector<uint32_t> ax(st.begin() + alo, st.begin() + alo + axis_stage_words(0, e));
      const LdsStaged sw{ax.data(), alo};
      const BufferWords gw{le.data()};
      for (int d = 0; d < 3; ++d) {
        bool okd;
        if (narrow) {
          uint16_t* c = reinterpret_cast<uint16_t*>(blk.data()) + cs * d;
          okd = split_axis_g<false>(sw, 0, e, P, d, wv, c);
          // the 32-bit IRLP loop and the device-buffer reader give the same codes
          std::vector<uint16_t> c2(e.n_local, 0x5A5Au);
          const bool i32 =
              irlp32_exact(P.pred_n[3 * d], P.pred_n[3 * d + 1], P.pred_n[3 * d + 2], P.pred_K[d]);
          const bool ok2 = i32 ? split_axis_g<true>(gw, 0, e, P, d, wv, c2.data())
                               : split_axis_g<false>(gw, 0, e, P, d, wv, c2.data());
          if (ok2 != okd || (okd && !std::equal(c2.begin(), c2.end(), c))) rc = 4;
          // the Rice-only codeword path, where both substreams are Rice coded
          if (axis_rice(sw, 0, e, P, d)) {
            std::vector<uint16_t> c3(e.n_local, 0x5A5Au);
            const bool ok3 = i32 ? split_axis_g<true, true>(sw, 0, e, P, d, wv, c3.data())
                                 : split_axis_g<false, true>(sw, 0, e, P, d, wv, c3.data());
            if (ok3 != okd || (okd && !std::equal(c3.begin(), c3.end(), c))) rc = 4;
          }
          for (uint32_t v = 0; v < e.n_local; ++v) codes3[3 * ((size_t)e.v_off + v) + d] = c[v];
        } else {
          int32_t* c = reinterpret_cast<int32_t*>(blk.data()) + cs * d;
          okd = split_axis_g<false>(sw, 0, e, P, d, wv, c);
          for (uint32_t v = 0; v < e.n_local; ++v) codes3[3 * ((size_t)e.v_off + v) + d] = c[v];
        }
        if (!okd) rc = 3;
      }
    } else if (split && narrow) {
      std::vector<uint32_t> blk(split_block_words(e.n_local, true), two ? 0xA5A5A5A5u : 0u);
      uint16_t* c16 = reinterpret_cast<uint16_t*>(blk.data() + e.n_local);
      const uint32_t cs = split_code_stride(e.n_local, true);
      ScratchSeen<uint16_t> seen{c16};
      if (two)
        for (uint32_t v = 0; v < e.n_local; ++v) blk[v] = walk_g[walk_slot(P, e, m) + v];
      else
        split_tokens(st.data(), e.start_bit, e, blk.data(), seen, tris.data() + 3ull * e.t_off);
      for (int d = 0; d < 3; ++d) {
        if (!split_axis(st.data(), (uint32_t)n_words, e.start_bit, e, P, d, blk.data(), c16 + cs * d))
          rc = 3;
        for (uint32_t v = 0; v < e.n_local; ++v)
          codes3[3 * ((size_t)e.v_off + v) + d] = c16[cs * d + v];
      }
    } else if (split) {
      std::vector<uint32_t> blk(split_block_words(e.n_local, false), two ? 0xA5A5A5A5u : 0u);
      int32_t* c32 = reinterpret_cast<int32_t*>(blk.data() + e.n_local);
      ScratchSeen<int32_t> seen{c32};
      if (two)
        for (uint32_t v = 0; v < e.n_local; ++v) blk[v] = walk_g[walk_slot(P, e, m) + v];
      else
        split_tokens(st.data(), e.start_bit, e, blk.data(), seen, tris.data() + 3ull * e.t_off);
      for (int d = 0; d < 3; ++d) {
        int32_t* cd = c32 + (size_t)e.n_local * d;
        if (!split_axis(st.data(), (uint32_t)n_words, e.start_bit, e, P, d, blk.data(), cd))
          rc = 3;
        for (uint32_t v = 0; v < e.n_local; ++v) codes3[3 * ((size_t)e.v_off + v) + d] = cd[v];
      }
    } else if (!decode_lane(st.data(), (uint32_t)n_words, e.start_bit, e, P,
                            codes.data() + e.v_off, tris.data() + 3ull * e.t_off)) {
      rc = 3;
    }
  }
  std::vector<float> V(3 * nv);
  for (uint64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const int32_t c = split ? codes3[3 * i + a] : field21(codes[i], a);
      const uint64_t mx = (1ull << P.g_bits[a]) - 1;
      const double vn = mx == 0 ? (double)P.g_min[a]
                                : (double)P.g_min[a] + ((double)c / (double)mx) * (double)P.g_range[a];
      V[3 * i + a] = (float)(vn * (double)P.scale + (double)P.center[a]);
    }
  FILE* fv = std::fopen(argv[2], "wb");
  FILE* ft = std::fopen(argv[3], "wb");
  if (!fv || !ft) return 1;
  std::fwrite(V.data(), 4, V.size(), fv);
  std::fwrite(tris.data(), 4, tris.size(), ft);
  std::fclose(fv);
  std::fclose(ft);
  return rc;
}
