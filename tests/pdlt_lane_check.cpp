// Host build of the PDLT v5 decoder's set-up and per-lane logic
// (meshpress_amd/csrc/pdlt_host.h, pdlt_lane.h): the header and boundary
// table parse, the meshlet pre-pass, the wave-group layout and, per group,
// the kernel's four phases (token lanes, boundary gather, axis lanes,
// output) over an LDS image of exactly the launch's size, under the
// sanitizers when built with them.
//
//   pdlt_lane_check <stream> <verts.f32 out> <tris.u32 out>
//                   [group_size lds_bytes [normalised [chain32 [two_pass]]]]
//
// two_pass = 1 runs the two-pass decode: the token pass over token groups
// of up to 64 meshlets (walks and seen sets to global arrays), then the
// chain pass that copies them back into the group blocks.  two_pass = 2
// (with k0 = the next argument, default 16) runs the chain token pass
// instead of the token lanes: the chain waves' layout and chain cuts
// (pdlt_layout_chains), the refs and one ChainLane per chain over each
// wave's LDS image, then the per-meshlet record resolution the kernel's
// wave performs, restated serially.
//
// Exit status: 0 ok, 2 stream rejected (message on stderr).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../meshpress_amd/csrc/pdlt_host.h"

using namespace stride;

int main(int argc, char** argv) {
  if (argc < 4) return 1;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 1;
  std::vector<uint8_t> data;
  uint8_t buf[65536];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + r);
  std::fclose(f);
  const int gs = argc > 5 ? std::atoi(argv[4]) : 16;
  const int budget = argc > 5 ? std::atoi(argv[5]) : 32768;
  const bool world = !(argc > 6 && std::atoi(argv[6]) == 1);
  const bool chain32 = argc > 7 && std::atoi(argv[7]) == 1;
  const int tp = argc > 8 ? std::atoi(argv[8]) : 0;
  const bool two_pass = tp == 1 || tp == 2;
  const int k0 = argc > 9 ? std::atoi(argv[9]) : 16;
  PdltHost H;
  std::string err;
  int rc = pdlt_parse(data.data(), data.size(), H, err);
  if (rc) {
    std::fprintf(stderr, "rejected (%d): %s\n", rc, err.c_str());
    return 2;
  }
  std::vector<PdltGroup> groups;
  uint32_t lds_bytes = 0;
  // the chain pass's blocks hold a refs slot only where pdlt_chain_block_refs keeps one
  const bool slot = !two_pass || pdlt_chain_block_refs(chain32);
  rc = pdlt_layout(H.tab, gs, budget, groups, lds_bytes, err, chain32 ? 3u : 6u, slot);
  if (rc) {
    std::fprintf(stderr, "layout (%d): %s\n", rc, err.c_str());
    return 2;
  }
  PdltParams P = H.params;
  P.world = world ? 1 : 0;
  P.chain32 = chain32 ? 1 : 0;
  std::vector<float> V(3 * H.n_v);
  std::vector<uint32_t> T(3 * H.n_t);
  for (size_t i = 0; i < H.bnd.size(); ++i) V[i] = pdlt_out(pdlt_bnd(H.bnd[i], P), P, (int)(i % 3));
  // two-pass: the token pass first, into exactly sized global arrays
  std::vector<uint32_t> walk_g(H.n_v - H.n_b + H.n_m), seen_g((size_t)kSeenWords * H.n_m);
  // the refs buffer the token pass fills (each wave's refs region copied out)
  std::vector<uint32_t> refs_g(H.n_refs, 0xDEADBEEFu);
  auto refs_out = [&](const uint32_t* lds, const PdltGroup& G) {
    for (uint32_t i = 0; i < G.seen_base - G.blk_base; ++i)
      refs_g.at(H.tab[G.m0].ref_base + i) = lds[G.blk_base + i];
  };
  if (tp == 2) {
    std::vector<PdltGroup> tg;
    std::vector<ChainEntry> ch;
    std::vector<RefChunk> rch;
    uint32_t tlds = 0;
    rc = pdlt_layout_chains(H.tab, H.words, k0, 64, tg, ch, rch, tlds, err);
    if (rc) {
      std::fprintf(stderr, "chain layout (%d): %s\n", rc, err.c_str());
      return 2;
    }
    for (const PdltGroup& G : tg) {
      std::vector<uint64_t> lds64((tlds + 7) / 8, 0x5A5A5A5A5A5A5A5Aull);
      uint32_t* lds = reinterpret_cast<uint32_t*>(lds64.data());
      // exactly the staged words: reads past them see the refs and records
      for (uint32_t i = 0; i < G.n_words; ++i) lds[i] = __builtin_bswap32(H.words.at(G.w_lo + i));
      const int64_t rel = -32 * (int64_t)G.w_lo;
      uint32_t c = G.pad[0];
      for (uint32_t m = G.m0; m < G.m1; ++m) {
        const PdltEntry& e = H.tab[m];
        const uint32_t nb = e.n_bnd, nl = e.n_bnd + e.n_int, nt = pdlt_tokens_of(e);
        const uint32_t lm = pdlt_chain_tokens(e, (uint32_t)k0), cnt = chain_count(nt, lm);
        uint32_t* refs = lds + G.blk_base + e.tok_ref_off;
        // the refs chunk by chunk, as the chunk lanes decode them
        for (uint32_t q = 0; q < pdlt_ref_chunks(nb); ++q) {
          const RefChunk rk = rch.at(e.chunk_base + q);
          if ((rk.pos >> 16) != q) return 5;
          pdlt_refs_chunk(lds, rel, e.start_bit, nb, rk, refs);
        }
        // phase 1: every chain on its own
        std::vector<uint16_t> rec(nt);
        std::vector<uint8_t> fin(16 * cnt);
        struct Sink {
          std::vector<uint16_t>* r;
          void operator()(uint32_t t, uint32_t x) const { r->at(t) = (uint16_t)x; }
        } sink{&rec};
        for (uint32_t q = 0; q < cnt; ++q, ++c) {
          const ChainEntry ce = ch.at(c);
          if (((ce.state >> 24) & 15u) != q) return 5;
          LdsCursor cur{lds, (uint32_t)((int64_t)e.strips_bit + rel) + (ce.pos & 0xFFFFu)};
          ChainLane<LdsCursor, true> cl;
          cl.init(cur, ce, chain_end(ce, nt, lm), nl, e.n_strips);
          while (!cl.done()) cl.step<true>(cur, sink);
          if ((cur.pc >> 5) + 1 >= G.n_words) return 5;   // the last window read stays staged
          for (int i = 0; i < 16; ++i)
            fin[16 * q + i] =
                (uint8_t)((i < 8 ? cl.fifo.lo >> (8 * i) : cl.fifo.hi >> (8 * (i - 8))) & 0xFFu);
        }
        // phase 2: placeholders, first sightings, triangles, walk, seen set
        for (uint32_t q = 1; q + 1 < cnt; ++q)
          for (int i = 0; i < 16; ++i) {
            const uint32_t x = fin[16 * q + i];
            if (x >= nl) fin[16 * q + i] = fin[16 * (q - 1) + ((x - nl) & 15u)];
          }
        std::vector<uint32_t> first(256, 0xFFFFFFFFu);
        for (uint32_t t = 0; t < nt; ++t) {
          uint32_t v = rec[t] & 0xFFu;
          if (v >= nl) {
            const uint32_t q = rec[t] >> 12;
            if (q == 0) return 5;
            v = fin[16 * (q - 1) + ((v - nl) & 15u)];
            rec[t] = (uint16_t)((rec[t] & 0xFF00u) | v);
          }
          if (v >= nl) return 5;
          first[v] = std::min(first[v], t);
        }
        auto gid = [&](uint32_t x) { return x < nb ? refs[x] : e.int_base - nb + x; };
        uint32_t* wout = walk_g.data() + pdlt_walk_off(e.int_base, m, H.n_b);
        uint32_t* tout = T.data() + 3ull * e.t_off;
        int last = -1;
        uint32_t wi = 0, ti = 0;
        for (uint32_t t = 0; t < nt; ++t) {
          const uint32_t r = rec[t], v = r & 0xFFu;
          const uint32_t v1 = rec[t >= 1 ? t - 1 : 0] & 0xFFu, v2 = rec[t >= 2 ? t - 2 : 0] & 0xFFu;
          const uint32_t p0 = rec[last >= 2 ? (uint32_t)last - 2 : 0] & 0xFFu;
          uint32_t a, b;
          const uint32_t entry = chain_entry(r, v1, v2, p0, a, b);
          if (v >= nb && first[v] == t) wout[wi++] = entry;
          if (chain_emits(r)) {
            tout[3 * ti] = gid(a);
            tout[3 * ti + 1] = gid(b);
            tout[3 * ti + 2] = gid(v);
            ++ti;
          }
          if ((r & kRecEc) == 0) last = (int)t;
        }
        if (ti != e.n_tris) return 5;
        uint32_t* so = seen_g.data() + (size_t)kSeenWords * m;
        for (uint32_t q = 0; q < 8; ++q) so[q] = 0;
        for (uint32_t i = 0; i < 256; ++i)
          if (i < nb || first[i] != 0xFFFFFFFFu) so[i >> 5] |= 1u << (i & 31);
        so[8] = wi;
      }
      if (c != G.pad[1]) return 5;
      refs_out(lds, G);
    }
  } else if (two_pass) {
    std::vector<PdltGroup> tg;
    uint32_t tlds = 0;
    rc = pdlt_layout_tokens(H.tab, budget, 64, tg, tlds, err);
    if (rc) {
      std::fprintf(stderr, "token layout (%d): %s\n", rc, err.c_str());
      return 2;
    }
    for (const PdltGroup& G : tg) {
      std::vector<uint64_t> lds64((tlds + 7) / 8, 0x5A5A5A5A5A5A5A5Aull);
      uint32_t* lds = reinterpret_cast<uint32_t*>(lds64.data());
      for (uint32_t i = 0; i < G.n_words; ++i) lds[i] = __builtin_bswap32(H.words.at(G.w_lo + i));
      for (uint32_t m = G.m0; m < G.m1; ++m) {
        const PdltEntry& e = H.tab[m];
        uint32_t* refs = lds + G.blk_base + e.tok_ref_off;
        uint32_t* seen = lds + G.seen_base + (m - G.m0);
        uint32_t n_emit = 0;
        pdlt_tokens<WalkSlot>(lds, -32 * (int64_t)G.w_lo, e, refs,
                              walk_g.data() + pdlt_walk_off(e.int_base, m, H.n_b),
                              BitSeen{seen, kTokenSeenStride}, &n_emit, T.data() + 3ull * e.t_off);
        uint32_t* so = seen_g.data() + (size_t)kSeenWords * m;
        for (int q = 0; q < 8; ++q) so[q] = seen[q * kTokenSeenStride];
        so[8] = n_emit;
      }
      refs_out(lds, G);
    }
  }
  // the kernels' block: three arrays of Tt per local (double for the float64
  // chain, float for the float32 chain), then refs (the one-kernel decode;
  // the chain pass reads them from the refs buffer and may keep the slot),
  // walk, seen set
  auto run_groups = [&](auto tag) {
  using Tt = decltype(tag);
  constexpr uint32_t RW = 3 * sizeof(Tt) / 4;
  for (const PdltGroup& G : groups) {
    std::vector<uint64_t> lds64((lds_bytes + 7) / 8, 0x5A5A5A5A5A5A5A5Aull);   // stale LDS
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds64.data());
    for (uint32_t i = 0; i < G.n_words; ++i) lds[i] = __builtin_bswap32(H.words.at(G.w_lo + i));
    const int64_t rel = -32 * (int64_t)G.w_lo;
    auto blk = [&](const PdltEntry& e) { return lds + G.blk_base + e.blk_off; };
    // phase A: token lanes, or the copy-back of the token pass and the refs
    for (uint32_t m = G.m0; m < G.m1; ++m) {
      const PdltEntry& e = H.tab[m];
      uint32_t* b = blk(e);
      const uint32_t nl = e.n_bnd + e.n_int;
      uint32_t* refs = b + RW * nl;
      uint32_t* walk = refs + (slot ? e.n_bnd : 0u);
      uint32_t* seen = walk + e.n_int;
      if (!two_pass) {
        pdlt_tokens(lds, rel, e, refs, walk, BitSeen{seen}, seen + 8, T.data() + 3ull * e.t_off);
        continue;
      }
      for (uint32_t i = 0; i < e.n_int; ++i) walk[i] = walk_g.at(pdlt_walk_off(e.int_base, m, H.n_b) + i);
      for (uint32_t i = 0; i < kSeenWords; ++i) seen[i] = seen_g.at((size_t)kSeenWords * m + i);
    }
    // phase B: boundary reconstructions
    for (uint32_t m = G.m0; m < G.m1; ++m) {
      const PdltEntry& e = H.tab[m];
      uint32_t* b = blk(e);
      const uint32_t nl = e.n_bnd + e.n_int;
      Tt* recon = reinterpret_cast<Tt*>(b);
      // the chain pass gathers from the token pass's refs
      const uint32_t* refs = two_pass ? refs_g.data() + e.ref_base : b + RW * nl;
      for (uint32_t i = 0; i < e.n_bnd; ++i)
        for (int d = 0; d < 3; ++d)
          recon[d * nl + i] = (Tt)pdlt_bnd(H.bnd.at(3ull * refs[i] + d), P);
    }
    // phase C: axis lanes
    for (uint32_t m = G.m0; m < G.m1; ++m) {
      const PdltEntry& e = H.tab[m];
      uint32_t* b = blk(e);
      const uint32_t nl = e.n_bnd + e.n_int;
      Tt* recon = reinterpret_cast<Tt*>(b);
      const uint32_t* walk = b + RW * nl + (slot ? e.n_bnd : 0u);
      const uint32_t n_emit = walk[e.n_int + 8];
      for (int d = 0; d < 3; ++d) pdlt_axis(lds, rel, e, d, P, recon + d * nl, walk, n_emit);
    }
    // phase D: interior vertices
    for (uint32_t m = G.m0; m < G.m1; ++m) {
      const PdltEntry& e = H.tab[m];
      const uint32_t* b = blk(e);
      const uint32_t nl = e.n_bnd + e.n_int;
      const Tt* rj = reinterpret_cast<const Tt*>(b);
      const uint32_t* sj = b + RW * nl + (slot ? e.n_bnd : 0u) + e.n_int;
      for (uint32_t t = 0; t < 3 * e.n_int; ++t) {
        const uint32_t i = t / 3, d = t - 3 * i, v = e.n_bnd + i;
        const bool reached = ((sj[v >> 5] >> (v & 31)) & 1u) != 0;
        V.at(3ull * e.int_base + t) = pdlt_out(reached ? (double)rj[d * nl + v] : 0.0, P, (int)d);
      }
    }
  }
  };
  if (chain32)
    run_groups(float{});
  else
    run_groups(double{});
  FILE* fv = std::fopen(argv[2], "wb");
  FILE* ft = std::fopen(argv[3], "wb");
  if (!fv || !ft) return 1;
  std::fwrite(V.data(), 4, V.size(), fv);
  std::fwrite(T.data(), 4, T.size(), ft);
  std::fclose(fv);
  std::fclose(ft);
  return 0;
}
