// STRIDE encoder front-end and DPR4 v3 writer (CPU, C++17).
//
// Restates, in native code, the reference's encoder front-end
// (encoder/paradelta_codec.py:758-882 prepare_paradelta_arrays, with its
// Numba callees) and the DPR4 writer (encoder/paradelta_v5_dup.py:239-464).
// The floating-point inputs that decide tie-breaks (normalised positions,
// face normals, cost constants) are computed by the Python host with numpy,
// exactly as the reference does, and passed in.
#include "stride_host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

namespace stride {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------

// Runs fn(i, thread) for i in [0, n) on up to n_threads threads, one per
// 128 items at most: a thread's start and join cost tens of microseconds,
// more than a small call's whole work.
template <class F>
static void parallel_for(int64_t n, int n_threads, F&& fn) {
  const int nt = (int)std::min<int64_t>(n_threads, n / 128);
  if (nt <= 1) {
    for (int64_t i = 0; i < n; ++i) fn(i, 0);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  const int64_t chunk = 16;
  for (int t = 0; t < nt; ++t) {
    th.emplace_back([&, t]() {
      for (;;) {
        int64_t s = next.fetch_add(chunk);
        if (s >= n) break;
        int64_t e = std::min(n, s + chunk);
        for (int64_t i = s; i < e; ++i) fn(i, t);
      }
    });
  }
  for (auto& x : th) x.join();
}

// Stable LSD radix sort of (key, payload) by key.
static void radix_sort_pairs(std::vector<uint64_t>& key, std::vector<int64_t>& val) {
  const size_t n = key.size();
  uint64_t maxk = 0;
  for (uint64_t k : key) maxk = std::max(maxk, k);
  std::vector<uint64_t> k2(n);
  std::vector<int64_t> v2(n);
  std::vector<size_t> cnt(1 << 16);
  for (int shift = 0; shift < 64 && (maxk >> shift) != 0; shift += 16) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; ++i) cnt[(key[i] >> shift) & 0xFFFF]++;
    size_t s = 0;
    for (auto& c : cnt) { size_t t = c; c = s; s += t; }
    for (size_t i = 0; i < n; ++i) {
      size_t d = cnt[(key[i] >> shift) & 0xFFFF]++;
      k2[d] = key[i];
      v2[d] = val[i];
    }
    key.swap(k2);
    val.swap(v2);
  }
}

// Small open-addressing int set/map (keys >= 0), sized for <= cap entries.
struct SmallMap {
  std::vector<int64_t> k;
  std::vector<int32_t> v;
  uint64_t mask = 0;
  void reset(size_t cap) {
    size_t s = 1;
    while (s < cap * 2 + 2) s <<= 1;
    k.assign(s, -1);
    v.assign(s, 0);
    mask = s - 1;
  }
  // returns slot
  size_t find_slot(int64_t key) const {
    size_t h = (uint64_t)(key * 2654435761ULL) & mask;
    while (k[h] != -1 && k[h] != key) h = (h + 1) & mask;
    return h;
  }
  bool contains(int64_t key) const { return k[find_slot(key)] == key; }
  int32_t get(int64_t key, int32_t dflt) const {
    size_t s = find_slot(key);
    return k[s] == key ? v[s] : dflt;
  }
  bool insert(int64_t key, int32_t val) {  // true if newly inserted
    size_t s = find_slot(key);
    if (k[s] == key) return false;
    k[s] = key;
    v[s] = val;
    return true;
  }
};

// ---------------------------------------------------------------------------
// Global triangle adjacency (utils/meshlet_plan_nb.py:19-101) and unique edge
// ids (utils/meshlet_gen_joint_nb.py:506-525).
//
// Edges are listed in three blocks, (v0,v1) for all tris, then (v1,v2), then
// (v0,v2); a stable sort by (min,max) key puts equal edges next to each other
// in input order, and every consecutive equal pair becomes one adjacency
// pair.  (The reference forms the stable order from an int64 composite
// key*3n+i, which wraps for meshes with more than about 1.1M vertices; this
// restatement uses the intended stable order at every size.)
// ---------------------------------------------------------------------------
static void global_adjacency(const int64_t* T, int64_t n, std::vector<int32_t>& off,
                             std::vector<int32_t>& idx, std::vector<int32_t>& edge_id) {
  const int64_t ne = 3 * n;
  std::vector<uint64_t> key(ne);
  std::vector<int64_t> ord(ne);
  int64_t max_v = 0;
  auto lohi = [&](int64_t e, int64_t& lo, int64_t& hi) {
    int64_t t = e % n, b = e / n;
    int64_t a = T[3 * t], bb = T[3 * t + 1], c = T[3 * t + 2];
    int64_t u, v;
    if (b == 0) { u = a; v = bb; } else if (b == 1) { u = bb; v = c; } else { u = a; v = c; }
    lo = std::min(u, v);
    hi = std::max(u, v);
  };
  for (int64_t e = 0; e < ne; ++e) {
    int64_t lo, hi;
    lohi(e, lo, hi);
    max_v = std::max(max_v, hi);
  }
  const uint64_t mult = (uint64_t)max_v + 1;
  for (int64_t e = 0; e < ne; ++e) {
    int64_t lo, hi;
    lohi(e, lo, hi);
    key[e] = (uint64_t)lo * mult + (uint64_t)hi;
    ord[e] = e;
  }
  radix_sort_pairs(key, ord);
  std::vector<int32_t> deg(n, 0);
  std::vector<int32_t> pa, pb;
  pa.reserve(ne / 2);
  pb.reserve(ne / 2);
  edge_id.assign(ne, 0);
  int32_t gid = 0;
  for (int64_t i = 0; i < ne; ++i) {
    if (i > 0 && key[i] != key[i - 1]) ++gid;
    int64_t e = ord[i];
    edge_id[(e % n) * 3 + e / n] = gid;
    if (i + 1 < ne && key[i] == key[i + 1]) {
      int32_t ta = (int32_t)(ord[i] % n), tb = (int32_t)(ord[i + 1] % n);
      pa.push_back(ta);
      pb.push_back(tb);
      deg[ta]++;
      deg[tb]++;
    }
  }
  off.assign(n + 1, 0);
  for (int64_t t = 0; t < n; ++t) off[t + 1] = off[t] + deg[t];
  idx.assign(off[n], 0);
  std::vector<int32_t> cur(off.begin(), off.end() - 1);
  for (size_t p = 0; p < pa.size(); ++p) {
    idx[cur[pa[p]]++] = pb[p];
    idx[cur[pb[p]]++] = pa[p];
  }
}

// ---------------------------------------------------------------------------
// Joint-cost greedy region growing (utils/meshlet_gen_joint_nb.py:31-430).
// ---------------------------------------------------------------------------
static double min_eig(double m00, double m01, double m02, double m11, double m12,
                      double m22) {
  double p1 = m01 * m01 + m02 * m02 + m12 * m12;
  double q = (m00 + m11 + m22) / 3.0;
  if (p1 < 1e-30) {
    double a = m00;
    if (m11 < a) a = m11;
    if (m22 < a) a = m22;
    return a;
  }
  double p2 = ((m00 - q) * (m00 - q) + (m11 - q) * (m11 - q) + (m22 - q) * (m22 - q) +
               2.0 * p1);
  double p = std::sqrt(p2 / 6.0);
  if (p < 1e-30) return q;
  double b00 = (m00 - q) / p, b11 = (m11 - q) / p, b22 = (m22 - q) / p;
  double b01 = m01 / p, b02 = m02 / p, b12 = m12 / p;
  double detB = (b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
                 b02 * (b01 * b12 - b11 * b02));
  double r = 0.5 * detB;
  if (r < -1.0) r = -1.0;
  else if (r > 1.0) r = 1.0;
  double phi = std::acos(r) / 3.0;
  return q + 2.0 * p * std::cos(phi + 2.0 * M_PI / 3.0);
}

struct Partitioner {
  const int64_t* T;
  const double* Vp;
  const double* FN;
  const int32_t* aoff;
  const int32_t* aidx;
  const int32_t* eid;
  int64_t n_t, n_v;
  int max_tris, max_verts;
  double K[6], NM[6];

  std::vector<uint8_t> visited, vert_in, tri_in, bnd_edge_in, in_frontier;
  std::vector<int32_t> depth, frontier;

  bool adj_has(int64_t t, int64_t x) const {
    for (int32_t k = aoff[t]; k < aoff[t + 1]; ++k)
      if (aidx[k] == x) return true;
    return false;
  }

  int grow(int64_t seed, std::vector<int32_t>& out) {
    visited[seed] = 1;
    tri_in[seed] = 1;
    depth[seed] = 0;
    int64_t sv[3] = {T[3 * seed], T[3 * seed + 1], T[3 * seed + 2]};
    for (int j = 0; j < 3; ++j) vert_in[sv[j]] = 1;
    int n_verts_in = 3;
    for (int j = 0; j < 3; ++j) bnd_edge_in[eid[3 * seed + j]] = 1;
    const double* p0 = Vp + 3 * sv[0];
    const double* p1 = Vp + 3 * sv[1];
    const double* p2 = Vp + 3 * sv[2];
    double sx = p0[0] + p1[0] + p2[0];
    double sy = p0[1] + p1[1] + p2[1];
    double sz = p0[2] + p1[2] + p2[2];
    double o00 = p0[0] * p0[0] + p1[0] * p1[0] + p2[0] * p2[0];
    double o11 = p0[1] * p0[1] + p1[1] * p1[1] + p2[1] * p2[1];
    double o22 = p0[2] * p0[2] + p1[2] * p1[2] + p2[2] * p2[2];
    double o01 = p0[0] * p0[1] + p1[0] * p1[1] + p2[0] * p2[1];
    double o02 = p0[0] * p0[2] + p1[0] * p1[2] + p2[0] * p2[2];
    double o12 = p0[1] * p0[2] + p1[1] * p1[2] + p2[1] * p2[2];
    double an_x = FN[3 * seed], an_y = FN[3 * seed + 1], an_z = FN[3 * seed + 2];
    int64_t tail = seed;
    int f_count = 0;
    for (int32_t k = aoff[seed]; k < aoff[seed + 1]; ++k) {
      int32_t nb = aidx[k];
      if (!visited[nb]) {
        frontier[f_count++] = nb;
        in_frontier[nb] = 1;
      }
    }
    out.clear();
    out.push_back((int32_t)seed);

    while (f_count > 0 && (int)out.size() < max_tris && n_verts_in < max_verts) {
      int64_t best_cand = -1;
      double best_cost = 1e300;
      for (int fi = 0; fi < f_count; ++fi) {
        int64_t cand = frontier[fi];
        if (visited[cand]) continue;
        int64_t c0 = T[3 * cand], c1 = T[3 * cand + 1], c2 = T[3 * cand + 2];
        int new_v_count = 0;
        if (!vert_in[c0]) new_v_count++;
        if (!vert_in[c1] && c1 != c0) new_v_count++;
        if (!vert_in[c2] && c2 != c0 && c2 != c1) new_v_count++;
        if (n_verts_in + new_v_count > max_verts) continue;

        // f1: plane residual of the hypothetical vertex set
        double ns_x = sx, ns_y = sy, ns_z = sz;
        double n00 = o00, n11 = o11, n22 = o22, n01 = o01, n02 = o02, n12 = o12;
        int n_new = 0;
        bool seen0 = false, seen1 = false;
        auto add = [&](int64_t v) {
          const double* p = Vp + 3 * v;
          double px = p[0], py = p[1], pz = p[2];
          ns_x += px; ns_y += py; ns_z += pz;
          n00 += px * px; n11 += py * py; n22 += pz * pz;
          n01 += px * py; n02 += px * pz; n12 += py * pz;
          n_new++;
        };
        if (!vert_in[c0]) { add(c0); seen0 = true; }
        if (!vert_in[c1] && !(seen0 && c1 == c0)) { add(c1); seen1 = true; }
        if (!vert_in[c2]) {
          bool dup = (seen0 && c2 == c0) || (seen1 && c2 == c1);
          if (!dup) add(c2);
        }
        int n_total = n_verts_in + n_new;
        double f1;
        if (n_total < 4) {
          f1 = 0.0;
        } else {
          double inv_n = 1.0 / n_total;
          double mx_ = ns_x * inv_n, my_ = ns_y * inv_n, mz_ = ns_z * inv_n;
          double denom = (double)(n_total - 1);
          if (denom < 1.0) denom = 1.0;
          double c00 = (n00 - n_total * mx_ * mx_) / denom;
          double c11 = (n11 - n_total * my_ * my_) / denom;
          double c22 = (n22 - n_total * mz_ * mz_) / denom;
          double c01 = (n01 - n_total * mx_ * my_) / denom;
          double c02 = (n02 - n_total * mx_ * mz_) / denom;
          double c12 = (n12 - n_total * my_ * mz_) / denom;
          double emin = min_eig(c00, c01, c02, c11, c12, c22);
          if (emin < 0.0) emin = 0.0;
          f1 = std::sqrt(emin);
        }
        // f2: perimeter delta
        int added = 0, removed = 0;
        for (int j = 0; j < 3; ++j) {
          if (bnd_edge_in[eid[3 * cand + j]]) removed++;
          else added++;
        }
        double f2 = (double)(added - removed);
        // f3: strip continuation
        double f3 = adj_has(tail, cand) ? 0.0 : 1.0;
        // f4: normal dissimilarity
        double nn2 = an_x * an_x + an_y * an_y + an_z * an_z;
        double f4;
        if (nn2 < 1e-24) {
          f4 = 1.0;
        } else {
          const double* fnc = FN + 3 * cand;
          double cs = (fnc[0] * an_x + fnc[1] * an_y + fnc[2] * an_z) / std::sqrt(nn2);
          f4 = 1.0 - cs;
        }
        // f5: shared edges; f6: bfs depth
        int shared = 0;
        int min_d = -1;
        for (int32_t k = aoff[cand]; k < aoff[cand + 1]; ++k) {
          int32_t nb = aidx[k];
          if (tri_in[nb]) {
            shared++;
            int d = depth[nb];
            if (min_d < 0 || d < min_d) min_d = d;
          }
        }
        double f5 = 1.0 - shared / 3.0;
        if (min_d < 0) min_d = 0;
        double f6 = (double)(min_d + 1);
        double cost = (K[0] * (f1 + NM[0]) + K[1] * (f2 + NM[1]) + K[2] * (f3 + NM[2]) +
                       K[3] * (f4 + NM[3]) + K[4] * (f5 + NM[4]) + K[5] * (f6 + NM[5]));
        if (cost < best_cost) {
          best_cost = cost;
          best_cand = cand;
        }
      }
      if (best_cand < 0) break;
      int64_t cand = best_cand;
      visited[cand] = 1;
      tri_in[cand] = 1;
      for (int j = 0; j < 3; ++j) {
        int32_t e = eid[3 * cand + j];
        bnd_edge_in[e] = !bnd_edge_in[e];
      }
      for (int j = 0; j < 3; ++j) {
        int64_t v = T[3 * cand + j];
        if (!vert_in[v]) {
          vert_in[v] = 1;
          const double* p = Vp + 3 * v;
          double px = p[0], py = p[1], pz = p[2];
          sx += px; sy += py; sz += pz;
          o00 += px * px; o11 += py * py; o22 += pz * pz;
          o01 += px * py; o02 += px * pz; o12 += py * pz;
          n_verts_in++;
        }
      }
      int cnt = (int)out.size() + 1;
      int prev = cnt - 1;
      const double* fnc = FN + 3 * cand;
      double inv_cnt = 1.0 / cnt;
      an_x = (an_x * prev + fnc[0]) * inv_cnt;
      an_y = (an_y * prev + fnc[1]) * inv_cnt;
      an_z = (an_z * prev + fnc[2]) * inv_cnt;
      if (adj_has(tail, cand)) tail = cand;
      int new_d = -1;
      for (int32_t k = aoff[cand]; k < aoff[cand + 1]; ++k) {
        int32_t nb = aidx[k];
        if (tri_in[nb] && nb != cand) {
          int d = depth[nb];
          if (new_d < 0 || d < new_d) new_d = d;
        }
      }
      if (new_d < 0) new_d = 0;
      depth[cand] = new_d + 1;
      out.push_back((int32_t)cand);
      for (int fi = 0; fi < f_count; ++fi) {
        if (frontier[fi] == cand) {
          frontier[fi] = frontier[f_count - 1];
          f_count--;
          break;
        }
      }
      in_frontier[cand] = 0;
      for (int32_t k = aoff[cand]; k < aoff[cand + 1]; ++k) {
        int32_t nb = aidx[k];
        if (!visited[nb] && !in_frontier[nb]) {
          frontier[f_count++] = nb;
          in_frontier[nb] = 1;
        }
      }
    }
    for (int fi = 0; fi < f_count; ++fi) in_frontier[frontier[fi]] = 0;
    return (int)out.size();
  }

  void run(std::vector<std::vector<int32_t>>& meshlets) {
    visited.assign(n_t, 0);
    vert_in.assign(n_v, 0);
    tri_in.assign(n_t, 0);
    depth.assign(n_t, -1);
    int32_t n_edges = 0;
    for (int64_t i = 0; i < 3 * n_t; ++i) n_edges = std::max(n_edges, eid[i] + 1);
    bnd_edge_in.assign(n_edges, 0);
    in_frontier.assign(n_t, 0);
    frontier.assign((size_t)max_tris * 4 + 32, 0);
    // seed order: stable argsort by degree (counting sort)
    int32_t maxdeg = 0;
    for (int64_t t = 0; t < n_t; ++t) maxdeg = std::max(maxdeg, aoff[t + 1] - aoff[t]);
    std::vector<int64_t> cnt(maxdeg + 2, 0);
    for (int64_t t = 0; t < n_t; ++t) cnt[aoff[t + 1] - aoff[t] + 1]++;
    for (int d = 1; d <= maxdeg + 1; ++d) cnt[d] += cnt[d - 1];
    std::vector<int32_t> seed_order(n_t);
    for (int64_t t = 0; t < n_t; ++t) seed_order[cnt[aoff[t + 1] - aoff[t]]++] = (int32_t)t;

    std::vector<int32_t> out;
    out.reserve(max_tris + 8);
    int64_t seed_ptr = 0;
    while (seed_ptr < n_t) {
      while (seed_ptr < n_t && visited[seed_order[seed_ptr]]) seed_ptr++;
      if (seed_ptr >= n_t) break;
      int64_t seed = seed_order[seed_ptr++];
      grow(seed, out);
      meshlets.push_back(out);
      for (int32_t t : out) {
        tri_in[t] = 0;
        depth[t] = -1;
        for (int j = 0; j < 3; ++j) {
          bnd_edge_in[eid[3 * t + j]] = 0;
          vert_in[T[3 * t + j]] = 0;
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// CPython set iteration order (Objects/setobject.c: set_add_entry,
// set_table_resize, set_insert_clean; LINEAR_PROBES 9, PERTURB_SHIFT 5).
// The multiseed strip cover iterates a Python set of seeds and keeps the
// first strictly better result, so the visiting order decides ties.
// ---------------------------------------------------------------------------
std::vector<int> pyset_order(const std::vector<int>& inserted) {
  std::vector<int64_t> table(8, -1);
  size_t mask = 7, fill = 0, used = 0;
  auto insert_clean = [](std::vector<int64_t>& tb, size_t msk, int64_t key) {
    size_t perturb = (size_t)key;
    size_t i = (size_t)key & msk;
    for (;;) {
      if (tb[i] == -1) { tb[i] = key; return; }
      if (i + 9 <= msk) {
        for (size_t j = 1; j <= 9; ++j)
          if (tb[i + j] == -1) { tb[i + j] = key; return; }
      }
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & msk;
    }
  };
  for (int key : inserted) {
    size_t i = (size_t)key & mask;
    size_t perturb = (size_t)key;
    int64_t* slot = nullptr;
    bool present = false;
    if (table[i] == -1) {
      slot = &table[i];
    } else {
      for (;;) {
        if (table[i] == key) { present = true; break; }
        bool done = false;
        if (i + 9 <= mask) {
          for (size_t j = 1; j <= 9; ++j) {
            if (table[i + j] == -1) { slot = &table[i + j]; done = true; break; }
            if (table[i + j] == key) { present = true; done = true; break; }
          }
        }
        if (done) break;
        perturb >>= 5;
        i = (i * 5 + 1 + perturb) & mask;
        if (table[i] == -1) { slot = &table[i]; break; }
      }
    }
    if (present) continue;
    *slot = key;
    fill++;
    used++;
    if (fill * 5 >= mask * 3) {
      size_t minused = used > 50000 ? used * 2 : used * 4;
      size_t newsize = 8;
      while (newsize <= minused) newsize <<= 1;
      std::vector<int64_t> nt(newsize, -1);
      for (int64_t k : table)
        if (k != -1) insert_clean(nt, newsize - 1, k);
      table.swap(nt);
      mask = newsize - 1;
    }
  }
  std::vector<int> order;
  for (int64_t k : table)
    if (k != -1) order.push_back((int)k);
  return order;
}

// ---------------------------------------------------------------------------
// Per-meshlet plan (encoder/paradelta_codec.py:347-439 with
// utils/meshlet_plan_nb.py eb_vertex_order_nb / greedy_nn_order_nb /
// local_face_adj_csr_nb and utils/amd_gts_nb.py multiseed strips).
// ---------------------------------------------------------------------------
struct PlanWork {
  SmallMap tri_map, seen, g2l;
};

static void eb_vertex_order(const std::vector<int32_t>& mt, const int64_t* T,
                            const int32_t* aoff, const int32_t* aidx, PlanWork& W,
                            std::vector<int64_t>& order) {
  const int n = (int)mt.size();
  W.tri_map.reset(n);
  for (int li = 0; li < n; ++li) W.tri_map.insert(mt[li], li);
  W.seen.reset(3 * n + 8);
  order.clear();
  std::vector<uint8_t> vis(n, 0);
  std::vector<int32_t> queue(n);
  for (int s = 0; s < n; ++s) {
    if (vis[s]) continue;
    int64_t sg = mt[s];
    vis[s] = 1;
    int qh = 0, qt = 0;
    queue[qt++] = (int32_t)sg;
    for (int j = 0; j < 3; ++j) {
      int64_t v = T[3 * sg + j];
      if (W.seen.insert(v, 0)) order.push_back(v);
    }
    while (qh < qt) {
      int32_t cur = queue[qh++];
      for (int32_t k = aoff[cur]; k < aoff[cur + 1]; ++k) {
        int32_t nb = aidx[k];
        int32_t li = W.tri_map.get(nb, -1);
        if (li >= 0 && !vis[li]) {
          vis[li] = 1;
          queue[qt++] = nb;
          for (int j = 0; j < 3; ++j) {
            int64_t v = T[3 * (int64_t)nb + j];
            if (W.seen.insert(v, 0)) order.push_back(v);
          }
        }
      }
    }
  }
  // Unreached vertices (cannot happen once every tri is visited; kept for
  // parity with the reference's fallback): appended in ascending id order.
  std::vector<int64_t> extra;
  for (int t = 0; t < n; ++t)
    for (int j = 0; j < 3; ++j) {
      int64_t v = T[3 * (int64_t)mt[t] + j];
      if (W.seen.insert(v, 0)) extra.push_back(v);
    }
  std::sort(extra.begin(), extra.end());
  order.insert(order.end(), extra.begin(), extra.end());
}

static void greedy_nn(std::vector<int64_t>& ids, const double* vn) {
  const int n = (int)ids.size();
  if (n <= 1) return;
  std::vector<double> pts(3 * n);
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) pts[3 * i + d] = vn[3 * ids[i] + d];
  std::vector<uint8_t> vis(n, 0);
  std::vector<int64_t> out(n);
  int cur = 0;
  vis[0] = 1;
  out[0] = ids[0];
  for (int k = 1; k < n; ++k) {
    double best_d = 1e30;
    int best_i = -1;
    for (int i = 0; i < n; ++i) {
      if (vis[i]) continue;
      double dx = pts[3 * i] - pts[3 * cur];
      double dy = pts[3 * i + 1] - pts[3 * cur + 1];
      double dz = pts[3 * i + 2] - pts[3 * cur + 2];
      double d = dx * dx + dy * dy + dz * dz;
      if (d < best_d) { best_d = d; best_i = i; }
    }
    out[k] = ids[best_i];
    vis[best_i] = 1;
    cur = best_i;
  }
  ids.swap(out);
}

static void local_face_adj(const std::vector<uint8_t>& tl, int n_tris,
                           std::vector<int32_t>& off, std::vector<int32_t>& idx) {
  // edges in discovery order: tri-major, then (a,b), (b,c), (c,a)
  SmallMap emap;
  emap.reset(3 * n_tris + 4);
  std::vector<int32_t> first(3 * n_tris), second(3 * n_tris, -1), count(3 * n_tris, 0);
  int n_unique = 0;
  for (int t = 0; t < n_tris; ++t) {
    int a = tl[3 * t], b = tl[3 * t + 1], c = tl[3 * t + 2];
    int uv[3][2] = {{a, b}, {b, c}, {c, a}};
    for (int e = 0; e < 3; ++e) {
      int lo = std::min(uv[e][0], uv[e][1]), hi = std::max(uv[e][0], uv[e][1]);
      int64_t key = ((int64_t)lo << 32) | hi;
      size_t s = emap.find_slot(key);
      if (emap.k[s] != key) {
        emap.k[s] = key;
        emap.v[s] = n_unique;
        first[n_unique] = t;
        count[n_unique] = 1;
        n_unique++;
      } else {
        int id = emap.v[s];
        count[id]++;
        if (count[id] == 2) second[id] = t;
      }
    }
  }
  std::vector<int32_t> deg(n_tris, 0), pa, pb;
  for (int e = 0; e < n_unique; ++e)
    if (count[e] == 2) {
      pa.push_back(first[e]);
      pb.push_back(second[e]);
      deg[first[e]]++;
      deg[second[e]]++;
    }
  off.assign(n_tris + 1, 0);
  for (int t = 0; t < n_tris; ++t) off[t + 1] = off[t] + deg[t];
  idx.assign(off[n_tris], 0);
  std::vector<int32_t> cur(off.begin(), off.end() - 1);
  for (size_t p = 0; p < pa.size(); ++p) {
    idx[cur[pa[p]]++] = pb[p];
    idx[cur[pb[p]]++] = pa[p];
  }
}

// Bidirectional greedy strip cover from a forced first seed
// (utils/amd_gts_nb.py:17-126).
static void strips_seeded(int n_f, const std::vector<int32_t>& off,
                          const std::vector<int32_t>& idx, const std::vector<int32_t>& deg,
                          int seed_first, std::vector<int32_t>& lens,
                          std::vector<int32_t>& tris) {
  std::vector<uint8_t> processed(n_f, 0);
  std::vector<int32_t> cur_val(deg);
  std::vector<int32_t> scratch(n_f);
  lens.clear();
  tris.clear();
  auto process = [&](int i) {
    processed[i] = 1;
    for (int32_t k = off[i]; k < off[i + 1]; ++k)
      if (!processed[idx[k]]) cur_val[idx[k]]--;
  };
  auto best_neighbor = [&](int tri) {
    int best = -1, k0b = 0, k1b = 0;
    for (int32_t k = off[tri]; k < off[tri + 1]; ++k) {
      int nb = idx[k];
      if (processed[nb]) continue;
      int la = 0;
      for (int32_t k2 = off[nb]; k2 < off[nb + 1]; ++k2) {
        int nb2 = idx[k2];
        if (nb2 != tri && !processed[nb2]) la += cur_val[nb2];
      }
      int k0 = cur_val[nb];
      if (best < 0 || k0 < k0b || (k0 == k0b && la < k1b)) {
        best = nb;
        k0b = k0;
        k1b = la;
      }
    }
    return best;
  };
  bool first = true;
  for (;;) {
    int seed;
    if (first) {
      seed = seed_first;
      first = false;
    } else {
      int best = -1, bv = 0;
      for (int i = 0; i < n_f; ++i) {
        if (processed[i]) continue;
        if (best < 0 || cur_val[i] < bv) { best = i; bv = cur_val[i]; }
      }
      if (best < 0) break;
      seed = best;
    }
    if (processed[seed]) continue;
    int n_fwd = 0, n_bwd = 0;
    scratch[n_fwd++] = seed;
    process(seed);
    int cur = seed;
    for (;;) {
      int nb = best_neighbor(cur);
      if (nb < 0) break;
      scratch[n_fwd++] = nb;
      process(nb);
      cur = nb;
    }
    cur = seed;
    for (;;) {
      int nb = best_neighbor(cur);
      if (nb < 0) break;
      scratch[n_fwd + n_bwd++] = nb;
      process(nb);
      cur = nb;
    }
    for (int j = n_bwd - 1; j >= 0; --j) tris.push_back(scratch[n_fwd + j]);
    for (int j = 0; j < n_fwd; ++j) tris.push_back(scratch[j]);
    lens.push_back(n_fwd + n_bwd);
  }
}

static void strips_multiseed(const std::vector<uint8_t>& tl, int n_f, const PrepParams& P,
                             Plan& plan) {
  plan.strip_len.clear();
  plan.strip_tris.clear();
  if (n_f == 0) return;
  std::vector<int32_t> off, idx;
  local_face_adj(tl, n_f, off, idx);
  std::vector<int32_t> deg(n_f);
  int amin = 0, amax = 0;
  for (int t = 0; t < n_f; ++t) {
    deg[t] = off[t + 1] - off[t];
    if (deg[t] < deg[amin]) amin = t;
    if (deg[t] > deg[amax]) amax = t;
  }
  std::vector<int> ins = {amin, amax};
  if (n_f > 4) {
    int cnt = std::min(6, n_f - 2);
    if (n_f >= P.seed_rows) throw std::runtime_error("strip seed table too small");
    for (int j = 0; j < cnt; ++j) ins.push_back(P.seed_table[6 * n_f + j]);
  }
  std::vector<int> order = pyset_order(ins);
  std::vector<int32_t> lens, tris;
  size_t best_n = (size_t)-1;
  for (int s : order) {
    strips_seeded(n_f, off, idx, deg, s, lens, tris);
    if (best_n == (size_t)-1 || lens.size() < best_n) {
      best_n = lens.size();
      plan.strip_len = lens;
      plan.strip_tris = tris;
    }
  }
}

// _root_orient (encoder/paradelta_codec.py:206-216).
static void root_orient(const uint8_t* tl, int root_id, int next_id, int out[3]) {
  int r[3] = {tl[3 * root_id], tl[3 * root_id + 1], tl[3 * root_id + 2]};
  out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
  if (next_id < 0) return;
  int nx[3] = {tl[3 * next_id], tl[3 * next_id + 1], tl[3 * next_id + 2]};
  auto in_next = [&](int v) { return v == nx[0] || v == nx[1] || v == nx[2]; };
  // set semantics: distinct root vertices shared with the next tri
  int sh[3], nsh = 0, third = -1;
  for (int j = 0; j < 3; ++j) {
    bool dup = false;
    for (int i = 0; i < j; ++i) dup |= (r[i] == r[j]);
    if (dup) continue;
    if (in_next(r[j])) sh[nsh++] = r[j];
    else third = r[j];
  }
  if (nsh != 2 || third < 0) return;
  out[0] = third;
  out[1] = std::min(sh[0], sh[1]);
  out[2] = std::max(sh[0], sh[1]);
}

// Strip step rule shared by the walk and the writer: returns edge code 0/1,
// fills (a,b,c,new_v) and the next prev triangle; -1 when the triangle does
// not share an edge containing the newest vertex.
static int strip_step(const int prev[3], const uint8_t* tri, int& a, int& b, int& c,
                      int& nv, int np[3]) {
  int t[3] = {tri[0], tri[1], tri[2]};
  auto in_prev = [&](int v) { return v == prev[0] || v == prev[1] || v == prev[2]; };
  int sh[3], nsh = 0, fresh = -1, nfresh = 0;
  for (int j = 0; j < 3; ++j) {
    bool dup = false;
    for (int i = 0; i < j; ++i) dup |= (t[i] == t[j]);
    if (dup) continue;
    if (in_prev(t[j])) sh[nsh++] = t[j];
    else { fresh = t[j]; nfresh++; }
  }
  if (nsh != 2 || nfresh != 1) return -1;
  auto same = [](int x0, int x1, int y0, int y1) {
    return (x0 == y0 && x1 == y1) || (x0 == y1 && x1 == y0);
  };
  nv = fresh;
  if (same(sh[0], sh[1], prev[1], prev[2])) {
    a = prev[1]; b = prev[2]; c = prev[0];
    np[0] = prev[1]; np[1] = prev[2]; np[2] = nv;
    return 0;
  }
  if (same(sh[0], sh[1], prev[0], prev[2])) {
    a = prev[0]; b = prev[2]; c = prev[1];
    np[0] = prev[0]; np[1] = prev[2]; np[2] = nv;
    return 1;
  }
  return -1;
}

static bool walk_plan(const Plan& pl, std::vector<WalkEntry>& walk, std::string& err) {
  const int n_local = (int)pl.l2g.size();
  std::vector<uint8_t> dec(n_local, 0);
  walk.clear();
  const uint8_t* tl = pl.tris_local.data();
  size_t pos = 0;
  for (size_t s = 0; s < pl.strip_len.size(); ++s) {
    int L = pl.strip_len[s];
    const int32_t* st = pl.strip_tris.data() + pos;
    pos += L;
    int root[3];
    root_orient(tl, st[0], L > 1 ? st[1] : -1, root);
    for (int j = 0; j < 3; ++j)
      if (!dec[root[j]]) {
        walk.push_back({(uint8_t)root[j], 0, 0, 0, 0});
        dec[root[j]] = 1;
      }
    int prev[3] = {root[0], root[1], root[2]};
    for (int i = 1; i < L; ++i) {
      int a, b, c, nv, np[3];
      if (strip_step(prev, tl + 3 * st[i], a, b, c, nv, np) < 0) {
        err = "oldest-edge share";
        return false;
      }
      if (!dec[nv]) {
        walk.push_back({(uint8_t)nv, 1, (uint8_t)a, (uint8_t)b, (uint8_t)c});
        dec[nv] = 1;
      }
      prev[0] = np[0]; prev[1] = np[1]; prev[2] = np[2];
    }
  }
  if ((int)walk.size() != n_local) {
    err = "walk does not reach every local vertex";
    return false;
  }
  return true;
}

static uint64_t spread3_21(uint64_t v) {
  v &= 0x1FFFFF;
  v = (v | (v << 32)) & 0x001F00000000FFFFULL;
  v = (v | (v << 16)) & 0x001F0000FF0000FFULL;
  v = (v | (v << 8)) & 0x100F00F00F00F00FULL;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

bool build_prep(const int64_t* tris, int64_t n_t, const double* vn, int64_t n_v,
                const double* fn, const int64_t* codes, const PrepParams& P, Prep& out) {
  out = Prep();
  out.n_v = n_v;
  out.n_t = n_t;
  if (n_t == 0 || n_v == 0) {
    out.error = "empty mesh";
    return false;
  }
  if (P.max_verts > 256 || P.max_verts < 3 || P.max_tris < 1 || P.max_tris > 65535) {
    out.error = "max_verts must be in [3, 256] and max_tris in [1, 65535]";
    return false;
  }
  for (int64_t i = 0; i < 3 * n_t; ++i)
    if (tris[i] < 0 || tris[i] >= n_v) {
      out.error = "triangle index out of range";
      return false;
    }
  std::vector<int32_t> aoff, aidx, eid;
  global_adjacency(tris, n_t, aoff, aidx, eid);

  Partitioner part;
  part.T = tris; part.Vp = vn; part.FN = fn;
  part.aoff = aoff.data(); part.aidx = aidx.data(); part.eid = eid.data();
  part.n_t = n_t; part.n_v = n_v;
  part.max_tris = P.max_tris; part.max_verts = P.max_verts;
  for (int i = 0; i < 6; ++i) { part.K[i] = P.K[i]; part.NM[i] = P.NM[i]; }
  std::vector<std::vector<int32_t>> meshlets;
  part.run(meshlets);
  part.visited.clear(); part.vert_in.clear(); part.tri_in.clear();
  part.bnd_edge_in.clear(); part.in_frontier.clear(); part.depth.clear();

  // Boundary vertices: used by >= 2 meshlets (utils/boundary_split.py:46-56).
  const int64_t n_m = (int64_t)meshlets.size();
  std::vector<int32_t> last(n_v, -1);
  std::vector<uint8_t> cnt(n_v, 0);
  for (int64_t m = 0; m < n_m; ++m)
    for (int32_t t : meshlets[m])
      for (int j = 0; j < 3; ++j) {
        int64_t v = tris[3 * (int64_t)t + j];
        if (last[v] != m) {
          last[v] = (int32_t)m;
          if (cnt[v] < 2) cnt[v]++;
        }
      }
  // Morton rank of boundary vertices: sorted ids, then stable sort by
  // Morton code (utils/boundary_bvh.py:41-52).
  std::vector<int64_t> bnd;
  for (int64_t v = 0; v < n_v; ++v)
    if (cnt[v] >= 2) bnd.push_back(v);
  std::vector<uint64_t> mk(bnd.size());
  for (size_t i = 0; i < bnd.size(); ++i) {
    const int64_t* c = codes + 3 * bnd[i];
    mk[i] = spread3_21((uint64_t)c[0]) | (spread3_21((uint64_t)c[1]) << 1) |
            (spread3_21((uint64_t)c[2]) << 2);
  }
  std::vector<int64_t> bord(bnd.size());
  for (size_t i = 0; i < bnd.size(); ++i) bord[i] = (int64_t)i;
  std::stable_sort(bord.begin(), bord.end(),
                   [&](int64_t x, int64_t y) { return mk[x] < mk[y]; });
  std::vector<int32_t> ref(n_v, -1);
  for (size_t i = 0; i < bord.size(); ++i) ref[bnd[bord[i]]] = (int32_t)i;
  out.boundary.resize(bord.size());
  for (size_t i = 0; i < bord.size(); ++i) out.boundary[i] = bnd[bord[i]];
  last.clear(); last.shrink_to_fit();

  out.plans.resize(n_m);
  out.walks.resize(n_m);
  std::vector<std::string> errs(std::max(1, P.n_threads));
  std::atomic<bool> failed{false};
  std::vector<PlanWork> work(std::max(1, P.n_threads));
  parallel_for(n_m, P.n_threads, [&](int64_t m, int tid) {
    if (failed.load()) return;
    try {
      const auto& mt = meshlets[m];
      Plan& pl = out.plans[m];
      PlanWork& W = work[tid];
      std::vector<int64_t> vorder;
      eb_vertex_order(mt, tris, aoff.data(), aidx.data(), W, vorder);
      std::vector<int64_t> bl, il;
      for (int64_t v : vorder) (cnt[v] >= 2 ? bl : il).push_back(v);
      std::sort(bl.begin(), bl.end(), [&](int64_t x, int64_t y) { return ref[x] < ref[y]; });
      greedy_nn(il, vn);
      pl.n_bnd = (int)bl.size();
      pl.l2g = bl;
      pl.l2g.insert(pl.l2g.end(), il.begin(), il.end());
      const int n_local = (int)pl.l2g.size();
      if (n_local > 256) throw std::runtime_error("meshlet has more than 256 vertices");
      W.g2l.reset(n_local);
      for (int i = 0; i < n_local; ++i) W.g2l.insert(pl.l2g[i], i);
      const int nt = (int)mt.size();
      pl.tris_local.resize(3 * nt);
      for (int t = 0; t < nt; ++t)
        for (int j = 0; j < 3; ++j)
          pl.tris_local[3 * t + j] = (uint8_t)W.g2l.get(tris[3 * (int64_t)mt[t] + j], 0);
      strips_multiseed(pl.tris_local, nt, P, pl);
      std::string e;
      if (!walk_plan(pl, out.walks[m], e)) throw std::runtime_error(e);
    } catch (const std::exception& ex) {
      errs[tid] = ex.what();
      failed = true;
    }
  });
  if (failed) {
    for (auto& e : errs)
      if (!e.empty()) out.error = e;
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Residuals (encoder/paradelta_v5_dup.py:343-392) and bit writer.
// ---------------------------------------------------------------------------
static inline int64_t irlp(int64_t a, int64_t b, int64_t c, const int32_t* n, int K) {
  int64_t s = (int64_t)n[0] * a + (int64_t)n[1] * b + (int64_t)n[2] * c;
  if (K == 0) return s;
  return (s + ((int64_t)1 << (K - 1))) >> K;
}

void compute_residuals(const Prep& P, const int64_t* codes, const int32_t* pred_n,
                       const int32_t* pred_K, Residuals& R, int n_threads) {
  const int64_t n_m = (int64_t)P.plans.size();
  R.anchor.assign(3 * n_m, 0);
  R.n_delta.assign(n_m, 0);
  R.n_para.assign(n_m, 0);
  R.off.assign(n_m + 1, 0);
  for (int64_t m = 0; m < n_m; ++m) {
    int nk0 = 0;
    for (const auto& w : P.walks[m]) nk0 += (w.kind == 0);
    R.n_delta[m] = std::max(0, nk0 - 1);
    R.n_para[m] = (int32_t)P.walks[m].size() - nk0;
    R.off[m + 1] = R.off[m] + 3 * (int64_t)(R.n_delta[m] + R.n_para[m]);
  }
  R.vals.assign(R.off[n_m], 0);
  std::vector<int64_t> mx(std::max(1, n_threads), 0);
  parallel_for(n_m, n_threads, [&](int64_t m, int tid) {
    const Plan& pl = P.plans[m];
    const auto& walk = P.walks[m];
    const int nd = R.n_delta[m], np = R.n_para[m];
    int64_t* dv = R.vals.data() + R.off[m];
    int64_t* pv = dv + 3 * nd;
    int64_t prev[3] = {0, 0, 0};
    bool first = true;
    int di = 0, pi = 0;
    int64_t local_max = 0;
    for (const auto& w : walk) {
      const int64_t* tc = codes + 3 * pl.l2g[w.v];
      if (w.kind == 0) {
        if (first) {
          for (int d = 0; d < 3; ++d) R.anchor[3 * m + d] = tc[d];
          first = false;
        } else {
          for (int d = 0; d < 3; ++d) {
            int64_t r = tc[d] - prev[d];
            dv[d * nd + di] = r;
            local_max = std::max(local_max, r < 0 ? -r : r);
          }
          di++;
        }
      } else {
        const int64_t* ta = codes + 3 * pl.l2g[w.a];
        const int64_t* tb = codes + 3 * pl.l2g[w.b];
        const int64_t* tcc = codes + 3 * pl.l2g[w.c];
        for (int d = 0; d < 3; ++d) {
          int64_t r = tc[d] - irlp(ta[d], tb[d], tcc[d], pred_n + 3 * d, pred_K[d]);
          pv[d * np + pi] = r;
          local_max = std::max(local_max, r < 0 ? -r : r);
        }
        pi++;
      }
      for (int d = 0; d < 3; ++d) prev[d] = tc[d];
    }
    mx[tid] = std::max(mx[tid], local_max);
  });
  R.max_abs = 0;
  for (int64_t v : mx) R.max_abs = std::max(R.max_abs, v);
}

// ---------------------------------------------------------------------------
// Per-meshlet features of the learned coder/predictor proposal.  The stream
// statistics come from the residuals R when given (the predictor the
// encoder settled on), else from the canonical parallelogram a + b - c, so
// the first pass does not depend on the predictor being chosen.  Sums are
// exact integers; the float conversions happen once at the end.
// ---------------------------------------------------------------------------
// Parallelogram samples (global codes of a, b, c and v for every kind-1
// walk entry) of the selected meshlets, axis-major columns.
void para_columns(const Prep& P, const int64_t* codes, const int32_t* sel, int64_t n_sel,
                  int32_t* cols, int64_t* off, int n_threads) {
  const int64_t n_rows = sel ? n_sel : (int64_t)P.plans.size();
  off[0] = 0;
  for (int64_t r = 0; r < n_rows; ++r) {
    const int64_t m = sel ? sel[r] : r;
    int64_t c = 0;
    for (const auto& w : P.walks[m]) c += w.kind != 0;
    off[r + 1] = off[r] + c;
  }
  if (!cols) return;
  const int64_t n = off[n_rows];
  parallel_for(n_rows, n_threads, [&](int64_t r, int) {
    const int64_t m = sel ? sel[r] : r;
    const Plan& pl = P.plans[m];
    int64_t i = off[r];
    for (const auto& w : P.walks[m]) {
      if (!w.kind) continue;
      const int64_t g[4] = {pl.l2g[w.a], pl.l2g[w.b], pl.l2g[w.c], pl.l2g[w.v]};
      for (int q = 0; q < 4; ++q)
        for (int d = 0; d < 3; ++d) cols[(q * 3 + d) * n + i] = (int32_t)codes[3 * g[q] + d];
      ++i;
    }
  });
}

static inline uint64_t zz_host(int64_t r) { return r < 0 ? (uint64_t)(-r) * 2 - 1 : (uint64_t)r * 2; }
static inline int bitlen1(uint64_t u) { return 63 - __builtin_clzll(u + 1); }   // floor(log2(u + 1))

void meshlet_features(const Prep& P, const int64_t* codes, const Residuals* R, const int32_t* sel,
                      int64_t n_sel, float* out, int n_threads) {
  const int64_t n_rows = sel ? n_sel : (int64_t)P.plans.size();
  parallel_for(n_rows, n_threads, [&](int64_t row, int) {
    const int64_t m = sel ? sel[row] : row;
    const Plan& pl = P.plans[m];
    const auto& walk = P.walks[m];
    float* f = out + (size_t)kFeatures * row;
    const int n_local = (int)walk.size();
    const int n_tris = (int)(pl.tris_local.size() / 3);
    const int n_strips = (int)pl.strip_len.size();
    int n_kind0 = 0;
    int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
    for (const auto& w : walk) {
      n_kind0 += w.kind == 0;
      const int64_t* tc = codes + 3 * pl.l2g[w.v];
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], tc[d]);
        hi[d] = std::max(hi[d], tc[d]);
      }
    }
    int64_t hist[4] = {0, 0, 0, 0};
    for (int L : pl.strip_len) hist[L <= 1 ? 0 : (L <= 3 ? 1 : (L <= 15 ? 2 : 3))] += L;
    // streams dx dy dz px py pz: count, sum of zigzag values, sum of bit lengths
    int64_t cnt[6] = {0, 0, 0, 0, 0, 0}, sbl[6] = {0, 0, 0, 0, 0, 0};
    uint64_t su[6] = {0, 0, 0, 0, 0, 0};
    int64_t sxy[3] = {0, 0, 0}, sxx[3] = {0, 0, 0};
    auto add = [&](int s, int64_t r) {
      const uint64_t u = zz_host(r);
      su[s] += u;
      sbl[s] += bitlen1(u);
      cnt[s]++;
    };
    int64_t prev[3] = {0, 0, 0};
    bool first = true;
    for (const auto& w : walk) {
      const int64_t* tc = codes + 3 * pl.l2g[w.v];
      if (w.kind == 0) {
        if (!first && !R)
          for (int d = 0; d < 3; ++d) add(d, tc[d] - prev[d]);
        first = false;
      } else {
        const int64_t* ta = codes + 3 * pl.l2g[w.a];
        const int64_t* tb = codes + 3 * pl.l2g[w.b];
        const int64_t* tq = codes + 3 * pl.l2g[w.c];
        for (int d = 0; d < 3; ++d) {
          if (!R) add(3 + d, tc[d] - (ta[d] + tb[d] - tq[d]));
          const int64_t x = ta[d] + tb[d] - 2 * tq[d], y = tc[d] - tq[d];
          sxy[d] += x * y;
          sxx[d] += x * x;
        }
      }
      for (int d = 0; d < 3; ++d) prev[d] = tc[d];
    }
    if (R) {
      const int nd = R->n_delta[m], np = R->n_para[m];
      const int64_t* v = R->vals.data() + R->off[m];
      for (int d = 0; d < 3; ++d) {
        for (int i = 0; i < nd; ++i) add(d, v[d * nd + i]);
        for (int i = 0; i < np; ++i) add(3 + d, v[3 * nd + d * np + i]);
      }
    }
    f[0] = (float)(n_local / 256.0);
    f[1] = (float)(n_tris / 256.0);
    f[2] = (float)(n_tris ? (double)n_strips / n_tris : 0.0);
    f[3] = (float)(n_local ? (double)n_kind0 / n_local : 0.0);
    for (int d = 0; d < 3; ++d)
      f[4 + d] = (float)(n_local ? std::log2(1.0 + (double)(hi[d] - lo[d])) / 16.0 : 0.0);
    for (int b = 0; b < 4; ++b) f[7 + b] = (float)(n_tris ? (double)hist[b] / n_tris : 0.0);
    for (int s = 0; s < 6; ++s) {
      f[11 + s] = (float)(std::log2(1.0 + (double)cnt[s]) / 8.0);
      f[17 + s] = (float)(cnt[s] ? (double)sbl[s] / (double)cnt[s] / 16.0 : 0.0);
      f[23 + s] = (float)(cnt[s] ? std::log2(1.0 + (double)su[s] / (double)cnt[s]) / 16.0 : 0.0);
    }
    for (int d = 0; d < 3; ++d) {
      const double s = sxx[d] > 0 ? (double)sxy[d] / (double)sxx[d] : 1.0;
      f[29 + d] = (float)(std::min(2.0, std::max(0.0, s)) / 2.0);
    }
  });
}

void BitWriter::put(uint64_t value, int n) {
  if (n <= 0) return;
  if (n < 64) value &= ((uint64_t)1 << n) - 1;
  nbits_ += (uint64_t)n;
  // feed in chunks of <= 56 bits so acc never overflows
  while (n > 0) {
    int take = n > 56 ? 56 : n;
    uint64_t part = (n - take) >= 64 ? 0 : (value >> (n - take));
    if (take < 64) part &= ((uint64_t)1 << take) - 1;
    acc_ = (acc_ << take) | part;
    nacc_ += take;
    n -= take;
    while (nacc_ >= 8) {
      buf_.push_back((uint8_t)(acc_ >> (nacc_ - 8)));
      nacc_ -= 8;
    }
    acc_ &= ((uint64_t)1 << nacc_) - 1;
  }
}

void BitWriter::put_rice(uint64_t u, int k) {
  uint64_t q = u >> k;
  while (q >= 32) { put(0, 32); q -= 32; }
  put(1, (int)q + 1);
  if (k > 0) put(u, k);
}

void BitWriter::put_eg(uint64_t u, int k) {
  uint64_t shifted = u >> k;
  int lb = 0;
  for (uint64_t x = shifted + 1; x > 1; x >>= 1) ++lb;
  if (lb > 0) put(0, lb);
  put(1, 1);
  if (lb > 0) put(shifted - ((((uint64_t)1) << lb) - 1), lb);
  if (k > 0) put(u, k);
}

std::vector<uint8_t> BitWriter::finish() {
  if (nacc_ > 0) {
    buf_.push_back((uint8_t)(acc_ << (8 - nacc_)));
    acc_ = 0;
    nacc_ = 0;
  }
  return buf_;
}

static inline uint64_t zigzag(int64_t v) {
  return v < 0 ? ((uint64_t)(-v) << 1) - 1 : ((uint64_t)v << 1);
}

static void emit_vert(BitWriter& w, int v, int* fifo, int& fn, int idx_bits) {
  int pos = -1;
  for (int i = 0; i < fn; ++i)
    if (fifo[i] == v) { pos = i; break; }
  if (pos >= 0) {
    w.put(0, 1);
    w.put((uint64_t)pos, kReuseBits);
    for (int i = pos; i + 1 < fn; ++i) fifo[i] = fifo[i + 1];
    fn--;
  } else {
    w.put(1, 1);
    w.put((uint64_t)v, idx_bits);
  }
  if (fn == kFifoSize) {
    for (int i = 0; i + 1 < kFifoSize; ++i) fifo[i] = fifo[i + 1];
    fn--;
  }
  fifo[fn++] = v;
}

std::vector<uint8_t> write_dup(const Prep& P, const Residuals& R, const Header& H,
                               const uint8_t* tagk) {
  const int64_t n_m = (int64_t)P.plans.size();
  BitWriter body;
  std::vector<uint64_t> rel(n_m);
  uint64_t total_local = 0;
  for (int64_t m = 0; m < n_m; ++m) {
    const Plan& pl = P.plans[m];
    rel[m] = body.bit_pos();
    const int n_local = (int)pl.l2g.size();
    const int n_tris = (int)pl.tris_local.size() / 3;
    total_local += n_local;
    body.put(n_local, 16);
    body.put(n_tris, 16);
    body.put(pl.strip_len.size(), 16);
    const int ib = idx_bits_for(n_local);
    int fifo[kFifoSize];
    int fn = 0;
    const uint8_t* tl = pl.tris_local.data();
    size_t pos = 0;
    for (size_t s = 0; s < pl.strip_len.size(); ++s) {
      int L = pl.strip_len[s];
      const int32_t* st = pl.strip_tris.data() + pos;
      pos += L;
      body.put(L, 16);
      int root[3];
      root_orient(tl, st[0], L > 1 ? st[1] : -1, root);
      for (int j = 0; j < 3; ++j) emit_vert(body, root[j], fifo, fn, ib);
      int prev[3] = {root[0], root[1], root[2]};
      for (int i = 1; i < L; ++i) {
        int a, b, c, nv, np[3];
        int ec = strip_step(prev, tl + 3 * st[i], a, b, c, nv, np);
        body.put((uint64_t)ec, 1);
        emit_vert(body, nv, fifo, fn, ib);
        prev[0] = np[0]; prev[1] = np[1]; prev[2] = np[2];
      }
    }
    for (int d = 0; d < 3; ++d) body.put((uint64_t)R.anchor[3 * m + d], H.g_bits[d]);
    const uint8_t* tk = tagk + 12 * m;
    for (int j = 0; j < 12; ++j) body.put(tk[j], 8);
    const int64_t* v = R.vals.data() + R.off[m];
    const int nd = R.n_delta[m], np = R.n_para[m];
    for (int s = 0; s < 6; ++s) {
      const int cntv = s < 3 ? nd : np;
      const int64_t* sv = s < 3 ? v + s * nd : v + 3 * nd + (s - 3) * np;
      const int tag = tk[2 * s], k = tk[2 * s + 1];
      for (int i = 0; i < cntv; ++i) {
        if (tag == 1) body.put_rice(zigzag(sv[i]), k);
        else body.put_eg(zigzag(sv[i]), k);
      }
    }
  }
  std::vector<uint8_t> body_bytes = body.finish();

  BitWriter w;
  w.put(kMagicDup, 32);
  w.put(kVersionDup, 8);
  w.put(R.max_abs > 32767 ? 1 : 0, 8);
  auto f32 = [&](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    w.put(u, 32);
  };
  for (int d = 0; d < 3; ++d) f32(H.center[d]);
  f32(H.scale);
  f32(H.per_coord_err);
  for (int d = 0; d < 3; ++d) f32(H.g_min[d]);
  for (int d = 0; d < 3; ++d) f32(H.g_range[d]);
  for (int d = 0; d < 3; ++d) w.put((uint64_t)H.g_bits[d], 8);
  for (int i = 0; i < 9; ++i) w.put((uint64_t)(H.pred_n[i] & 0xFFFF), 16);
  for (int d = 0; d < 3; ++d) w.put((uint64_t)(H.pred_K[d] & 0xFF), 8);
  w.put(total_local, 32);
  w.put((uint64_t)P.n_t, 32);
  w.put((uint64_t)n_m, 32);
  const uint64_t region = w.bit_pos() + 32ULL * (uint64_t)n_m;
  for (int64_t m = 0; m < n_m; ++m) w.put(region + rel[m], 32);
  std::vector<uint8_t> out = w.finish();
  out.insert(out.end(), body_bytes.begin(), body_bytes.end());
  return out;
}

// ---------------------------------------------------------------------------
// PDLT v5 writer (encoder/paradelta_v5.py:144-283; the per-meshlet chain and
// bit layout of encoder/paradelta_v5_nb.py:20-539).
// ---------------------------------------------------------------------------
namespace {

// g_min + code / (2^bits - 1) * g_range in f64 (paradelta_codec.py:167-175).
inline double dequant64(int64_t code, const V5Header& H, int d) {
  const int64_t mx = ((int64_t)1 << H.g_bits[d]) - 1;
  if (mx == 0) return (double)H.g_min[d];
  return (double)H.g_min[d] + (double)code / (double)mx * (double)H.g_range[d];
}

// The float32 closed-loop chain of one meshlet (interior_pass_strip_nb).
// w3 == nullptr: plain parallelogram a + b - c, fit samples appended to
// rows; otherwise LIN3 with the stored weights, codes written (n_int x 3).
void v5_chain(const Prep& P, int64_t m, const int64_t* gcodes, const double* vn,
              const V5Header& H, const float* w3, int64_t* codes, std::vector<double>* rows) {
  const Plan& pl = P.plans[m];
  const int n_bnd = pl.n_bnd;
  const int n_local = (int)pl.l2g.size();
  float recon[256][3];
  for (int l = 0; l < n_bnd; ++l)
    for (int d = 0; d < 3; ++d) recon[l][d] = (float)dequant64(gcodes[3 * pl.l2g[l] + d], H, d);
  float fb[3] = {0.f, 0.f, 0.f};
  if (n_bnd > 0) {
    float s[3] = {0.f, 0.f, 0.f};
    for (int l = 0; l < n_bnd; ++l)
      for (int d = 0; d < 3; ++d) s[d] = s[d] + recon[l][d];
    const float nb = (float)n_bnd;
    for (int d = 0; d < 3; ++d) fb[d] = s[d] / nb;
  }
  const float delta = (float)(2.0 * (double)H.per_coord_err);
  int i = 0;
  for (const WalkEntry& e : P.walks[m]) {
    const int v = e.v;
    if (v < n_bnd) continue;
    const int64_t g = pl.l2g[v];
    float pred[3];
    for (int d = 0; d < 3; ++d) {
      if (e.kind == 0) {
        pred[d] = fb[d];
      } else if (!w3) {
        pred[d] = (recon[e.a][d] + recon[e.b][d]) - recon[e.c][d];
      } else {
        const float t0 = w3[0] * recon[e.a][d];
        const float t1 = w3[1] * recon[e.b][d];
        const float t2 = w3[2] * recon[e.c][d];
        pred[d] = (t0 + t1) + t2;
      }
    }
    for (int d = 0; d < 3; ++d) {
      const float tv = (float)vn[3 * g + d];
      const float q = (tv - pred[d]) / delta;
      const int64_t c = (int64_t)std::nearbyint(q);   // half to even, as np.round
      if (codes) codes[3 * i + d] = c;
      recon[v][d] = pred[d] + (float)c * delta;
    }
    if (rows && e.kind == 1) {
      for (int d = 0; d < 3; ++d) rows->push_back((double)recon[e.a][d]);
      for (int d = 0; d < 3; ++d) rows->push_back((double)recon[e.b][d]);
      for (int d = 0; d < 3; ++d) rows->push_back((double)recon[e.c][d]);
      for (int d = 0; d < 3; ++d) rows->push_back(vn[3 * g + d]);
    }
    ++i;
  }
  (void)n_local;
}

inline uint64_t zz64(int64_t v) { return v < 0 ? ((uint64_t)(-(v + 1)) << 1) + 1 : ((uint64_t)v << 1); }

// Lowest Rice cost over k = 0..k_max (lowest k on ties).
int best_rice(const uint64_t* u, size_t n, int k_max, int64_t& bits) {
  int best_k = 0;
  bits = INT64_MAX;
  for (int k = 0; k <= k_max; ++k) {
    int64_t b = (int64_t)n * (1 + k);
    for (size_t i = 0; i < n; ++i) b += (int64_t)(u[i] >> k);
    if (b < bits) { bits = b; best_k = k; }
  }
  return best_k;
}

int best_eg(const uint64_t* u, size_t n, int k_max, int64_t& bits) {
  int best_k = 0;
  bits = INT64_MAX;
  for (int k = 0; k <= k_max; ++k) {
    int64_t b = 0;
    for (size_t i = 0; i < n; ++i) b += 2 * (63 - __builtin_clzll((u[i] >> k) + 1)) + 1 + k;
    if (b < bits) { bits = b; best_k = k; }
  }
  return best_k;
}

void v5_meshlet(BitWriter& w, const Plan& pl, const std::vector<int32_t>& ref, const int64_t* codes) {
  const int n_bnd = pl.n_bnd;
  const int n_local = (int)pl.l2g.size();
  const int n_int = n_local - n_bnd;
  const int n_tris = (int)pl.tris_local.size() / 3;
  w.put((uint64_t)n_bnd, 16);
  w.put((uint64_t)n_int, 16);
  w.put((uint64_t)n_tris, 16);
  w.put(pl.strip_len.size(), 16);
  if (n_bnd > 0) {
    w.put((uint64_t)ref[pl.l2g[0]], 32);
    if (n_bnd > 1) {
      std::vector<uint64_t> diffs(n_bnd - 1);
      for (int i = 0; i + 1 < n_bnd; ++i)
        diffs[i] = (uint64_t)(ref[pl.l2g[i + 1]] - ref[pl.l2g[i]] - 1);
      int64_t bits;
      const int k = best_rice(diffs.data(), diffs.size(), 11, bits);
      w.put((uint64_t)k, 8);
      for (uint64_t x : diffs) w.put_rice(x, k);
    }
  }
  const int ib = idx_bits_for(n_local);
  int fifo[kFifoSize];
  int fn = 0;
  const uint8_t* tl = pl.tris_local.data();
  size_t pos = 0;
  for (size_t s = 0; s < pl.strip_len.size(); ++s) {
    const int L = pl.strip_len[s];
    const int32_t* st = pl.strip_tris.data() + pos;
    pos += L;
    w.put((uint64_t)L, 16);
    int root[3];
    root_orient(tl, st[0], L > 1 ? st[1] : -1, root);
    for (int j = 0; j < 3; ++j) emit_vert(w, root[j], fifo, fn, ib);
    int prev[3] = {root[0], root[1], root[2]};
    for (int i = 1; i < L; ++i) {
      int a, b, c, nv, np[3];
      const int ec = strip_step(prev, tl + 3 * st[i], a, b, c, nv, np);
      w.put((uint64_t)ec, 1);
      emit_vert(w, nv, fifo, fn, ib);
      prev[0] = np[0]; prev[1] = np[1]; prev[2] = np[2];
    }
  }
  if (n_int == 0) return;
  std::vector<uint64_t> u(n_int);
  for (int d = 0; d < 3; ++d) {
    int64_t mn = codes[d], mx = codes[d];
    for (int i = 0; i < n_int; ++i) {
      mn = std::min(mn, codes[3 * i + d]);
      mx = std::max(mx, codes[3 * i + d]);
      u[i] = zz64(codes[3 * i + d]);
    }
    const int64_t rng = mx - mn;
    const int bw = rng > 0 ? 64 - __builtin_clzll((uint64_t)rng + 1) : 1;
    int64_t best = (mn < -32768 || mn > 32767) ? INT64_MAX : 8 + 16 + 8 + (int64_t)n_int * bw;
    int64_t rb, eb;
    const int rk = best_rice(u.data(), u.size(), 11, rb);
    const int ek = best_eg(u.data(), u.size(), 7, eb);
    int tag = 0;
    if (16 + rb < best) { best = 16 + rb; tag = 1; }
    if (16 + eb < best) { best = 16 + eb; tag = 2; }
    w.put((uint64_t)tag, 8);
    if (tag == 0) {
      w.put((uint64_t)mn & 0xFFFF, 16);
      w.put((uint64_t)bw, 8);
      for (int i = 0; i < n_int; ++i) w.put((uint64_t)(codes[3 * i + d] - mn), bw);
    } else {
      const int k = tag == 1 ? rk : ek;
      w.put((uint64_t)k, 8);
      for (int i = 0; i < n_int; ++i) {
        if (tag == 1) w.put_rice(u[i], k);
        else w.put_eg(u[i], k);
      }
    }
  }
}

}  // namespace

void v5_samples(const Prep& P, const int64_t* codes, const double* vn, const V5Header& H,
                std::vector<double>& rows, int n_threads) {
  const int64_t n_m = (int64_t)P.plans.size();
  std::vector<std::vector<double>> per(n_m);
  parallel_for(n_m, n_threads, [&](int64_t m, int) { v5_chain(P, m, codes, vn, H, nullptr, nullptr, &per[m]); });
  size_t n = 0;
  for (auto& r : per) n += r.size();
  rows.clear();
  rows.reserve(n);
  for (auto& r : per) rows.insert(rows.end(), r.begin(), r.end());
}

std::vector<uint8_t> write_v5(const Prep& P, const int64_t* codes, const double* vn,
                              const V5Header& H, const float* lin3, int n_threads) {
  const int64_t n_m = (int64_t)P.plans.size();
  std::vector<int64_t> coff(n_m + 1, 0);
  for (int64_t m = 0; m < n_m; ++m)
    coff[m + 1] = coff[m] + 3 * (int64_t)(P.plans[m].l2g.size() - P.plans[m].n_bnd);
  std::vector<int64_t> rc(coff[n_m]);
  parallel_for(n_m, n_threads, [&](int64_t m, int) {
    v5_chain(P, m, codes, vn, H, lin3, rc.data() + coff[m], nullptr);
  });
  int64_t max_abs = 0;
  for (int64_t c : rc) max_abs = std::max(max_abs, c < 0 ? -c : c);
  std::vector<int32_t> ref(P.n_v, -1);
  for (size_t i = 0; i < P.boundary.size(); ++i) ref[P.boundary[i]] = (int32_t)i;
  BitWriter body;
  std::vector<uint64_t> rel(n_m);
  for (int64_t m = 0; m < n_m; ++m) {
    rel[m] = body.bit_pos();
    v5_meshlet(body, P.plans[m], ref, rc.data() + coff[m]);
  }
  std::vector<uint8_t> body_bytes = body.finish();

  BitWriter w;
  w.put(0x504C4454u, 32);
  w.put(5, 8);
  w.put(max_abs > 32767 ? 1 : 0, 8);
  auto f32 = [&](float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    w.put(x, 32);
  };
  for (int d = 0; d < 3; ++d) f32(H.center[d]);
  f32(H.scale);
  f32(H.per_coord_err);
  for (int d = 0; d < 3; ++d) f32(H.g_min[d]);
  for (int d = 0; d < 3; ++d) f32(H.g_range[d]);
  for (int d = 0; d < 3; ++d) w.put((uint64_t)H.g_bits[d], 8);
  const uint64_t n_b = P.boundary.size();
  w.put((uint64_t)P.n_v, 32);
  w.put((uint64_t)P.n_t, 32);
  w.put(n_b, 32);
  w.put((uint64_t)n_m, 32);
  for (int j = 0; j < 3; ++j) f32(lin3[j]);
  if (n_b > 0) {
    std::vector<uint64_t> u(n_b - 1);
    for (int d = 0; d < 3; ++d) {
      w.put((uint64_t)codes[3 * P.boundary[0] + d], H.g_bits[d]);
      if (n_b > 1) {
        for (uint64_t i = 0; i + 1 < n_b; ++i)
          u[i] = zz64(codes[3 * P.boundary[i + 1] + d] - codes[3 * P.boundary[i] + d]);
        int64_t bits;
        const int k = best_rice(u.data(), u.size(), 11, bits);
        w.put((uint64_t)k, 8);
        for (uint64_t x : u) w.put_rice(x, k);
      }
    }
  }
  const int pad = (int)((8 - (w.bit_pos() & 7)) & 7);
  if (pad) w.put(0, pad);
  const uint64_t region = w.bit_pos() + 32ULL * (uint64_t)n_m;
  for (int64_t m = 0; m < n_m; ++m) w.put(region + rel[m], 32);
  std::vector<uint8_t> out = w.finish();
  out.insert(out.end(), body_bytes.begin(), body_bytes.end());
  return out;
}

}  // namespace stride
