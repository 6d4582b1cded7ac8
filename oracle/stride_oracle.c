/*
 * stride_oracle.c - plain C restatement of the reference's CPU decoder and
 * encoder cost model for the STRIDE (DPR4 v3) path.
 *
 * TEST INFRASTRUCTURE ONLY: linked by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline as the checker / CPU baseline; never by the product
 * library.  Built by oracle/Makefile into oracle/liboracle.so.
 *
 * Follows (maletsden/meshpress):
 *   oracle_decode       encoder/paradelta_v5_dup.py:471-618 (decode_dup), with
 *                       utils/bit_codec.py:192-257 (BitReader) and
 *                       encoder/paradelta_codec.py:330-340 (_read_vert)
 *   oracle_stream_costs utils/residual_entropy.py:39-54 (_rice_bits,
 *                       _exp_golomb_bits) over Rice k 0..11, EG k 0..7
 *   oracle_irlp_cost    encoder/_irlp_fit.py:120-183 (_cost_seeds_njit)
 *   oracle_pick         encoder/paradelta_v5_dup.py:159-171 (_pick_best_k)
 *   oracle_irlp_cost_mt, oracle_pick_streams: the same over host threads
 *
 * n_threads == 1 reads the stream sequentially as decode_dup does; with more
 * threads, meshlets are decoded independently from the offset table.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_OK 0
#define ORACLE_E_MAGIC -10
#define ORACLE_E_VERSION -11
#define ORACLE_E_TRUNC -12
#define ORACLE_E_CORRUPT -13

typedef struct {
  const uint8_t* d;
  uint64_t nbits;
  uint64_t pos;
  int err;
} rd_t;

static inline uint64_t rd_get(rd_t* r, int n) {
  uint64_t v = 0;
  if (n <= 0) return 0;
  if (r->pos + (uint64_t)n > r->nbits) {
    r->err = 1;
    r->pos += (uint64_t)n;
    return 0;
  }
  while (n > 0) {
    uint64_t byte = r->pos >> 3;
    int in = (int)(r->pos & 7);
    int take = 8 - in < n ? 8 - in : n;
    uint32_t chunk = ((uint32_t)r->d[byte] >> (8 - in - take)) & ((1u << take) - 1u);
    v = (v << take) | chunk;
    r->pos += (uint64_t)take;
    n -= take;
  }
  return v;
}

static inline uint64_t rd_unary(rd_t* r) {
  uint64_t q = 0;
  for (;;) {
    if (r->pos >= r->nbits) {
      r->err = 1;
      return q;
    }
    if (rd_get(r, 1)) return q;
    q++;
  }
}

static inline int64_t unzz(uint64_t u) { return (int64_t)(u >> 1) ^ -(int64_t)(u & 1); }

typedef struct {
  uint32_t magic, version, code_width;
  float center[3], scale, pce, g_min[3], g_range[3];
  int g_bits[3];
  int64_t pred_n[9], pred_K[3];
  uint32_t n_v, n_t, n_m;
  uint64_t body_bit;
} hdr_t;

static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static int read_hdr(const uint8_t* d, size_t len, hdr_t* h) {
  rd_t r = {d, 8ull * len, 0, 0};
  h->magic = (uint32_t)rd_get(&r, 32);
  if (r.err) return ORACLE_E_TRUNC;
  if (h->magic != 0x44505234u) return ORACLE_E_MAGIC;
  h->version = (uint32_t)rd_get(&r, 8);
  if (h->version != 3) return ORACLE_E_VERSION;
  h->code_width = (uint32_t)rd_get(&r, 8);
  for (int i = 0; i < 3; ++i) h->center[i] = u2f((uint32_t)rd_get(&r, 32));
  h->scale = u2f((uint32_t)rd_get(&r, 32));
  h->pce = u2f((uint32_t)rd_get(&r, 32));
  for (int i = 0; i < 3; ++i) h->g_min[i] = u2f((uint32_t)rd_get(&r, 32));
  for (int i = 0; i < 3; ++i) h->g_range[i] = u2f((uint32_t)rd_get(&r, 32));
  for (int i = 0; i < 3; ++i) h->g_bits[i] = (int)rd_get(&r, 8);
  for (int i = 0; i < 9; ++i) {
    uint32_t u = (uint32_t)rd_get(&r, 16);
    h->pred_n[i] = (u & 0x8000u) ? (int64_t)u - 0x10000 : (int64_t)u;
  }
  for (int i = 0; i < 3; ++i) h->pred_K[i] = (int64_t)rd_get(&r, 8);
  h->n_v = (uint32_t)rd_get(&r, 32);
  h->n_t = (uint32_t)rd_get(&r, 32);
  h->n_m = (uint32_t)rd_get(&r, 32);
  r.pos += (uint64_t)(-(int64_t)r.pos & 7);
  r.pos += 32ull * h->n_m;
  if (r.err || r.pos > r.nbits) return ORACLE_E_TRUNC;
  h->body_bit = r.pos;
  return ORACLE_OK;
}

static uint64_t offset_of(const uint8_t* d, uint32_t m) {
  const uint8_t* p = d + 86 + 4 * (size_t)m;
  return ((uint64_t)p[0] << 24) | ((uint64_t)p[1] << 16) | ((uint64_t)p[2] << 8) | p[3];
}

/* FIFO as decode_dup's deque(maxlen=16) of ints, oldest first. */
typedef struct {
  int v[16];
  int n;
} fifo_t;

static void fifo_move_back(fifo_t* f, int v) {
  int p = -1;
  for (int i = 0; i < f->n; ++i)
    if (f->v[i] == v) { p = i; break; }
  if (p >= 0) {
    for (int i = p; i + 1 < f->n; ++i) f->v[i] = f->v[i + 1];
    f->n--;
  } else if (f->n == 16) {
    for (int i = 0; i + 1 < 16; ++i) f->v[i] = f->v[i + 1];
    f->n--;
  }
  f->v[f->n++] = v;
}

static int read_vert(rd_t* r, fifo_t* f, int ib, int* v) {
  if (rd_get(r, 1) == 0) {
    int fi = (int)rd_get(r, 5);
    if (fi >= f->n) return ORACLE_E_CORRUPT; /* list(deque)[fi] -> IndexError */
    *v = f->v[fi];
  } else {
    *v = (int)rd_get(r, ib);
  }
  fifo_move_back(f, *v);
  return r->err ? ORACLE_E_TRUNC : ORACLE_OK;
}

static int64_t irlp(int64_t a, int64_t b, int64_t c, const int64_t* n, int64_t K) {
  int64_t s = n[0] * a + n[1] * b + n[2] * c;
  return K == 0 ? s : (s + ((int64_t)1 << (K - 1))) >> K;
}

/* Decodes one meshlet at r->pos. codes: n_local x 3; tris: n_tris x 3 local
 * ids.  Returns n_local / n_tris through out params. */
static int decode_meshlet(rd_t* r, const hdr_t* h, int64_t* codes, int64_t* tris,
                          int* nl_out, int* nt_out, int cap_local, int cap_tris) {
  int nl = (int)rd_get(r, 16), nt = (int)rd_get(r, 16), ns = (int)rd_get(r, 16);
  if (r->err) return ORACLE_E_TRUNC;
  if (nl > cap_local || nt > cap_tris) return ORACLE_E_CORRUPT;
  int ib = 1;
  while ((1 << ib) < nl + 1) ++ib;
  fifo_t f;
  f.n = 0;
  /* the walk holds distinct local ids, all below 2^idx_bits */
  const size_t wcap = ((size_t)1 << ib) + 1;
  int* walk_v = (int*)malloc(sizeof(int) * wcap);
  int* walk_r = (int*)malloc(sizeof(int) * 3 * wcap);
  uint8_t* seen = (uint8_t*)calloc(wcap, 1);
  int nw = 0, tc = 0, rc = ORACLE_OK;
  for (int s = 0; s < ns && rc == ORACLE_OK; ++s) {
    int L = (int)rd_get(r, 16);
    int p[3];
    for (int j = 0; j < 3 && rc == ORACLE_OK; ++j) rc = read_vert(r, &f, ib, &p[j]);
    if (rc) break;
    if (tc >= cap_tris) { rc = ORACLE_E_CORRUPT; break; }
    tris[3 * tc] = p[0]; tris[3 * tc + 1] = p[1]; tris[3 * tc + 2] = p[2];
    tc++;
    for (int j = 0; j < 3; ++j)
      if (!seen[p[j]]) { seen[p[j]] = 1; walk_v[nw] = p[j]; walk_r[3 * nw] = -1; nw++; }
    for (int t = 0; t < L - 1; ++t) {
      int ec = (int)rd_get(r, 1), nv;
      rc = read_vert(r, &f, ib, &nv);
      if (rc) break;
      int a, b, c;
      if (ec == 0) { a = p[1]; b = p[2]; c = p[0]; p[0] = p[1]; p[1] = p[2]; p[2] = nv; }
      else { a = p[0]; b = p[2]; c = p[1]; p[1] = p[2]; p[2] = nv; }
      if (tc >= cap_tris) { rc = ORACLE_E_CORRUPT; break; }
      tris[3 * tc] = a; tris[3 * tc + 1] = b; tris[3 * tc + 2] = nv;
      tc++;
      if (!seen[nv]) {
        seen[nv] = 1;
        walk_v[nw] = nv; walk_r[3 * nw] = a; walk_r[3 * nw + 1] = b; walk_r[3 * nw + 2] = c;
        nw++;
      }
    }
  }
  if (rc == ORACLE_OK) {
    int64_t anchor[3];
    for (int d = 0; d < 3; ++d) anchor[d] = (int64_t)rd_get(r, h->g_bits[d]);
    int tag[6], k[6];
    for (int s = 0; s < 6; ++s) { tag[s] = (int)rd_get(r, 8); k[s] = (int)rd_get(r, 8); }
    int nk0 = 0;
    for (int i = 0; i < nw; ++i) nk0 += walk_r[3 * i] < 0;
    int nd = nk0 > 0 ? nk0 - 1 : 0, np = nw - nk0;
    int64_t* res = (int64_t*)malloc(sizeof(int64_t) * (size_t)(3 * (nd + np) + 1));
    for (int s = 0; s < 6; ++s) {
      int cnt = s < 3 ? nd : np;
      int64_t* dst = res + (s < 3 ? s * nd : 3 * nd + (s - 3) * np);
      for (int i = 0; i < cnt; ++i) {
        uint64_t u;
        if (tag[s] == 1) {
          uint64_t q = rd_unary(r);
          u = (q << k[s]) | rd_get(r, k[s]);
        } else {
          uint64_t lb = rd_unary(r);
          uint64_t sh = lb ? ((1ull << lb) - 1 + rd_get(r, (int)lb)) : 0;
          u = (sh << k[s]) | rd_get(r, k[s]);
        }
        dst[i] = unzz(u);
      }
    }
    if (r->err) rc = ORACLE_E_TRUNC;
    if (rc == ORACLE_OK && tc != nt) rc = ORACLE_E_CORRUPT;
    if (rc == ORACLE_OK) {
      for (int i = 0; i < nl; ++i) codes[3 * i] = codes[3 * i + 1] = codes[3 * i + 2] = 0;
      int64_t prev[3] = {0, 0, 0};
      int first = 1, di = 0, pi = 0;
      for (int i = 0; i < nw; ++i) {
        int v = walk_v[i];
        if (v >= nl) { rc = ORACLE_E_CORRUPT; break; }
        int64_t* cv = codes + 3 * v;
        if (walk_r[3 * i] < 0) {
          if (first) { cv[0] = anchor[0]; cv[1] = anchor[1]; cv[2] = anchor[2]; first = 0; }
          else {
            for (int d = 0; d < 3; ++d) cv[d] = prev[d] + res[d * nd + di];
            di++;
          }
        } else {
          const int64_t* ca = codes + 3 * walk_r[3 * i];
          const int64_t* cb = codes + 3 * walk_r[3 * i + 1];
          const int64_t* cc = codes + 3 * walk_r[3 * i + 2];
          for (int d = 0; d < 3; ++d)
            cv[d] = irlp(ca[d], cb[d], cc[d], h->pred_n + 3 * d, h->pred_K[d]) + res[3 * nd + d * np + pi];
          pi++;
        }
        prev[0] = cv[0]; prev[1] = cv[1]; prev[2] = cv[2];
      }
    }
    free(res);
  }
  free(walk_v);
  free(walk_r);
  free(seen);
  *nl_out = nl;
  *nt_out = tc;
  return rc;
}

static void dequant(const hdr_t* h, const int64_t* codes, int n, float* V) {
  for (int d = 0; d < 3; ++d) {
    const uint64_t mx = (1ull << h->g_bits[d]) - 1;
    const double gmin = h->g_min[d], grng = h->g_range[d], sc = h->scale, ce = h->center[d];
    for (int i = 0; i < n; ++i) {
      double vn = mx == 0 ? gmin : gmin + ((double)codes[3 * i + d] / (double)mx) * grng;
      V[3 * i + d] = (float)(vn * sc + ce);
    }
  }
}

/* n_v = sum n_local, n_t = sum n_tris, from the meshlet headers. */
int oracle_sizes(const uint8_t* d, size_t len, uint32_t* n_m, uint64_t* n_v, uint64_t* n_t) {
  hdr_t h;
  int rc = read_hdr(d, len, &h);
  if (rc) return rc;
  uint64_t sv = 0, st = 0;
  for (uint32_t m = 0; m < h.n_m; ++m) {
    rd_t r = {d, 8ull * len, offset_of(d, m), 0};
    sv += rd_get(&r, 16);
    st += rd_get(&r, 16);
    if (r.err) return ORACLE_E_TRUNC;
  }
  *n_m = h.n_m;
  *n_v = sv;
  *n_t = st;
  return ORACLE_OK;
}

typedef struct {
  const uint8_t* d;
  size_t len;
  const hdr_t* h;
  const uint64_t* vb;
  const uint64_t* tb;
  float* V;
  int64_t* T;
  int64_t* codes;
  uint32_t m0, m1;
  int rc;
} job_t;

static void* decode_range(void* arg) {
  job_t* j = (job_t*)arg;
  int64_t codes[3 * 65536];
  int64_t* tris = (int64_t*)malloc(sizeof(int64_t) * 3 * 65536);
  for (uint32_t m = j->m0; m < j->m1 && j->rc == 0; ++m) {
    rd_t r = {j->d, 8ull * j->len, offset_of(j->d, m), 0};
    int nl, nt;
    int rc = decode_meshlet(&r, j->h, codes, tris, &nl, &nt, 65535, 65535);
    if (rc) { j->rc = rc; break; }
    uint64_t vb = j->vb[m], tb = j->tb[m];
    dequant(j->h, codes, nl, j->V + 3 * vb);
    if (j->codes) memcpy(j->codes + 3 * vb, codes, sizeof(int64_t) * 3 * (size_t)nl);
    for (int t = 0; t < 3 * nt; ++t) j->T[3 * tb + t] = (int64_t)vb + tris[t];
  }
  free(tris);
  return NULL;
}

/* Decodes to V float32 (n_v x 3), T int64 (n_t x 3) and optionally the
 * integer codes (n_v x 3). */
int oracle_decode(const uint8_t* d, size_t len, float* V, int64_t* T, int64_t* codes_out,
                  int n_threads) {
  hdr_t h;
  int rc = read_hdr(d, len, &h);
  if (rc) return rc;
  if (n_threads <= 1) {
    rd_t r = {d, 8ull * len, h.body_bit, 0};
    int64_t* codes = (int64_t*)malloc(sizeof(int64_t) * 3 * 65536);
    int64_t* tris = (int64_t*)malloc(sizeof(int64_t) * 3 * 65536);
    uint64_t vb = 0, tb = 0;
    for (uint32_t m = 0; m < h.n_m && rc == 0; ++m) {
      int nl, nt;
      rc = decode_meshlet(&r, &h, codes, tris, &nl, &nt, 65535, 65535);
      if (rc) break;
      dequant(&h, codes, nl, V + 3 * vb);
      if (codes_out) memcpy(codes_out + 3 * vb, codes, sizeof(int64_t) * 3 * (size_t)nl);
      for (int t = 0; t < 3 * nt; ++t) T[3 * tb + t] = (int64_t)vb + tris[t];
      vb += (uint64_t)nl;
      tb += (uint64_t)nt;
    }
    free(codes);
    free(tris);
    return rc;
  }
  uint64_t* vb = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)h.n_m + 1));
  uint64_t* tb = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)h.n_m + 1));
  vb[0] = tb[0] = 0;
  for (uint32_t m = 0; m < h.n_m; ++m) {
    rd_t r = {d, 8ull * len, offset_of(d, m), 0};
    vb[m + 1] = vb[m] + rd_get(&r, 16);
    tb[m + 1] = tb[m] + rd_get(&r, 16);
    if (r.err) { free(vb); free(tb); return ORACLE_E_TRUNC; }
  }
  if (n_threads > 256) n_threads = 256;
  pthread_t th[256];
  job_t jobs[256];
  for (int t = 0; t < n_threads; ++t) {
    jobs[t] = (job_t){d, len, &h, vb, tb, V, T, codes_out,
                      (uint32_t)((uint64_t)h.n_m * t / n_threads),
                      (uint32_t)((uint64_t)h.n_m * (t + 1) / n_threads), 0};
  }
  /* per-thread stacks hold 1.5 MiB of codes: create with a larger stack */
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, 8u << 20);
  for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], &attr, decode_range, &jobs[t]);
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
  pthread_attr_destroy(&attr);
  for (int t = 0; t < n_threads; ++t)
    if (jobs[t].rc) { rc = jobs[t].rc; break; }
  free(vb);
  free(tb);
  return rc;
}

/* The 20 body costs of one stream: Rice k = 0..11, then EG k = 0..7. */
void oracle_stream_costs(const int64_t* vals, int64_t n, int64_t* costs) {
  for (int j = 0; j < 20; ++j) costs[j] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t v = vals[i];
    uint64_t u = v < 0 ? ((uint64_t)(-v) << 1) - 1 : ((uint64_t)v << 1);
    for (int k = 0; k < 12; ++k) costs[k] += (int64_t)(u >> k) + 1 + k;
    for (int k = 0; k < 8; ++k) {
      uint64_t x = (u >> k) + 1;
      int lb = 63 - __builtin_clzll(x);
      costs[12 + k] += 2 * lb + 1 + k;
    }
  }
}

/* (tag, k) choice per stream with _pick_best_k's tie-breaks. */
void oracle_pick(const int64_t* vals, int64_t n, int* tag, int* k, int64_t* bits) {
  int64_t c[20];
  if (n == 0) { *tag = 1; *k = 0; *bits = 0; return; }
  oracle_stream_costs(vals, n, c);
  int rk = 0, ek = 0;
  for (int j = 1; j < 12; ++j) if (c[j] < c[rk]) rk = j;
  for (int j = 1; j < 8; ++j) if (c[12 + j] < c[12 + ek]) ek = j;
  if (c[rk] <= c[12 + ek]) { *tag = 1; *k = rk; *bits = c[rk]; }
  else { *tag = 2; *k = ek; *bits = c[12 + ek]; }
}

/* Sum over meshlets of min-coder bits for S candidates (n0, n1, n2) at K. */
void oracle_irlp_cost(const int64_t* a, const int64_t* b, const int64_t* c, const int64_t* v,
                      const int64_t* off, int64_t n_m, const int64_t* cand, int S, int K,
                      int64_t* out) {
  int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * 65536);
  for (int s = 0; s < S; ++s) {
    int64_t total = 0;
    const int64_t* n3 = cand + 3 * s;
    for (int64_t m = 0; m < n_m; ++m) {
      int64_t cnt = off[m + 1] - off[m];
      if (cnt <= 0) continue;
      for (int64_t i = 0; i < cnt && i < 65536; ++i) {
        int64_t j = off[m] + i;
        tmp[i] = v[j] - irlp(a[j], b[j], c[j], n3, K);
      }
      int64_t cs[20];
      oracle_stream_costs(tmp, cnt < 65536 ? cnt : 65536, cs);
      int64_t best = cs[0];
      for (int q = 1; q < 20; ++q) if (cs[q] < best) best = cs[q];
      total += best;
    }
    out[s] = total;
  }
  free(tmp);
}

/* ------------------------------------------------------------------------
 * Threaded forms of the encoder search cost model: the CPU side of the
 * exhaustive-search baseline (bench.py's encode-side cpu_baseline).  Same
 * arithmetic as oracle_irlp_cost and oracle_pick; the work is split over
 * n_threads host threads by meshlet (costs) or by stream (picks). */
typedef struct {
  const int64_t *a, *b, *c, *v, *off, *cand;
  int64_t m0, m1;
  int S, K;
  int64_t* part;   /* S partial sums of this thread */
} irlp_job_t;

static void* irlp_range(void* arg) {
  irlp_job_t* j = (irlp_job_t*)arg;
  int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * 65536);
  for (int s = 0; s < j->S; ++s) {
    const int64_t* n3 = j->cand + 3 * s;
    int64_t total = 0;
    for (int64_t m = j->m0; m < j->m1; ++m) {
      int64_t cnt = j->off[m + 1] - j->off[m];
      if (cnt <= 0) continue;
      if (cnt > 65536) cnt = 65536;
      for (int64_t i = 0; i < cnt; ++i) {
        int64_t q = j->off[m] + i;
        tmp[i] = j->v[q] - irlp(j->a[q], j->b[q], j->c[q], n3, j->K);
      }
      int64_t cs[20];
      oracle_stream_costs(tmp, cnt, cs);
      int64_t best = cs[0];
      for (int t = 1; t < 20; ++t) if (cs[t] < best) best = cs[t];
      total += best;
    }
    j->part[s] = total;
  }
  free(tmp);
  return NULL;
}

/* oracle_irlp_cost over n_threads threads (meshlets split evenly). */
void oracle_irlp_cost_mt(const int64_t* a, const int64_t* b, const int64_t* c, const int64_t* v,
                         const int64_t* off, int64_t n_m, const int64_t* cand, int S, int K,
                         int64_t* out, int n_threads) {
  if (n_threads < 1) n_threads = 1;
  if (n_threads > 256) n_threads = 256;
  pthread_t th[256];
  irlp_job_t jobs[256];
  int64_t* part = (int64_t*)calloc((size_t)n_threads * (size_t)(S > 0 ? S : 1), sizeof(int64_t));
  for (int t = 0; t < n_threads; ++t) {
    jobs[t] = (irlp_job_t){a, b, c, v, off, cand, n_m * t / n_threads, n_m * (t + 1) / n_threads,
                           S, K, part + (size_t)t * (size_t)S};
    pthread_create(&th[t], NULL, irlp_range, &jobs[t]);
  }
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
  for (int s = 0; s < S; ++s) {
    int64_t total = 0;
    for (int t = 0; t < n_threads; ++t) total += part[(size_t)t * (size_t)S + s];
    out[s] = total;
  }
  free(part);
}

typedef struct {
  const int64_t *vals, *off;
  int64_t s0, s1;
  uint8_t* tagk;
  int64_t* bits;
} pick_job_t;

static void* pick_range(void* arg) {
  pick_job_t* j = (pick_job_t*)arg;
  for (int64_t s = j->s0; s < j->s1; ++s) {
    int tag, k;
    int64_t bits;
    oracle_pick(j->vals + j->off[s], j->off[s + 1] - j->off[s], &tag, &k, &bits);
    j->tagk[2 * s] = (uint8_t)tag;
    j->tagk[2 * s + 1] = (uint8_t)k;
    j->bits[s] = bits;
  }
  return NULL;
}

/* oracle_pick for streams [off[s], off[s+1]) of vals, s < n_streams, over
 * n_threads threads: (tag, k) bytes and body bits per stream. */
void oracle_pick_streams(const int64_t* vals, const int64_t* off, int64_t n_streams,
                         uint8_t* tagk, int64_t* bits, int n_threads) {
  if (n_threads < 1) n_threads = 1;
  if (n_threads > 256) n_threads = 256;
  pthread_t th[256];
  pick_job_t jobs[256];
  for (int t = 0; t < n_threads; ++t) {
    jobs[t] = (pick_job_t){vals, off, n_streams * t / n_threads, n_streams * (t + 1) / n_threads,
                           tagk, bits};
    pthread_create(&th[t], NULL, pick_range, &jobs[t]);
  }
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
}
