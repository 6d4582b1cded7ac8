"""CPU oracle for the STRIDE (DPR4 v3) codec path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it
as the checker.  It restates, in plain Python + numpy, the reference's
algorithm for the path (maletsden/meshpress):

  bit codec        utils/bit_codec.py:16-257       BitWriter / BitReader
  coder costs      utils/residual_entropy.py:34-54, encoder/paradelta_codec.py:124-146
  coder choice     encoder/paradelta_v5_dup.py:159-171    _pick_best_k
  FIFO tokens      encoder/paradelta_codec.py:206-229,330-340
  walk             encoder/paradelta_v5_dup.py:114-152    _walk_meshlet
  IRLP predictor   encoder/paradelta_v5_dup.py:91-104     _irlp_pred
  writer           encoder/paradelta_v5_dup.py:239-464    _write_meshlet_dup / encode_dup
  decoder          encoder/paradelta_v5_dup.py:471-618    decode_dup
  IRLP cost/fit    encoder/_irlp_fit.py:102-394           encoder_cost_batch / fit_predictor

How each part is pinned (see tests/test_oracle.py):
  * bit codec: the reference's import-time known-answer test
    (utils/bit_codec.py:262-288) and closed-form field layouts;
  * coder choice and costs: brute force over the 20 candidates, exact bit
    counts against the emitted stream length;
  * decode_dup / encode_dup: golden vectors written by the reference's own
    decode_dup/encode_dup (tests/golden/), hand-derived byte fixtures, the
    round-trip invariant decoded codes == global_codes[l2g] (the reference's
    PASS criterion, scripts/verify_stride_dup.py:16);
  * dequantisation: the IEEE f64 formula evaluated independently;
  * IRLP cost: brute force per meshlet; fit_predictor: golden (n, K) from the
    reference's own search on the same samples.
"""
from __future__ import annotations

import math
import struct
from collections import deque

import numpy as np

MAGIC_DUP = 0x44505234
VERSION_DUP = 3
FIFO_SIZE = 16
REUSE_BITS = 5            # ceil(log2(FIFO_SIZE + 1))
RICE_K_MAX = 11
EG_K_MAX = 7
CANON_N = np.array([[1, 1, -1]] * 3, dtype=np.int64)
CANON_K = np.zeros(3, dtype=np.int64)


# ---------------------------------------------------------------------------
# Bit codec (utils/bit_codec.py:16-257).  MSB first; bit 7 of byte 0 first.
# ---------------------------------------------------------------------------
class BitWriter:
    def __init__(self):
        self.bits: list[int] = []

    def put(self, value: int, n: int) -> None:
        if n <= 0:
            return
        value &= (1 << n) - 1
        self.bits.extend((value >> (n - 1 - i)) & 1 for i in range(n))

    def put_rice(self, u: int, k: int) -> None:
        if u < 0:
            raise ValueError("Rice(k) expects non-negative input")
        q = u >> k
        self.bits.extend([0] * q)
        self.bits.append(1)
        self.put(u & ((1 << k) - 1), k)

    def put_eg(self, u: int, k: int) -> None:
        if u < 0:
            raise ValueError("EG(k) expects non-negative input")
        x = (u >> k) + 1
        lb = x.bit_length() - 1
        self.bits.extend([0] * lb)
        self.bits.append(1)
        self.put(x - (1 << lb), lb)
        self.put(u & ((1 << k) - 1), k)

    def put_zigzag(self, v: int, n: int) -> None:
        self.put(2 * v if v >= 0 else -2 * v - 1, n)

    def put_f32(self, v: float) -> None:
        self.put(struct.unpack(">I", struct.pack(">f", v))[0], 32)

    def bit_pos(self) -> int:
        return len(self.bits)

    def tobytes(self) -> bytes:
        b = self.bits + [0] * ((-len(self.bits)) % 8)
        out = bytearray()
        for i in range(0, len(b), 8):
            v = 0
            for j in range(8):
                v = (v << 1) | b[i + j]
            out.append(v)
        return bytes(out)


class BitReader:
    def __init__(self, data: bytes, pos: int = 0):
        self.data = data
        self.pos = pos

    def bit(self) -> int:
        byte = self.data[self.pos >> 3]          # IndexError past the end
        b = (byte >> (7 - (self.pos & 7))) & 1
        self.pos += 1
        return b

    def get(self, n: int) -> int:
        v = 0
        for _ in range(n):
            v = (v << 1) | self.bit()
        return v

    def unary(self) -> int:
        q = 0
        while self.bit() == 0:
            q += 1
        return q

    def rice(self, k: int) -> int:
        q = self.unary()
        return (q << k) | self.get(k)

    def eg(self, k: int) -> int:
        lb = self.unary()
        shifted = (1 << lb) - 1 + self.get(lb) if lb else 0
        return (shifted << k) | self.get(k)

    def zigzag(self, n: int) -> int:
        u = self.get(n)
        return (u >> 1) ^ -(u & 1)

    def f32(self) -> float:
        return struct.unpack(">f", struct.pack(">I", self.get(32)))[0]


def zigzag(v):
    v = np.asarray(v, dtype=np.int64)
    return np.where(v >= 0, 2 * v, -2 * v - 1)


def unzigzag(u: int) -> int:
    return (u >> 1) ^ -(u & 1)


# ---------------------------------------------------------------------------
# Coder costs and choice
# ---------------------------------------------------------------------------
def rice_bits(u: np.ndarray, k: int) -> int:
    """Total Rice(k) bits: sum(u >> k) + n (1 + k)."""
    return int((u >> k).sum()) + len(u) * (1 + k) if len(u) else 0


def eg_bits(u: np.ndarray, k: int) -> int:
    """Total Exp-Golomb(k) bits: sum(2 floor(log2((u >> k) + 1)) + 1 + k)."""
    if not len(u):
        return 0
    x = (u >> k) + 1
    lb = np.array([int(t).bit_length() - 1 for t in x], dtype=np.int64)
    return int((2 * lb + 1 + k).sum())


def stream_costs(vals) -> list[int]:
    """The 20 candidate body costs: Rice k = 0..11, then Exp-Golomb k = 0..7."""
    u = zigzag(vals)
    return [rice_bits(u, k) for k in range(RICE_K_MAX + 1)] + \
           [eg_bits(u, k) for k in range(EG_K_MAX + 1)]


def pick_tag_k(vals) -> tuple[int, int, int]:
    """(tag, k, bits) for one stream: lowest k wins ties, Rice wins ties vs EG,
    empty stream -> Rice k=0 (paradelta_v5_dup.py:159-171)."""
    if len(vals) == 0:
        return 1, 0, 0
    c = stream_costs(vals)
    rk = min(range(RICE_K_MAX + 1), key=lambda k: (c[k], k))
    ek = min(range(EG_K_MAX + 1), key=lambda k: (c[RICE_K_MAX + 1 + k], k))
    rb, eb = c[rk], c[RICE_K_MAX + 1 + ek]
    return (1, rk, rb) if rb <= eb else (2, ek, eb)


# ---------------------------------------------------------------------------
# Connectivity helpers
# ---------------------------------------------------------------------------
def idx_bits_for(n_local: int) -> int:
    return max(1, (n_local).bit_length())    # == ceil(log2(n_local + 1))


def root_orient(tris_local, root_id, next_id):
    root = [int(x) for x in tris_local[root_id]]
    if next_id is None:
        return root
    shared = set(root) & set(int(x) for x in tris_local[next_id])
    if len(shared) != 2:
        return root
    third = (set(root) - shared).pop()
    lo, hi = sorted(shared)
    return [third, lo, hi]


def strip_step(prev, tri):
    """(edge_code, (a, b, c), new_v, next_prev) or None if the triangle does
    not share an edge containing the newest vertex."""
    ts = set(int(x) for x in tri)
    shared = ts & set(prev)
    if len(shared) != 2 or len(ts - shared) != 1:
        return None
    nv = next(iter(ts - shared))
    if shared == {prev[1], prev[2]}:
        return 0, (prev[1], prev[2], prev[0]), nv, [prev[1], prev[2], nv]
    if shared == {prev[0], prev[2]}:
        return 1, (prev[0], prev[2], prev[1]), nv, [prev[0], prev[2], nv]
    return None


def emit_vert(w: BitWriter, v: int, fifo: deque, idx_bits: int) -> None:
    if v in fifo:
        w.put(0, 1)
        w.put(list(fifo).index(v), REUSE_BITS)
        fifo.remove(v)
    else:
        w.put(1, 1)
        w.put(v, idx_bits)
    fifo.append(v)


def read_vert(r: BitReader, fifo: deque, idx_bits: int) -> int:
    if r.bit() == 0:
        v = list(fifo)[r.get(REUSE_BITS)]     # IndexError on a bad FIFO index
    else:
        v = r.get(idx_bits)
    if v in fifo:
        fifo.remove(v)
    fifo.append(v)
    return v


def walk_meshlet(tris_local, strips):
    """[(v, kind, refs)] with kind 'none' (strip-root vertex) or 'para'."""
    n = int(np.max(tris_local)) + 1 if len(tris_local) else 0
    seen = np.zeros(max(n, 1), dtype=bool)
    order = []
    for strip in strips:
        root = root_orient(tris_local, strip[0], strip[1] if len(strip) > 1 else None)
        for v in root:
            if not seen[v]:
                order.append((v, "none", None))
                seen[v] = True
        prev = list(root)
        for ti in strip[1:]:
            st = strip_step(prev, tris_local[ti])
            if st is None:
                continue
            _, refs, nv, prev = st
            if not seen[nv]:
                order.append((nv, "para", refs))
                seen[nv] = True
    return order


def irlp_pred(a, b, c, n3x3, K3):
    out = np.empty(3, dtype=np.int64)
    for d in range(3):
        s = int(n3x3[d][0]) * int(a[d]) + int(n3x3[d][1]) * int(b[d]) + int(n3x3[d][2]) * int(c[d])
        K = int(K3[d])
        out[d] = s if K == 0 else (s + (1 << (K - 1))) >> K
    return out


# ---------------------------------------------------------------------------
# Encoder (given plans)
# ---------------------------------------------------------------------------
def meshlet_residuals(plan, global_codes, pred_n, pred_K):
    """(anchor[3], delta (n_delta,3), para (n_para,3)) in walk order."""
    l2g = np.asarray(plan["local_to_global"], dtype=np.int64)
    tc = global_codes[l2g].astype(np.int64)
    anchor, prev = None, None
    deltas, paras = [], []
    for v, kind, refs in walk_meshlet(np.asarray(plan["ml_tris_local"]), plan["strips"]):
        if kind == "none":
            if anchor is None:
                anchor = tc[v].copy()
            else:
                deltas.append(tc[v] - prev)
        else:
            a, b, c = refs
            paras.append(tc[v] - irlp_pred(tc[a], tc[b], tc[c], pred_n, pred_K))
        prev = tc[v]
    if anchor is None:
        anchor = np.zeros(3, dtype=np.int64)
    d = np.array(deltas, dtype=np.int64).reshape(-1, 3)
    p = np.array(paras, dtype=np.int64).reshape(-1, 3)
    return anchor, d, p


def encode_dup(prep: dict, pred_n=None, pred_K=None, tagk=None) -> bytes:
    """DPR4 v3 bytes for a prep dict (the reference's field names).  The
    predictor defaults to canonical; tagk (n_m, 12) overrides the exhaustive
    coder choice."""
    pred_n = CANON_N if pred_n is None else np.asarray(pred_n, dtype=np.int64)
    pred_K = CANON_K if pred_K is None else np.asarray(pred_K, dtype=np.int64)
    g_bits = [int(x) for x in prep["g_bits"]]
    body = BitWriter()
    rel = []
    total_local = 0
    max_abs = 0
    for m, plan in enumerate(prep["plans"]):
        tl = np.asarray(plan["ml_tris_local"], dtype=np.int64)
        n_local = len(plan["local_to_global"])
        anchor, dl, pa = meshlet_residuals(plan, prep["global_codes"], pred_n, pred_K)
        for arr in (dl, pa):
            if arr.size:
                max_abs = max(max_abs, int(np.abs(arr).max()))
        rel.append(body.bit_pos())
        total_local += n_local
        body.put(n_local, 16)
        body.put(len(tl), 16)
        body.put(len(plan["strips"]), 16)
        ib = idx_bits_for(n_local)
        fifo: deque = deque(maxlen=FIFO_SIZE)
        for strip in plan["strips"]:
            body.put(len(strip), 16)
            root = root_orient(tl, strip[0], strip[1] if len(strip) > 1 else None)
            for v in root:
                emit_vert(body, v, fifo, ib)
            prev = list(root)
            for ti in strip[1:]:
                st = strip_step(prev, tl[ti])
                if st is None:
                    raise RuntimeError("oldest-edge share")
                ec, _, nv, prev = st
                body.put(ec, 1)
                emit_vert(body, nv, fifo, ib)
        for d in range(3):
            body.put(int(anchor[d]), g_bits[d])
        if tagk is None:
            tk = []
            for arr in (dl, pa):
                for d in range(3):
                    t, k, _ = pick_tag_k(arr[:, d])
                    tk += [t, k]
        else:
            tk = [int(x) for x in tagk[m]]
        for x in tk:
            body.put(x, 8)
        for s in range(6):
            arr = dl if s < 3 else pa
            tag, k = tk[2 * s], tk[2 * s + 1]
            for u in zigzag(arr[:, s % 3]):
                (body.put_rice if tag == 1 else body.put_eg)(int(u), k)
    w = BitWriter()
    w.put(MAGIC_DUP, 32)
    w.put(VERSION_DUP, 8)
    w.put(1 if max_abs > 32767 else 0, 8)
    for c in prep["center"]:
        w.put_f32(float(c))
    w.put_f32(float(prep["scale"]))
    w.put_f32(float(prep["per_coord_err"]))
    for v in prep["g_min"]:
        w.put_f32(float(v))
    for v in prep["g_range"]:
        w.put_f32(float(v))
    for b in g_bits:
        w.put(b, 8)
    for d in range(3):
        for j in range(3):
            w.put(int(pred_n[d][j]) & 0xFFFF, 16)
    for d in range(3):
        w.put(int(pred_K[d]) & 0xFF, 8)
    w.put(total_local, 32)
    w.put(int(prep["n_t"]), 32)
    w.put(len(prep["plans"]), 32)
    region = w.bit_pos() + 32 * len(prep["plans"])
    for r in rel:
        w.put(region + r, 32)
    return w.tobytes() + body.tobytes()


# ---------------------------------------------------------------------------
# Decoder
# ---------------------------------------------------------------------------
def read_header(data: bytes) -> dict:
    r = BitReader(data)
    magic = r.get(32)
    if magic != MAGIC_DUP:
        raise ValueError(f"bad magic 0x{magic:08X}")
    version = r.get(8)
    if version != VERSION_DUP:
        raise ValueError(f"version {version}")
    h = {"code_width": r.get(8),
         "center": [r.f32() for _ in range(3)], "scale": r.f32(), "per_coord_err": r.f32(),
         "g_min": [r.f32() for _ in range(3)], "g_range": [r.f32() for _ in range(3)],
         "g_bits": [r.get(8) for _ in range(3)]}
    pn = []
    for _ in range(9):
        u = r.get(16)
        pn.append(u - 0x10000 if u & 0x8000 else u)
    h["pred_n"] = np.array(pn, dtype=np.int64).reshape(3, 3)
    h["pred_K"] = np.array([r.get(8) for _ in range(3)], dtype=np.int64)
    h["n_v"], h["n_t"], h["n_meshlets"] = r.get(32), r.get(32), r.get(32)
    h["offsets"] = [r.get(32) for _ in range(h["n_meshlets"])]
    h["body_bit"] = r.pos
    return h


def decode_meshlet(r: BitReader, h: dict):
    """Decode one meshlet at the reader's position: (codes (n_local,3) int64,
    local tris [(a,b,c)])."""
    g_bits = h["g_bits"]
    n_local, n_tris, n_strips = r.get(16), r.get(16), r.get(16)
    ib = idx_bits_for(n_local)
    fifo: deque = deque(maxlen=FIFO_SIZE)
    tris, walk, seen = [], [], set()
    for _ in range(n_strips):
        L = r.get(16)
        p = [read_vert(r, fifo, ib) for _ in range(3)]
        tris.append(tuple(p))
        for v in p:
            if v not in seen:
                walk.append((v, None))
                seen.add(v)
        for _ in range(L - 1):
            ec = r.bit()
            nv = read_vert(r, fifo, ib)
            if ec == 0:
                refs = (p[1], p[2], p[0])
                p = [p[1], p[2], nv]
            else:
                refs = (p[0], p[2], p[1])
                p = [p[0], p[2], nv]
            tris.append((refs[0], refs[1], nv))
            if nv not in seen:
                walk.append((nv, refs))
                seen.add(nv)
    anchor = [r.get(g_bits[d]) for d in range(3)]
    tk = [r.get(8) for _ in range(12)]
    n_k0 = sum(1 for _, refs in walk if refs is None)
    n_delta, n_para = max(0, n_k0 - 1), len(walk) - n_k0
    streams = []
    for s in range(6):
        tag, k = tk[2 * s], tk[2 * s + 1]
        cnt = n_delta if s < 3 else n_para
        rd = r.rice if tag == 1 else r.eg       # any tag other than 1 reads as EG
        streams.append([unzigzag(rd(k)) for _ in range(cnt)])
    codes = np.zeros((n_local, 3), dtype=np.int64)
    prev, first, di, pi = None, True, 0, 0
    for v, refs in walk:
        if refs is None:
            if first:
                codes[v] = anchor
                first = False
            else:
                codes[v] = prev + np.array([streams[d][di] for d in range(3)])
                di += 1
        else:
            a, b, c = refs
            codes[v] = irlp_pred(codes[a], codes[b], codes[c], h["pred_n"], h["pred_K"]) + \
                np.array([streams[3 + d][pi] for d in range(3)])
            pi += 1
        prev = codes[v]
    return codes, tris


def dequantize(codes: np.ndarray, h: dict) -> np.ndarray:
    """float32 world positions: ((g_min + (code / mx) * g_range) * scale + center)
    in float64, rounded once to float32 (paradelta_v5_dup.py:607-618)."""
    vn = np.zeros(codes.shape, dtype=np.float64)
    for d in range(3):
        mx = (1 << int(h["g_bits"][d])) - 1
        gmin, grng = float(h["g_min"][d]), float(h["g_range"][d])
        vn[:, d] = gmin if mx == 0 else gmin + codes[:, d].astype(np.float64) / mx * grng
    center = np.array(h["center"], dtype=np.float64)
    return (vn * float(h["scale"]) + center).astype(np.float32)


def decode_dup(data: bytes, return_codes: bool = False):
    """(V float32 (n_v,3), T int64 (n_t,3)) exactly as the reference decoder."""
    h = read_header(data)
    r = BitReader(data, h["body_bit"])
    all_codes, all_tris, base = [], [], 0
    for _ in range(h["n_meshlets"]):
        codes, tris = decode_meshlet(r, h)
        all_codes.append(codes)
        all_tris.extend((base + a, base + b, base + c) for a, b, c in tris)
        base += len(codes)
    codes = np.concatenate(all_codes) if all_codes else np.zeros((0, 3), dtype=np.int64)
    V = dequantize(codes, h)
    T = np.array(all_tris, dtype=np.int64).reshape(-1, 3)
    return (V, T, codes) if return_codes else (V, T)


# ---------------------------------------------------------------------------
# IRLP cost and fit (encoder/_irlp_fit.py)
# ---------------------------------------------------------------------------
def cost_min_coder(u: np.ndarray) -> int:
    """min over the 20 coders of the body bits of one meshlet's stream."""
    if len(u) == 0:
        return 0
    best = None
    for k in range(RICE_K_MAX + 1):
        c = int((u >> k).sum()) + len(u) * (k + 1)
        best = c if best is None or c < best else best
    for k in range(EG_K_MAX + 1):
        x = (u >> k) + 1
        lb = np.floor(np.log2(x.astype(np.float64))).astype(np.int64)
        c = int((2 * lb + 1 + k).sum())
        best = c if c < best else best
    return best


def irlp_residual(a, b, c, v, n3, K):
    s = n3[0] * a + n3[1] * b + n3[2] * c
    pred = s if K == 0 else (s + (1 << (K - 1))) >> K
    return v - pred


def irlp_cost(a, b, c, v, off, n3, K) -> int:
    """Sum over meshlets of min-coder bits of one axis's parallelogram
    residuals under predictor (n3, K) (_cost_seeds_njit)."""
    a, b, c, v = (np.asarray(x, dtype=np.int64) for x in (a, b, c, v))
    u = zigzag(irlp_residual(a, b, c, v, [int(x) for x in n3], int(K)))
    return sum(cost_min_coder(u[off[m]:off[m + 1]]) for m in range(len(off) - 1))


def irls_l1(X, y, n_iter=12, eps=1.0):
    w, *_ = np.linalg.lstsq(X, y, rcond=None)
    for _ in range(n_iter):
        r = y - X @ w
        sw = np.sqrt(1.0 / np.maximum(np.abs(r), eps))
        w_new, *_ = np.linalg.lstsq(X * sw[:, None], y * sw, rcond=None)
        if np.max(np.abs(w_new - w)) < 1e-6:
            return w_new
        w = w_new
    return w


def search_axis(a, b, c, v, off, w_axis, cost_fn=None):
    """(K, (n0, n1, n2), bits): K in 5..9, 10 seeds, 2 rounds of +-1
    coordinate descent, skip a K whose best seed exceeds 1.05 x best."""
    cost_fn = cost_fn or (lambda cands, K: np.array(
        [irlp_cost(a, b, c, v, off, n3, K) for n3 in cands], dtype=np.int64))
    best = None
    moves = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))
    for K in (5, 6, 7, 8, 9):
        D = 1 << K
        base = [int(round(w_axis[i] * D)) for i in range(3)]
        seeds = [tuple(base), (D, D, -D)]
        for s0 in (-1, 1):
            for s1 in (-1, 1):
                for s2 in (-1, 1):
                    seeds.append((base[0] + s0, base[1] + s1, base[2] + s2))
        cur = np.array(seeds, dtype=np.int64)
        cur_cost = cost_fn(cur, K)
        if best is not None and int(cur_cost.min()) > 1.05 * best[2]:
            continue
        active = np.ones(len(cur), dtype=bool)
        for _ in range(2):
            if not active.any():
                break
            sel = np.where(active)[0]
            cands = np.repeat(cur[sel], 6, axis=0)
            for j, (i, sg) in enumerate(moves):
                cands[j::6, i] += sg
            cc = cost_fn(cands, K)
            nxt = np.zeros(len(cur), dtype=bool)
            for slot, si in enumerate(sel):
                g = cc[6 * slot:6 * slot + 6]
                j = int(g.argmin())
                if g[j] < cur_cost[si]:
                    cur[si] = cands[6 * slot + j]
                    cur_cost[si] = g[j]
                    nxt[si] = True
            active = nxt
        bi = int(cur_cost.argmin())
        if best is None or int(cur_cost[bi]) < best[2]:
            best = (K, tuple(int(x) for x in cur[bi]), int(cur_cost[bi]))
    return best


def fit_predictor(A, B, C, V, off, cost_fn_factory=None, subsample=4000, full_cost_fn=None):
    """(n 3x3, K 3) per-axis IRLP fit with canonical fallback on all samples.
    cost_fn_factory(a, b, c, v, off) -> cost_fn(cands, K) replaces the search's
    candidate costs, full_cost_fn(a, b, c, v, off, n3, K) -> bits the fallback
    comparison's; both default to irlp_cost (same values, other code)."""
    full_cost_fn = full_cost_fn or irlp_cost
    A, B, C, V = (np.asarray(x, dtype=np.int64) for x in (A, B, C, V))
    off = np.asarray(off, dtype=np.int64)
    n_m = len(off) - 1
    if n_m == 0 or len(A) == 0:
        return CANON_N.copy(), CANON_K.copy()
    if n_m > subsample:
        sel = np.sort(np.random.default_rng(0).choice(n_m, size=subsample, replace=False))
        rows = np.concatenate([np.arange(off[i], off[i + 1]) for i in sel])
        As, Bs, Cs, Vs = A[rows], B[rows], C[rows], V[rows]
        off_s = np.concatenate(([0], np.cumsum(np.diff(off)[sel])))
    else:
        As, Bs, Cs, Vs, off_s = A, B, C, V, off
    n_fit = np.zeros((3, 3), dtype=np.int64)
    K_fit = np.zeros(3, dtype=np.int64)
    for d in range(3):
        X = np.stack([As[:, d], Bs[:, d], Cs[:, d]], axis=1).astype(np.float64)
        w = irls_l1(X, Vs[:, d].astype(np.float64))
        cf = cost_fn_factory(As[:, d], Bs[:, d], Cs[:, d], Vs[:, d], off_s) if cost_fn_factory else None
        K, ns, _ = search_axis(As[:, d], Bs[:, d], Cs[:, d], Vs[:, d], off_s, w, cf)
        n_fit[d] = ns
        K_fit[d] = K
    for d in range(3):
        b_fit = full_cost_fn(A[:, d], B[:, d], C[:, d], V[:, d], off, n_fit[d], int(K_fit[d]))
        b_can = full_cost_fn(A[:, d], B[:, d], C[:, d], V[:, d], off, CANON_N[d], 0)
        if b_can < b_fit:
            n_fit[d] = CANON_N[d]
            K_fit[d] = 0
    return n_fit, K_fit
