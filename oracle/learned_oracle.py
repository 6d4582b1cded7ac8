"""CPU oracle of the learned proposal (test infrastructure: only tests/,
the smoke check and the bench CPU baseline use it; nothing in the product
imports it).

Restates, in numpy:

* features(walks, l2g, strip_len, codes): the 32 per-meshlet features
  (meshpress_amd/csrc/stride_host.cpp meshlet_features, layout in
  include/stride_hip.h stride_encoder_features), from the exported plans.
* the MLP 32 -> 128 -> 128 -> 176 of meshpress_amd/csrc/stride_mlp.hip:
  forward, grouped softmax cross-entropy, backward, Adam.  With
  emulate=True the bf16 roundings of the kernels are applied at the same
  points (weights, activations, logit gradients, hidden gradients) and
  products are summed in float64 (the kernels sum in float32, so results
  agree to float32 rounding); with emulate=False everything is float64,
  which is what the finite-difference checks use to pin the gradients.
* Adam: Kingma & Ba (2015), Algorithm 1, bias-corrected, float32.

Pins: bf16 rounding against hand-worked round-to-nearest-even cases;
gradients against central finite differences of the float64 loss; Adam
against its closed form after one step (-lr * sign(g)) and a hand-computed
second step; features against hand-built meshlets with known sums.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 32, 128, 176
HEAD_OFF = (0, 17, 34, 51, 71, 91, 111, 131, 151, 171)
N_HEADS = 9
IGNORE = 255
N_W = N_IN * N_HID + N_HID * N_HID + N_HID * N_OUT
N_P = N_W + N_HID + N_HID + N_OUT


# ---------------------------------------------------------------- features
def _zz(r):
    r = int(r)
    return -2 * r - 1 if r < 0 else 2 * r


def _irlp(a, b, c, n, K):
    """paradelta_v5_dup.py:91-104: (n0 a + n1 b + n2 c + 2^(K-1)) >> K (floor)."""
    s = n[0] * a + n[1] * b + n[2] * c
    return s if K == 0 else (s + (1 << (K - 1))) >> K


def features(walks, l2g, strip_len, codes, pred=None):
    """walks: per meshlet list of (v, kind, a, b, c); l2g: per meshlet list of
    global ids; strip_len: per meshlet list of strip lengths; codes (N, 3)
    int64; pred: None for the canonical a + b - c, else (pred_n 3x3, pred_K 3)
    for the residuals the stream statistics use.  float32 (n_meshlets, 32)."""
    out = np.zeros((len(walks), 32), np.float32)
    for m, walk in enumerate(walks):
        lg = l2g[m]
        tc = lambda i: [int(x) for x in codes[lg[i]]]  # noqa: E731
        n_local = len(walk)
        n_tris = int(sum(strip_len[m]))
        n_strips = len(strip_len[m])
        n_kind0 = sum(1 for w in walk if w[1] == 0)
        cs = np.array([tc(w[0]) for w in walk], dtype=np.int64).reshape(-1, 3)
        hist = [0, 0, 0, 0]
        for L in strip_len[m]:
            hist[0 if L <= 1 else (1 if L <= 3 else (2 if L <= 15 else 3))] += L
        streams = [[] for _ in range(6)]
        sxy, sxx = [0, 0, 0], [0, 0, 0]
        prev, first = None, True
        for (v, kind, a, b, c) in walk:
            t = tc(v)
            if kind == 0:
                if not first:
                    for d in range(3):
                        streams[d].append(t[d] - prev[d])
                first = False
            else:
                ta, tb, tq = tc(a), tc(b), tc(c)
                for d in range(3):
                    if pred is None:
                        p_ = ta[d] + tb[d] - tq[d]
                    else:
                        p_ = _irlp(ta[d], tb[d], tq[d], [int(x) for x in pred[0][d]], int(pred[1][d]))
                    streams[3 + d].append(t[d] - p_)
                    x, y = ta[d] + tb[d] - 2 * tq[d], t[d] - tq[d]
                    sxy[d] += x * y
                    sxx[d] += x * x
            prev = t
        f = out[m]
        f[0] = n_local / 256.0
        f[1] = n_tris / 256.0
        f[2] = n_strips / n_tris if n_tris else 0.0
        f[3] = n_kind0 / n_local if n_local else 0.0
        for d in range(3):
            f[4 + d] = np.log2(1.0 + float(cs[:, d].max() - cs[:, d].min())) / 16.0 if n_local else 0.0
        for b in range(4):
            f[7 + b] = hist[b] / n_tris if n_tris else 0.0
        for s in range(6):
            us = [_zz(r) for r in streams[s]]
            f[11 + s] = np.log2(1.0 + len(us)) / 8.0
            if us:
                # floor(log2(u + 1)) is the bit length of u + 1, minus one
                f[17 + s] = sum((u + 1).bit_length() - 1 for u in us) / len(us) / 16.0
                f[23 + s] = np.log2(1.0 + sum(us) / len(us)) / 16.0
        for d in range(3):
            s_ = sxy[d] / sxx[d] if sxx[d] > 0 else 1.0
            f[29 + d] = min(2.0, max(0.0, s_)) / 2.0
    return out


# ---------------------------------------------------------------- bf16
def bf16_round(x):
    """float32 -> nearest bf16 value (ties to even), as float32; NaN stays NaN."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(x.shape)


def split(p):
    o1 = N_IN * N_HID
    o2 = o1 + N_HID * N_HID
    o3 = o2 + N_HID * N_OUT
    return (p[:o1].reshape(N_IN, N_HID), p[o1:o2].reshape(N_HID, N_HID),
            p[o2:o3].reshape(N_HID, N_OUT), p[o3:o3 + N_HID], p[o3 + N_HID:o3 + 2 * N_HID],
            p[o3 + 2 * N_HID:])


def _q(a, emulate):
    return bf16_round(np.asarray(a, np.float32)).astype(np.float64) if emulate else np.asarray(a, np.float64)


# ---------------------------------------------------------------- network
def forward(x, p, emulate=True):
    """x (n, 32) -> (h1, h2, z) as float64 arrays (bf16 values when emulating)."""
    W1, W2, W3, b1, b2, b3 = split(np.asarray(p, np.float64 if not emulate else np.float32))
    x = _q(x, emulate)
    W1, W2, W3 = _q(W1, emulate), _q(W2, emulate), _q(W3, emulate)
    b1, b2, b3 = (np.asarray(b, np.float64) for b in (b1, b2, b3))
    h1 = _q(np.maximum(x @ W1 + b1, 0.0), emulate)
    h2 = _q(np.maximum(h1 @ W2 + b2, 0.0), emulate)
    z = h2 @ W3 + b3
    if emulate:
        z = z.astype(np.float32).astype(np.float64)
    return h1, h2, z


def xent(z, labels, scale, emulate=True):
    """(dz (n, 176), loss (n,)): grouped softmax cross-entropy, heads with
    label IGNORE skipped, dz = scale * (softmax - onehot)."""
    z = np.asarray(z, np.float64)
    n = len(z)
    dz = np.zeros((n, N_OUT))
    loss = np.zeros(n)
    for h in range(N_HEADS):
        o, e = HEAD_OFF[h], HEAD_OFF[h + 1]
        y = labels[:, h].astype(np.int64)
        ok = (y != IGNORE) & (y < e - o)
        zh = z[:, o:e]
        mx = zh.max(axis=1, keepdims=True)
        ex = np.exp(zh - mx)
        s = ex.sum(axis=1, keepdims=True)
        pr = ex / s
        yy = np.where(ok, y, 0)
        lh = np.log(s[:, 0]) + mx[:, 0] - zh[np.arange(n), yy]
        loss += np.where(ok, lh, 0.0)
        g = pr.copy()
        g[np.arange(n), yy] -= 1.0
        dz[:, o:e] = np.where(ok[:, None], scale * g, 0.0)
    if emulate:
        dz = _q(dz, True)
    return dz, loss


def backward(x, h1, h2, dz, p, emulate=True, return_hidden=False):
    """Gradient (N_P,) of the loss behind dz, in the kernels' layout."""
    W1, W2, W3, _, _, _ = split(np.asarray(p, np.float32 if emulate else np.float64))
    x = _q(x, emulate)
    W2, W3 = _q(W2, emulate), _q(W3, emulate)
    h1, h2, dz = (np.asarray(a, np.float64) for a in (h1, h2, dz))
    g = np.zeros(N_P)
    gW1, gW2, gW3, gb1, gb2, gb3 = split(g)
    gW3[:] = h2.T @ dz
    gb3[:] = dz.sum(axis=0)
    dh2 = _q((dz @ W3.T) * (h2 > 0), emulate)
    gW2[:] = h1.T @ dh2
    gb2[:] = dh2.sum(axis=0)
    dh1 = _q((dh2 @ W2.T) * (h1 > 0), emulate)
    gW1[:] = x.T @ dh1
    gb1[:] = dh1.sum(axis=0)
    if return_hidden:
        return g, dh2, dh1
    return g


def loss64(p, x, labels):
    """Mean float64 loss (finite-difference target)."""
    _, _, z = forward(x, p, emulate=False)
    _, l = xent(z, labels, 1.0, emulate=False)
    return l.sum() / len(x)


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32 with the kernels' operation order."""
    f = np.float32
    p, m, v, g = (np.asarray(a, np.float32) for a in (p, m, v, g))
    c1 = f(1.0 / (1.0 - b1 ** step))
    c2 = f(1.0 / (1.0 - b2 ** step))
    m2 = f(b1) * m + (f(1) - f(b1)) * g
    v2 = f(b2) * v + (f(1) - f(b2)) * g * g
    p2 = p - f(lr) * (m2 * c1) / (np.sqrt(v2 * c2) + f(eps))
    return p2.astype(np.float32), m2.astype(np.float32), v2.astype(np.float32)


def heads(z):
    """(pred_prob (n, 51), coder class (n, 6), coder class probability (n, 6))."""
    z = np.asarray(z, np.float64)
    n = len(z)
    pp = np.zeros((n, 51))
    cls = np.zeros((n, 6), np.uint8)
    cp = np.zeros((n, 6))
    for h in range(N_HEADS):
        o, e = HEAD_OFF[h], HEAD_OFF[h + 1]
        zh = z[:, o:e]
        am = zh.argmax(axis=1)
        ex = np.exp(zh - zh.max(axis=1, keepdims=True))
        pr = ex / ex.sum(axis=1, keepdims=True)
        if h < 3:
            pp[:, o:e] = pr
        else:
            cls[:, h - 3] = am
            cp[:, h - 3] = pr[np.arange(n), am]
    return pp, cls, cp
