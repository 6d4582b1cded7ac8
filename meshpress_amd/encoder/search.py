"""GPU-backed encoder search: exhaustive per-stream coder choice and the
per-mesh IRLP predictor fit.

Mirrors (maletsden/meshpress):
  _pick_best_k            encoder/paradelta_v5_dup.py:159-171
  fit_predictor           encoder/_irlp_fit.py:364-394
  _search_axis            encoder/_irlp_fit.py:265-316
  _irls_l1 / _fit_irls    encoder/_irlp_fit.py:239-257
The candidate costs are evaluated by the stride_irlp_cost kernel; the search
control flow (seeds, coordinate descent, early stop) stays on the host.
"""
from __future__ import annotations

import numpy as np

from .. import _lib

KS = (5, 6, 7, 8, 9)
SEARCH_RADIUS = 1
COORD_DESCENT_ROUNDS = 2
IRLS_ITERS = 12
IRLS_EPS = 1.0
SUBSAMPLE_MESHLETS = 4000
EARLY_STOP_SLACK = 1.05
CANON_N = np.array([[1, 1, -1]] * 3, dtype=np.int64)
CANON_K = np.zeros(3, dtype=np.int64)
N_CODERS = 20          # Rice k 0..11, Exp-Golomb k 0..7


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the STRIDE encoder search runs on a ROCm GPU; torch.cuda is unavailable")
    return torch


def _dev(a, dtype):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def select_codes(vals: np.ndarray, off: np.ndarray):
    """(tag,k) uint8 (n_streams, 2) and body bits int64 (n_streams,)."""
    torch = _torch()
    n = len(off) - 1
    d_vals = _dev(vals if len(vals) else np.zeros(1, np.int32), torch.int32)
    d_off = _dev(off, torch.int64)
    d_tagk = torch.empty((max(n, 1), 2), dtype=torch.uint8, device="cuda")
    d_bits = torch.empty(max(n, 1), dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().stride_select_codes(_lib.ptr(d_vals), _lib.ptr(d_off), n, _lib.ptr(d_tagk),
                                              _lib.ptr(d_bits), _lib.stream_ptr()))
    return d_tagk[:n].cpu().numpy(), d_bits[:n].cpu().numpy()


def stream_costs(vals: np.ndarray, off: np.ndarray) -> np.ndarray:
    """All 20 candidate body costs per stream, int64 (n_streams, 20)."""
    torch = _torch()
    n = len(off) - 1
    d_vals = _dev(vals if len(vals) else np.zeros(1, np.int32), torch.int32)
    d_off = _dev(off, torch.int64)
    d_c = torch.empty((max(n, 1), N_CODERS), dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().stride_stream_costs(_lib.ptr(d_vals), _lib.ptr(d_off), n, _lib.ptr(d_c),
                                              _lib.stream_ptr()))
    return d_c[:n].cpu().numpy()


class IrlpCost:
    """Device-resident samples of one axis; evaluates candidate costs."""

    def __init__(self, a, b, c, v, off):
        torch = _torch()
        self.n_m = len(off) - 1
        z = np.zeros(1, np.int32)
        self.a = _dev(a if len(a) else z, torch.int32)
        self.b = _dev(b if len(b) else z, torch.int32)
        self.c = _dev(c if len(c) else z, torch.int32)
        self.v = _dev(v if len(v) else z, torch.int32)
        self.off = _dev(off, torch.int64)
        self.evals = 0

    def __call__(self, cands: np.ndarray, K: int) -> np.ndarray:
        torch = _torch()
        cands = np.ascontiguousarray(np.asarray(cands, dtype=np.int32).reshape(-1, 3))
        S = len(cands)
        d_c = _dev(cands, torch.int32)
        d_out = torch.empty(max(S, 1), dtype=torch.int64, device="cuda")
        _lib.check(_lib.lib().stride_irlp_cost(
            _lib.ptr(self.a), _lib.ptr(self.b), _lib.ptr(self.c), _lib.ptr(self.v), _lib.ptr(self.off),
            self.n_m, _lib.ptr(d_c), S, int(K), _lib.ptr(d_out), _lib.stream_ptr()))
        self.evals += S
        return d_out[:S].cpu().numpy()


class IrlpCostAxes:
    """Device-resident samples of all three axes, uploaded once (columns
    (4, 3, n): a, b, c, v by axis, and int64 meshlet offsets).  run() takes
    the candidates of several axes, uploads them in one copy, launches one
    cost evaluation per axis and reads every cost back in one transfer."""

    def __init__(self, cols, off):
        torch = _torch()
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.n = int(cols.shape[2])
        self.n_m = len(off) - 1
        self.cols = _dev(cols if self.n else np.zeros((4, 3, 1), np.int32), torch.int32)
        self.off = _dev(off, torch.int64)
        self.evals = 0

    def run(self, items, K: int):
        """[(axis, candidates (S, 3))] -> [int64 costs (S,)], in order."""
        torch = _torch()
        cs = [np.asarray(c, dtype=np.int32).reshape(-1, 3) for _, c in items]
        total = sum(len(c) for c in cs)
        if total == 0:
            return [np.zeros(0, np.int64) for _ in cs]
        d_c = _dev(np.concatenate(cs), torch.int32)
        d_out = torch.empty(total, dtype=torch.int64, device="cuda")
        row = 4 * max(self.n, 1)                 # bytes of one (column, axis) row
        base = self.cols.data_ptr()
        o = 0
        for (d, _), c in zip(items, cs):
            col = [base + row * (3 * j + d) for j in range(4)]
            _lib.check(_lib.lib().stride_irlp_cost(
                col[0], col[1], col[2], col[3], _lib.ptr(self.off), self.n_m,
                d_c.data_ptr() + 12 * o, len(c), int(K), d_out.data_ptr() + 8 * o,
                _lib.stream_ptr()))
            o += len(c)
        self.evals += total
        return np.split(d_out.cpu().numpy(), np.cumsum([len(c) for c in cs])[:-1])


def irls_l1(X, y, n_iter=IRLS_ITERS, eps=IRLS_EPS):
    w, *_ = np.linalg.lstsq(X, y, rcond=None)
    for _ in range(n_iter):
        r = y - X @ w
        sw = np.sqrt(1.0 / np.maximum(np.abs(r), eps))
        w_new, *_ = np.linalg.lstsq(X * sw[:, None], y * sw, rcond=None)
        if np.max(np.abs(w_new - w)) < 1e-6:
            return w_new
        w = w_new
    return w


def search_axis(cost, w_axis):
    """(K, (n0, n1, n2), bits) under min(Rice, EG) cost."""
    best = None
    moves = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))
    for K in KS:
        D = 1 << K
        base = [int(round(w_axis[i] * D)) for i in range(3)]
        seeds = [tuple(base), (D, D, -D)]
        for s0 in (-SEARCH_RADIUS, SEARCH_RADIUS):
            for s1 in (-SEARCH_RADIUS, SEARCH_RADIUS):
                for s2 in (-SEARCH_RADIUS, SEARCH_RADIUS):
                    seeds.append((base[0] + s0, base[1] + s1, base[2] + s2))
        cur = np.array(seeds, dtype=np.int64)
        cur_cost = cost(cur, K)
        if best is not None and int(cur_cost.min()) > EARLY_STOP_SLACK * best[2]:
            continue
        active = np.ones(len(cur), dtype=bool)
        for _ in range(COORD_DESCENT_ROUNDS):
            if not active.any():
                break
            sel = np.where(active)[0]
            cands = np.repeat(cur[sel], 6, axis=0)
            for j, (i, sg) in enumerate(moves):
                cands[j::6, i] += sg
            cc = cost(cands, K)
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


def subsample_sel(n_m, n_keep=SUBSAMPLE_MESHLETS):
    """Sorted meshlet selection of the IRLP fit (encoder/_irlp_fit.py:72-97:
    default_rng(0).choice without replacement), or None to keep every meshlet."""
    if n_m <= n_keep:
        return None
    return np.sort(np.random.default_rng(0).choice(n_m, size=n_keep, replace=False))


def as_columns(A, B, C, V):
    """(n, 3) sample arrays -> int32 columns (4, 3, n)."""
    return np.ascontiguousarray(np.stack([np.asarray(x).reshape(-1, 3).T for x in (A, B, C, V)]),
                                dtype=np.int32)


def subsample(A, B, C, V, off, n_keep=SUBSAMPLE_MESHLETS):
    n_m = len(off) - 1
    sel = subsample_sel(n_m, n_keep)
    if sel is None:
        return A, B, C, V, off
    rows = np.concatenate([np.arange(off[i], off[i + 1]) for i in sel])
    counts = np.diff(off)[sel]
    return A[rows], B[rows], C[rows], V[rows], np.concatenate(([0], np.cumsum(counts)))


def _axis_cost(cols, off, d):
    return IrlpCost(cols[0, d], cols[1, d], cols[2, d], cols[3, d], off)


def fit_axis_cols(sub, full, d):
    """Exhaustive fit of axis d (the per-axis body of fit_predictor).
    sub, full: (columns (4, 3, n), offsets) of the subsample and of every
    meshlet.  ((n0, n1, n2) int64, K, bits over all samples, evaluations)."""
    cols, off_s = sub
    X = np.stack([cols[0, d], cols[1, d], cols[2, d]], axis=1).astype(np.float64)
    w = irls_l1(X, cols[3, d].astype(np.float64))
    cost = _axis_cost(cols, off_s, d)
    K, ns, _ = search_axis(cost, w)
    fcost = _axis_cost(full[0], full[1], d)
    b_fit = int(fcost(np.array(ns)[None], int(K))[0])
    b_can = int(fcost(CANON_N[d][None], 0)[0])
    if b_can < b_fit:
        return CANON_N[d].copy(), 0, b_can, cost.evals + fcost.evals
    return np.array(ns, dtype=np.int64), int(K), b_fit, cost.evals + fcost.evals


def fit_predictor_cols(sub, full, stats=None):
    """(n 3x3 int64, K 3 int64) from subsample and full sample columns."""
    n_fit = CANON_N.copy()
    K_fit = CANON_K.copy()
    if len(full[1]) <= 1 or full[0].shape[2] == 0:
        if stats is not None:
            stats["irlp_evals"] = 0
        return n_fit, K_fit
    evals = 0
    for d in range(3):
        n_fit[d], K_fit[d], _, ev = fit_axis_cols(sub, full, d)
        evals += ev
    if stats is not None:
        stats["irlp_evals"] = evals
    return n_fit, K_fit


def fit_predictor(A, B, C, V, off, stats=None):
    """(n 3x3 int64, K 3 int64): exhaustive IRLP search on the GPU
    (encoder/_irlp_fit.py:364-394) over (n, 3) sample arrays."""
    off = np.asarray(off, dtype=np.int64)
    As, Bs, Cs, Vs, off_s = subsample(A, B, C, V, off)
    return fit_predictor_cols((as_columns(As, Bs, Cs, Vs), np.asarray(off_s, np.int64)),
                              (as_columns(A, B, C, V), off), stats)


def fit_predictor_native(nat, stats=None):
    """fit_predictor with the samples taken from the native encoder handle."""
    sel = subsample_sel(nat.n_meshlets)
    full = nat.para_columns()
    sub = full if sel is None else nat.para_columns(sel)
    return fit_predictor_cols(sub, full, stats)
