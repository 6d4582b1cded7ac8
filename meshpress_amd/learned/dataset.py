"""Training labels for the proposal network, produced by the exhaustive
search on the GPU.

Per meshlet:
  features  float32 (32,)         stride_encoder_features
  labels    uint8 (9,)            three predictor heads, six coder heads
                                  (IGNORE where the head has nothing to code)

Predictor head of axis d: the index j of the predictor variant
VARIANTS[j] (shift VARIANT_K) with the fewest parallelogram-residual bits
on that meshlet (minimum over the 20 coders, the cost of
encoder/_irlp_fit.py:120-183); lowest j on ties.  Variant j is the
parallelogram shrunk towards the third vertex,
    pred = c + s (a + b - 2c),  s = 1 - j/32,
with integer weights (256 s, 256 s, 256 - 512 s) / 256, so j = 0 is the
canonical a + b - c and the weights always sum to 2^K.

Coder head of stream s: the class of the (tag, k) that the exhaustive
choice (_pick_best_k, paradelta_v5_dup.py:159-171) picks for the residuals
under the exhaustively fitted predictor: k for Rice(k), 12 + k for
Exp-Golomb(k).
"""
from __future__ import annotations

import numpy as np

from ..encoder import search
from .mlp import IGNORE, N_HEADS, PRED_CLASSES

VARIANT_K = 8
VARIANTS = np.array([[256 - 8 * j, 256 - 8 * j, 16 * j - 256] for j in range(PRED_CLASSES)],
                    dtype=np.int32)


def coder_class(tagk: np.ndarray) -> np.ndarray:
    """(tag, k) rows -> class index (Rice k -> k, Exp-Golomb k -> 12 + k)."""
    tagk = np.asarray(tagk).reshape(-1, 2)
    return np.where(tagk[:, 0] == 1, tagk[:, 1], 12 + tagk[:, 1]).astype(np.uint8)


def class_tagk(cls: np.ndarray) -> np.ndarray:
    """Inverse of coder_class: (n, 2) uint8 (tag, k)."""
    cls = np.asarray(cls, dtype=np.int64).reshape(-1)
    out = np.empty((len(cls), 2), np.uint8)
    out[:, 0] = np.where(cls < 12, 1, 2)
    out[:, 1] = np.where(cls < 12, cls, cls - 12)
    return out


def variant_costs(cols, off) -> np.ndarray:
    """int64 (n_meshlets, 3, PRED_CLASSES): per-meshlet bits of every variant;
    cols (4, 3, n) sample columns (a, b, c, v by axis), off meshlet offsets."""
    torch = search._torch()
    from .. import _lib
    n_m = len(off) - 1
    out = np.zeros((n_m, 3, PRED_CLASSES), np.int64)
    if n_m == 0:
        return out
    d_off = search._dev(np.asarray(off, np.int64), torch.int64)
    d_c = search._dev(VARIANTS, torch.int32)
    d_out = torch.empty((n_m, PRED_CLASSES), dtype=torch.int64, device="cuda")
    z = np.zeros(1, np.int32)
    for d in range(3):
        dcols = [search._dev(cols[q, d] if cols.shape[2] else z, torch.int32) for q in range(4)]
        _lib.check(_lib.lib().stride_irlp_cost_meshlets(
            *[_lib.ptr(c) for c in dcols], _lib.ptr(d_off), n_m, _lib.ptr(d_c), PRED_CLASSES,
            VARIANT_K, _lib.ptr(d_out), _lib.stream_ptr()))
        out[:, d, :] = d_out.cpu().numpy()
    return out


def label_prep(prep: dict, stats: dict | None = None):
    """(features (2 n_m, 32) float32, labels (2 n_m, 9) uint8) for one
    prepared mesh: the first n_m rows carry the canonical-pass features with
    the predictor labels (coder heads ignored), the next n_m rows the
    features of the residuals under the exhaustively fitted predictor with
    the coder labels (predictor heads ignored) -- the two passes the
    proposal runs the network on."""
    nat = prep["native"]
    n_m = nat.n_meshlets
    F0 = nat.features(0)
    cols, off = nat.para_columns()
    n_para = np.diff(off)
    Y0 = np.full((n_m, N_HEADS), IGNORE, np.uint8)
    vc = variant_costs(cols, off)
    best = vc.argmin(axis=2).astype(np.uint8)          # first minimum = lowest j
    has = n_para > 0
    Y0[has, 0:3] = best[has]
    pred_n, pred_K = search.fit_predictor_native(nat)
    vals, soff, _ = nat.residuals(pred_n, pred_K)
    F1 = nat.features(1)
    tagk, _ = search.select_codes(vals, soff)
    cls = coder_class(tagk).reshape(n_m, 6)
    lens = np.diff(soff).reshape(n_m, 6)
    Y1 = np.full((n_m, N_HEADS), IGNORE, np.uint8)
    Y1[:, 3:9] = np.where(lens > 0, cls, IGNORE)
    if stats is not None:
        stats.update(pred_n=pred_n.tolist(), pred_K=pred_K.tolist(), n_meshlets=int(n_m))
    return np.concatenate([F0, F1]), np.concatenate([Y0, Y1])
