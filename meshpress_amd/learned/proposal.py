"""Encoder proposals: the learned one (MLP on the GPU, scored on the GPU,
exhaustive fallback) and the network-free baseline it is compared with.

LearnedProposal replaces the two searches of encode_dup:

* predictor (encoder/_irlp_fit.py:364-394): the network runs on the
  canonical-pass features; per axis its per-meshlet class probabilities
  are summed with weights n_para; the canonical variant and the `top`
  best-voted variants are scored exactly by the IRLP cost kernel on the
  parallelogram samples of the meshlets the exhaustive fit would search
  (its 4,000-meshlet subsample), the best one is refined by
  `refine_rounds` rounds of +-1 moves on its three weights, and with
  fallback an axis whose vote share is below `tau_pred` also runs the
  exhaustive per-axis search, keeping whichever costs fewer bits over every
  meshlet;
* coders (paradelta_v5_dup.py:159-171): the network runs again on the
  features of the residuals under the chosen predictor and each residual
  stream takes its arg-max (tag, k); with fallback a stream whose class
  probability is below `tau_coder` is scored exhaustively instead.  The
  residual values themselves never leave the native encoder unless a
  stream falls back.

Every choice is a valid DPR4 v3 parameter, so the decoder is unchanged.
"""
from __future__ import annotations

import time
from pathlib import Path

import numpy as np

from ..encoder import search
from .dataset import VARIANT_K, VARIANTS, class_tagk
from .mlp import PRED_CLASSES, ProposalMLP

WEIGHTS = Path(__file__).resolve().parent / "weights" / "proposal.npz"
MOVES = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))


def gather_streams(vals: np.ndarray, soff: np.ndarray, idx: np.ndarray):
    """Values and offsets of the streams `idx` (in that order)."""
    soff = np.asarray(soff, np.int64)
    lens = soff[idx + 1] - soff[idx]
    new_off = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    pos = np.repeat(soff[idx] - new_off[:-1], lens) + np.arange(new_off[-1], dtype=np.int64)
    return vals[pos], new_off


class LearnedProposal:
    def __init__(self, mlp: ProposalMLP | None = None, weights=None, top: int = 2,
                 refine_rounds: int = 1, tau_pred: float = 0.2, tau_coder: float = 0.9):
        if mlp is None:
            path = Path(weights) if weights is not None else WEIGHTS
            if not path.exists():
                raise RuntimeError(f"proposal weights {path} are missing (train them with "
                                   "python -m meshpress_amd.learned.train)")
            mlp = ProposalMLP.load(path)
        self.mlp = mlp
        self.top = int(top)
        self.refine_rounds = int(refine_rounds)
        self.tau_pred = float(tau_pred)
        self.tau_coder = float(tau_coder)

    # ------------------------------------------------------------ predictor
    def fit_predictor(self, nat, stats=None, fallback=False):
        """(pred_n 3x3, pred_K 3) for the mesh behind the native handle.
        Votes and candidate scores use the meshlets of the IRLP fit's
        subsample (every meshlet of a mesh with at most 4,000)."""
        t0 = time.perf_counter()
        sel = search.subsample_sel(nat.n_meshlets)
        cols, off = nat.para_columns(sel)
        pp, _, _ = self.mlp.infer(nat.features(0, sel), want=("pred",))
        w = np.diff(off).astype(np.float64)
        pred_n = search.CANON_N.copy()
        pred_K = search.CANON_K.copy()
        evals, conf, fb = 0, [], []
        full = None
        if cols.shape[2] and w.sum() > 0:
            # one upload of the three axes' samples; each stage (the
            # candidates of every axis, then each refinement round) is one
            # candidate upload, one launch per axis and one read-back
            cost = search.IrlpCostAxes(cols, off)
            S, order, cand = [], [], []
            for d in range(3):
                S.append(w @ pp[:, PRED_CLASSES * d:PRED_CLASSES * (d + 1)].astype(np.float64))
                order.append(np.argsort(-S[d], kind="stable"))
                cand.append(sorted({0, *order[d][:self.top].tolist()}))
            ccs = cost.run([(d, VARIANTS[cand[d]]) for d in range(3)], VARIANT_K)
            best_n, best_c, canonical = [], [], []
            for d in range(3):
                i = int(np.argmin(ccs[d]))
                best_n.append(VARIANTS[cand[d][i]].astype(np.int64))
                best_c.append(int(ccs[d][i]))
                canonical.append(cand[d][i] == 0)
            best_K = VARIANT_K
            active = [True, True, True]
            for _ in range(self.refine_rounds):
                axes = [d for d in range(3) if active[d]]
                if not axes:
                    break
                mvs = []
                for d in axes:
                    mv = np.repeat(best_n[d][None], len(MOVES), axis=0)
                    for j, (k, sg) in enumerate(MOVES):
                        mv[j, k] += sg
                    mvs.append(mv)
                mcs = cost.run(list(zip(axes, mvs)), VARIANT_K)
                for d, mv, mc in zip(axes, mvs, mcs):
                    j = int(np.argmin(mc))
                    if int(mc[j]) >= best_c[d]:
                        active[d] = False
                        continue
                    best_n[d], best_c[d], canonical[d] = mv[j].copy(), int(mc[j]), False
            evals += cost.evals
            for d in range(3):
                bn, bK, canon = best_n[d], best_K, canonical[d]
                share = float(S[d][order[d][0]] / S[d].sum()) if S[d].sum() > 0 else 1.0
                conf.append(share)
                if fallback and share < self.tau_pred:
                    # the exhaustive per-axis search; both results are then
                    # compared over every meshlet
                    if full is None:
                        full = nat.para_columns() if sel is not None else (cols, off)
                    n_e, K_e, c_e, ev = search.fit_axis_cols((cols, off), full, d)
                    fcost = search.IrlpCost(full[0][0, d], full[0][1, d], full[0][2, d],
                                            full[0][3, d], full[1])
                    c_l = int(fcost(bn[None], int(bK))[0])
                    evals += ev + fcost.evals
                    fb.append(d)
                    if c_e <= c_l:
                        bn, bK = n_e, K_e
                        canon = K_e == 0 and np.array_equal(n_e, search.CANON_N[d])
                if canon:
                    pred_n[d], pred_K[d] = search.CANON_N[d], 0
                else:
                    pred_n[d], pred_K[d] = bn, bK
        if stats is not None:
            stats.update(irlp_evals=evals, pred_confidence=conf, pred_fallback_axes=fb,
                         proposal_predictor_s=time.perf_counter() - t0)
        return pred_n, pred_K

    # ------------------------------------------------------------ coders
    def select_codes(self, nat, soff, stats=None, fallback=False):
        """(tag, k) per residual stream of the last residuals() call."""
        t0 = time.perf_counter()
        _, cls, cp = self.mlp.infer(nat.features(1), want=("cls", "cp") if fallback else ("cls",))
        soff = np.asarray(soff, np.int64)
        lens = np.diff(soff)
        tagk = class_tagk(cls.reshape(-1))
        tagk[lens == 0] = (1, 0)
        n_fb = 0
        if fallback:
            idx = np.nonzero((cp.reshape(-1) < self.tau_coder) & (lens > 0))[0]
            n_fb = len(idx)
            if n_fb:
                sv, so = gather_streams(nat.residual_values(), soff, idx)
                tk, _ = search.select_codes(sv, so)
                tagk[idx] = tk
        if stats is not None:
            stats.update(coder_fallback_streams=int(n_fb), coder_streams=int((lens > 0).sum()),
                         proposal_coder_s=time.perf_counter() - t0)
        return tagk


class BaselineProposal:
    """The method without the network: canonical predictor, and Rice(k) with
    k = floor(log2(mean u)) of each stream's zigzag values (the closed-form
    Rice parameter estimate), read from the same features."""

    def fit_predictor(self, nat, stats=None, fallback=False):
        return search.CANON_N.copy(), search.CANON_K.copy()

    def select_codes(self, nat, soff, stats=None, fallback=False):
        F = nat.features(1)
        mean_u = np.exp2(F[:, 23:29].astype(np.float64) * 16.0) - 1.0
        k = np.floor(np.log2(np.maximum(mean_u, 1.0))).astype(np.int64).clip(0, 11)
        tagk = np.zeros((k.size, 2), np.uint8)
        tagk[:, 0] = 1
        tagk[:, 1] = k.reshape(-1)
        lens = np.diff(np.asarray(soff, np.int64))
        tagk[lens == 0] = (1, 0)
        return tagk
