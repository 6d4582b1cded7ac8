"""STRIDE encoder surface (reference: encoder/implementation/stride.py:18-53).

`STRIDEEncoder(**kw).encode(model) -> CompressedModel` with the reference's
keywords and defaults.  The front-end runs natively (prep.py); the coder
search runs on the GPU:

  selector="exhaustive"       IRLP grid search + per-stream (tag, k) scoring
                              on the GPU (the reference's exhaustive search);
  selector="learned"          the learned proposal (learned/proposal.py)
                              picks the predictor and coders, scored on
                              the GPU (packaged weights unless `proposal`
                              is given);
  selector="learned+fallback" as "learned", with exhaustive search wherever
                              the proposal's confidence is low.

Output is always a valid DPR4 v3 stream; bits per vertex/triangle use the
reference's formulas (len(data) * 8 / n_v, / n_t).
"""
from __future__ import annotations

import time

import numpy as np

from ..utils.types import CompressedModel, Model
from . import search
from .prep import header_f32, prepare

SELECTORS = ("exhaustive", "learned", "learned+fallback")


class STRIDEEncoder:
    def __init__(self, *, max_verts: int = 256, max_tris: int = 256,
                 precision_error: float = 0.0005, precision_mode: str = "world",
                 gen_method: str = "joint_learned", strip_method: str = "multiseed",
                 predictor: str = "generalized", verbose: bool = False,
                 selector: str = "exhaustive", proposal=None, n_threads=None):
        if selector not in SELECTORS:
            raise ValueError(f"selector must be one of {SELECTORS}, got {selector!r}")
        self.max_verts = max_verts
        self.max_tris = max_tris
        self.precision_error = precision_error
        self.precision_mode = precision_mode
        self.gen_method = gen_method
        self.strip_method = strip_method
        self.predictor = predictor
        self.verbose = verbose
        self.selector = selector
        self.proposal = proposal
        self.n_threads = n_threads
        self.last_stats: dict = {}

    def prepare(self, model: Model) -> dict:
        return prepare(model.vertices_np, model.triangles_np, max_verts=self.max_verts,
                       max_tris=self.max_tris, precision_error=self.precision_error,
                       precision_mode=self.precision_mode, gen_method=self.gen_method,
                       strip_method=self.strip_method, n_threads=self.n_threads)

    def encode_prepared(self, prep: dict) -> bytes:
        """encode_dup (encoder/paradelta_v5_dup.py:316-464) for a prep dict."""
        if self.predictor not in ("generalized", "canonical"):
            raise ValueError(f"predictor must be 'generalized' or 'canonical', got {self.predictor!r}")
        st = self.last_stats = {}
        nat = prep["native"]
        learned = None
        if self.selector != "exhaustive":
            if self.proposal is None:
                from ..learned.proposal import LearnedProposal
                self.proposal = LearnedProposal()
            learned = self.proposal
        fb = self.selector == "learned+fallback"
        # search time = predictor choice + coder choice; the residual pass
        # between them is the same work for every selector and is timed apart
        t0 = time.perf_counter()
        if self.predictor == "canonical":
            pred_n, pred_K = search.CANON_N.copy(), search.CANON_K.copy()
        elif learned is not None:
            pred_n, pred_K = learned.fit_predictor(nat, stats=st, fallback=fb)
        else:
            pred_n, pred_K = search.fit_predictor_native(nat, stats=st)
        t1 = time.perf_counter()
        _, soff, max_abs = nat.residuals(pred_n, pred_K, export=False)
        t2 = time.perf_counter()
        if learned is not None:
            tagk = learned.select_codes(nat, soff, stats=st, fallback=fb)
        else:
            tagk, _ = search.select_codes(nat.residual_values(), soff)
        t3 = time.perf_counter()
        data = nat.write(header_f32(prep), prep["g_bits"], pred_n, pred_K,
                         tagk.reshape(prep["n_meshlets"], 12))
        t4 = time.perf_counter()
        st.update(predictor_s=t1 - t0, residual_s=t2 - t1, coder_s=t3 - t2, write_s=t4 - t3,
                  search_s=(t1 - t0) + (t3 - t2),
                  pred_n=pred_n.tolist(), pred_K=pred_K.tolist(), max_abs=max_abs)
        if self.verbose:
            print(f"  dup: n_v_src={prep['n_v']:,} n_meshlets={prep['n_meshlets']:,} "
                  f"max|code|={max_abs:,} BPV(src)={8 * len(data) / prep['n_v']:.2f}")
        return data

    def encode(self, model: Model) -> CompressedModel:
        prep = self.prepare(model)
        data = self.encode_prepared(prep)
        n_v, n_t = int(prep["n_v"]), int(prep["n_t"])
        return CompressedModel(data=data, bits_per_vertex=len(data) * 8 / n_v,
                               bits_per_triangle=len(data) * 8 / n_t)
