"""Encoder front-end: numpy normalisation/quantisation plus the native
partition/plan/strip/walk front-end of libstride_hip.so.

Mirrors prepare_paradelta_arrays (encoder/paradelta_codec.py:758-882).  The
floating-point steps whose rounding decides later tie-breaks (centre, scale,
quantisation, face normals) use the same numpy expressions as the
reference; the combinatorial steps run in C++ (stride_encoder_create).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from functools import lru_cache

import numpy as np

from .. import _lib

# Joint-cost meshlet growth constants (utils/meshlet_gen_joint.py:29-37 and
# :259-266): feature order plane_resid, boundary_perim, strip_cont,
# normal_sim, shared_edges, bfs_depth.
LEARNED_WEIGHTS = (0.0008, 0.0008, 0.2086, 0.0008, 0.2638, 0.5250)
FEATURE_NORMS = ((1e-3, 1e-3), (1.0, 1.0), (0.7, 0.5), (0.1, 0.1), (0.6, 0.2), (5.0, 3.0))


def cost_constants():
    """(K, NM): K_i = w_i / std_i, NM_i = -mean_i."""
    K = np.array([w * (1.0 / max(1e-9, float(s))) for w, (_, s) in zip(LEARNED_WEIGHTS, FEATURE_NORMS)],
                 dtype=np.float64)
    NM = np.array([-float(m) for m, _ in FEATURE_NORMS], dtype=np.float64)
    return K, NM


@lru_cache(maxsize=8)
def strip_seed_table(max_tris: int) -> np.ndarray:
    """Row n: the 6 extra strip seeds default_rng(0).choice(n, min(6, n-2),
    replace=False) drawn for an n-triangle meshlet (utils/amd_gts_nb.py:160-165)."""
    rows = max_tris + 1
    tab = np.zeros((rows, 6), dtype=np.int32)
    for n in range(5, rows):
        s = np.random.default_rng(0).choice(n, size=min(6, n - 2), replace=False)
        tab[n, :len(s)] = s
    return tab


def quantize_global(vn: np.ndarray, per_coord_err: float):
    """(codes int64 (N,3), g_min f32, g_range f32, g_bits int32) as
    _quantize_global (encoder/paradelta_codec.py:149-165): round half to even."""
    g_min = vn.min(axis=0).astype(np.float64)
    g_range = (vn.max(axis=0) - g_min).astype(np.float64)
    step = 2.0 * per_coord_err
    g_bits = np.zeros(3, dtype=np.int32)
    codes = np.zeros((len(vn), 3), dtype=np.int64)
    for d in range(3):
        rng = float(g_range[d])
        if rng <= 0:
            g_bits[d] = 1
            continue
        nb = max(1, int(np.ceil(np.log2(rng / step + 1))))
        g_bits[d] = nb
        mx = (1 << nb) - 1
        codes[:, d] = np.round((vn[:, d] - g_min[d]) / rng * mx).clip(0, mx).astype(np.int64)
    return codes, g_min.astype(np.float32), g_range.astype(np.float32), g_bits


def face_normals(verts: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """Unit normals (utils/meshlet_generator.py:49-55)."""
    v0, v1, v2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    n = np.cross(v1 - v0, v2 - v0)
    return n / (np.linalg.norm(n, axis=1, keepdims=True) + 1e-12)


def default_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class NativePrep:
    """Handle over stride_encoder: meshlets, plans, strips and walks."""

    def __init__(self, tris, vn, fn, codes, max_verts, max_tris, n_threads=None):
        L = _lib.lib()
        self._L = L
        tris = np.ascontiguousarray(tris, dtype=np.int64)
        vn = np.ascontiguousarray(vn, dtype=np.float64)
        fn = np.ascontiguousarray(fn, dtype=np.float64)
        codes = np.ascontiguousarray(codes, dtype=np.int64)
        K, NM = cost_constants()
        tab = strip_seed_table(int(max_tris))
        h = C.c_void_p()
        _lib.check(L.stride_encoder_create(
            _lib.ptr(tris), len(tris), _lib.ptr(vn), len(vn), _lib.ptr(fn), _lib.ptr(codes),
            int(max_verts), int(max_tris), _lib.ptr(K), _lib.ptr(NM), _lib.ptr(tab), len(tab),
            int(n_threads or default_threads()), C.byref(h)))
        self._h = h
        v = [C.c_int64() for _ in range(5)]
        _lib.check(L.stride_encoder_counts(h, *[C.byref(x) for x in v]))
        (self.n_meshlets, self.n_local_total, self.n_tris_total, self.n_strips_total,
         self.n_para_total) = [int(x.value) for x in v]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.stride_encoder_destroy(h)
            self._h = None

    def local_to_global(self) -> np.ndarray:
        """Global vertex id of every decoded vertex (meshlet-major local
        order, the order of the decoder's output), int64 (n_local_total,)."""
        nm = self.n_meshlets
        bufs = [np.zeros(nm, np.int32) for _ in range(4)]
        l2g = np.zeros(self.n_local_total, np.int64)
        tl = np.zeros((self.n_tris_total, 3), np.uint8)
        slen = np.zeros(self.n_strips_total, np.int32)
        stris = np.zeros(self.n_tris_total, np.int32)
        walk = np.zeros((self.n_local_total, 5), np.uint8)
        _lib.check(self._L.stride_encoder_export_plans(
            self._h, *[_lib.ptr(a) for a in (*bufs, l2g, tl, slen, stris, walk)]))
        return l2g

    def export_plans(self):
        """Reference-shaped plans: list of dicts with local_to_global,
        ml_tris_local, strips, n_bnd, n_int, n_tris_m, n_strips."""
        nm = self.n_meshlets
        n_local = np.zeros(nm, np.int32)
        n_tris = np.zeros(nm, np.int32)
        n_strips = np.zeros(nm, np.int32)
        n_bnd = np.zeros(nm, np.int32)
        l2g = np.zeros(self.n_local_total, np.int64)
        tl = np.zeros((self.n_tris_total, 3), np.uint8)
        slen = np.zeros(self.n_strips_total, np.int32)
        stris = np.zeros(self.n_tris_total, np.int32)
        walk = np.zeros((self.n_local_total, 5), np.uint8)
        _lib.check(self._L.stride_encoder_export_plans(
            self._h, *[_lib.ptr(a) for a in (n_local, n_tris, n_strips, n_bnd, l2g, tl, slen,
                                              stris, walk)]))
        plans = []
        pl = pt = ps = 0
        for m in range(nm):
            a, b, c = int(n_local[m]), int(n_tris[m]), int(n_strips[m])
            lens = slen[ps:ps + c]
            strips, q = [], 0
            for L_ in lens:
                strips.append([int(x) for x in stris[pt + q:pt + q + int(L_)]])
                q += int(L_)
            plans.append({
                "local_to_global": [int(x) for x in l2g[pl:pl + a]],
                "ml_tris_local": tl[pt:pt + b].astype(np.int64),
                "strips": strips,
                "n_bnd": int(n_bnd[m]), "n_int": a - int(n_bnd[m]),
                "n_tris_m": b, "n_strips": c,
                "walk": walk[pl:pl + a].copy(),
            })
            pl += a
            pt += b
            ps += c
        return plans

    def para_samples(self):
        """(A, B, C, V) int32 (n_para, 3) and meshlet offsets int64 (n_m + 1)."""
        n = self.n_para_total
        A, B, Cc, V = (np.zeros((n, 3), np.int32) for _ in range(4))
        off = np.zeros(self.n_meshlets + 1, np.int64)
        _lib.check(self._L.stride_encoder_para_samples(
            self._h, _lib.ptr(A), _lib.ptr(B), _lib.ptr(Cc), _lib.ptr(V), _lib.ptr(off)))
        return A, B, Cc, V, off

    def residuals(self, pred_n, pred_K, export: bool = True):
        """Computes the residual streams for (pred_n, pred_K); returns (int32
        values or None when export is False, int64 stream offsets (6 n_m + 1),
        max |residual|)."""
        pn = np.ascontiguousarray(np.asarray(pred_n, dtype=np.int32).reshape(9))
        pk = np.ascontiguousarray(np.asarray(pred_K, dtype=np.int32).reshape(3))
        nv, mx = C.c_int64(), C.c_int64()
        _lib.check(self._L.stride_encoder_residuals(self._h, _lib.ptr(pn), _lib.ptr(pk),
                                                    C.byref(nv), C.byref(mx)))
        self._n_vals = int(nv.value)
        vals = self.residual_values() if export else None
        off = np.zeros(6 * self.n_meshlets + 1, np.int64)
        _lib.check(self._L.stride_encoder_export_residuals(self._h, None, _lib.ptr(off)))
        return vals, off, int(mx.value)

    def residual_values(self) -> np.ndarray:
        """int32 values of the residual streams of the last residuals() call."""
        vals = np.zeros(self._n_vals, np.int32)
        _lib.check(self._L.stride_encoder_export_residuals(self._h, _lib.ptr(vals), None))
        return vals

    def features(self, source: int = 0, sel=None) -> np.ndarray:
        """Proposal features, float32 (rows, 32), of the meshlets `sel` (all
        when None) (stride_encoder_features): stream statistics of the
        canonical predictor (source 0) or of the last residuals() call
        (source 1)."""
        s = None if sel is None else np.ascontiguousarray(sel, np.int32)
        rows = self.n_meshlets if s is None else len(s)
        out = np.zeros((rows, 32), np.float32)
        _lib.check(self._L.stride_encoder_features(self._h, int(source), _lib.ptr(s), rows,
                                                   _lib.ptr(out)))
        return out

    def para_columns(self, sel=None):
        """IRLP fit samples of the meshlets `sel` (all when None) as int32
        columns (4, 3, n) -- a, b, c, v by axis -- and int64 row offsets."""
        s = None if sel is None else np.ascontiguousarray(sel, np.int32)
        rows = self.n_meshlets if s is None else len(s)
        off = np.zeros(rows + 1, np.int64)
        _lib.check(self._L.stride_encoder_para_columns(self._h, _lib.ptr(s), rows, None,
                                                       _lib.ptr(off)))
        cols = np.zeros((4, 3, int(off[-1])), np.int32)
        _lib.check(self._L.stride_encoder_para_columns(self._h, _lib.ptr(s), rows, _lib.ptr(cols),
                                                       _lib.ptr(off)))
        return cols, off

    def v5_prepare(self, vn, f13, g_bits):
        """Plain-parallelogram pass of the PDLT v5 writer; (n_boundary, n_samples)."""
        vn = np.ascontiguousarray(vn, dtype=np.float64)
        f13 = np.ascontiguousarray(np.asarray(f13, dtype=np.float32).reshape(13))
        gb = np.ascontiguousarray(np.asarray(g_bits, dtype=np.int32).reshape(3))
        nb, ns = C.c_int64(), C.c_int64()
        _lib.check(self._L.stride_encoder_v5_prepare(self._h, _lib.ptr(vn), _lib.ptr(f13),
                                                     _lib.ptr(gb), C.byref(nb), C.byref(ns)))
        self._v5_samples = int(ns.value)
        return int(nb.value), int(ns.value)

    def v5_samples(self):
        """LIN3 fit samples: a, b, c reconstructions and true positions,
        float64 (n_samples, 3) each."""
        out = [np.zeros((self._v5_samples, 3), np.float64) for _ in range(4)]
        _lib.check(self._L.stride_encoder_v5_samples(self._h, *[_lib.ptr(a) for a in out]))
        return out

    def write_v5(self, lin3) -> bytes:
        w = np.ascontiguousarray(np.asarray(lin3, dtype=np.float32).reshape(3))
        n = C.c_size_t()
        _lib.check(self._L.stride_encoder_write_v5(self._h, _lib.ptr(w), None, 0, C.byref(n)))
        out = np.zeros(int(n.value), np.uint8)
        _lib.check(self._L.stride_encoder_write_v5(self._h, _lib.ptr(w), _lib.ptr(out), len(out),
                                                   C.byref(n)))
        return out.tobytes()

    def write(self, f13, g_bits, pred_n, pred_K, tagk) -> bytes:
        f13 = np.ascontiguousarray(np.asarray(f13, dtype=np.float32).reshape(13))
        gb = np.ascontiguousarray(np.asarray(g_bits, dtype=np.int32).reshape(3))
        pn = np.ascontiguousarray(np.asarray(pred_n, dtype=np.int32).reshape(9))
        pk = np.ascontiguousarray(np.asarray(pred_K, dtype=np.int32).reshape(3))
        tk = np.ascontiguousarray(np.asarray(tagk, dtype=np.uint8).reshape(-1))
        n = C.c_size_t()
        _lib.check(self._L.stride_encoder_write(self._h, _lib.ptr(f13), _lib.ptr(gb), _lib.ptr(pn),
                                                _lib.ptr(pk), _lib.ptr(tk), None, 0, C.byref(n)))
        out = np.zeros(int(n.value), np.uint8)
        _lib.check(self._L.stride_encoder_write(self._h, _lib.ptr(f13), _lib.ptr(gb), _lib.ptr(pn),
                                                _lib.ptr(pk), _lib.ptr(tk), _lib.ptr(out), len(out),
                                                C.byref(n)))
        return out.tobytes()


def prepare(verts, tris, *, max_verts=256, max_tris=256, precision_error=0.0005,
            precision_mode="world", gen_method="joint_learned", strip_method="multiseed",
            n_threads=None) -> dict:
    """Prep dict with the reference's field names plus `native` (NativePrep)."""
    if gen_method != "joint_learned":
        raise ValueError(f"gen_method={gen_method!r} is not supported; only 'joint_learned'")
    if strip_method != "multiseed":
        raise ValueError(f"strip_method={strip_method!r} is not supported; only 'multiseed'")
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int64)
    n_v, n_t = len(verts), len(tris)
    if n_v == 0 or n_t == 0:
        raise ValueError("mesh has no vertices or no triangles")
    center = verts.mean(axis=0)
    vc = verts - center
    scale = float(np.max(np.linalg.norm(vc, axis=1)))
    vn = vc / scale
    if precision_mode == "world":
        per_coord_err = precision_error / scale / math.sqrt(3)
    elif precision_mode == "bbox_frac":
        bbox_extent = float((verts.max(0) - verts.min(0)).max())
        per_coord_err = (precision_error * bbox_extent) / (2.0 * scale)
    else:
        raise ValueError(f"unknown precision_mode={precision_mode!r}; "
                         "expected 'world' or 'bbox_frac'")
    codes, g_min, g_range, g_bits = quantize_global(vn, per_coord_err)
    fn = face_normals(vn, tris)
    native = NativePrep(tris, vn, fn, codes, max_verts, max_tris, n_threads)
    return {
        "center": center, "scale": scale, "per_coord_err": per_coord_err,
        "g_min": g_min, "g_range": g_range, "g_bits": g_bits,
        "n_v": n_v, "n_t": n_t, "n_meshlets": native.n_meshlets,
        "global_codes": codes, "vn": vn, "native": native,
        "max_verts": max_verts, "max_tris": max_tris, "precision_error": precision_error,
    }


def header_f32(prep) -> np.ndarray:
    """center[3], scale, per_coord_err, g_min[3], g_range[3] as written (f32)."""
    f = np.zeros(13, dtype=np.float32)
    f[0:3] = np.asarray(prep["center"], dtype=np.float64).astype(np.float32)
    f[3] = np.float32(prep["scale"])
    f[4] = np.float32(prep["per_coord_err"])
    f[5:8] = prep["g_min"]
    f[8:11] = prep["g_range"]
    return f
