"""Encode many meshes at once: the CPU front-end of every mesh runs on a
pool of host threads (the native prep and writer release the GIL), the GPU
search calls of all threads share the device.

    blobs, stats = encode_many(sources, n_workers=16, precision_error=...,
                               precision_mode="bbox_frac")

`sources` holds (verts, tris) pairs or zero-argument callables returning
one, so mesh generation also runs on the pool.  Each mesh is encoded by its
own STRIDEEncoder(**encoder_kw) with one native thread, and its result is
exactly what `STRIDEEncoder(**encoder_kw).encode(model)` returns for that
mesh alone.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

from ..utils.types import Model
from .prep import default_threads
from .stride import STRIDEEncoder


def _encode_one(src, encoder_kw):
    t0 = time.perf_counter()
    V, T = src() if callable(src) else src
    t1 = time.perf_counter()
    enc = STRIDEEncoder(**{**encoder_kw, "n_threads": 1})
    prep = enc.prepare(Model.from_arrays(V, T))
    t2 = time.perf_counter()
    data = enc.encode_prepared(prep)
    t3 = time.perf_counter()
    st = {"n_v": int(prep["n_v"]), "n_t": int(prep["n_t"]), "n_meshlets": int(prep["n_meshlets"]),
          "bytes": len(data), "gen_s": t1 - t0, "prep_s": t2 - t1, "search_write_s": t3 - t2,
          "g_bits": [int(x) for x in prep["g_bits"]]}
    st.update({k: v for k, v in enc.last_stats.items() if k in ("pred_n", "pred_K", "max_abs")})
    return data, st


def encode_many(sources, *, n_workers=None, progress=None, **encoder_kw):
    """([bytes], [per-mesh stats]) in the order of `sources`."""
    sources = list(sources)
    n_workers = max(1, int(n_workers or default_threads()))
    out = [None] * len(sources)
    with ThreadPoolExecutor(max_workers=n_workers) as pool:
        futs = [pool.submit(_encode_one, s, encoder_kw) for s in sources]
        for i, f in enumerate(futs):
            out[i] = f.result()
            if progress is not None:
                progress(i + 1, len(sources))
    return [o[0] for o in out], [o[1] for o in out]
