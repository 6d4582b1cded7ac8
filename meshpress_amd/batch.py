"""Batched decode of many DPR4 v3 streams (stride_batch_* in libstride_hip.so).

All streams share one device buffer and one global meshlet table; a decode
is a single launch (optionally a persistent grid pulling meshlets from a
device work queue) or a replay of a captured hipGraph.  Outputs are the
concatenation of every stream's decode_dup result; each mesh's indices stay
relative to its own vertex array, so `split()` returns exactly what decoding
each stream alone returns.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class BatchDecoder:
    def __init__(self, blobs):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("BatchDecoder needs a ROCm GPU (torch.cuda is unavailable)")
        L = _lib.lib()
        self._L = L
        self.blobs = [bytes(b) for b in blobs]
        n = len(self.blobs)
        if n == 0:
            raise ValueError("need at least one stream")
        arrs = [np.frombuffer(b, dtype=np.uint8) for b in self.blobs]
        ptrs = (C.c_void_p * n)(*[a.ctypes.data if len(a) else None for a in arrs])
        lens = (C.c_size_t * n)(*[len(a) for a in arrs])
        self.device = torch.device("cuda", torch.cuda.current_device())
        _lib.check(L.stride_init(self.device.index))
        h = C.c_void_p()
        _lib.check(L.stride_batch_create(ptrs, lens, n, C.byref(h)))
        self._h = h
        nv, nt, nm = C.c_uint64(), C.c_uint64(), C.c_uint32()
        _lib.check(L.stride_batch_sizes(h, C.byref(nv), C.byref(nt), C.byref(nm)))
        self.n_meshes = n
        self.n_v_total, self.n_t_total, self.n_meshlets = int(nv.value), int(nt.value), int(nm.value)
        self.v_base = np.zeros(n + 1, np.uint64)
        self.t_base = np.zeros(n + 1, np.uint64)
        _lib.check(L.stride_batch_offsets(h, _lib.ptr(self.v_base), _lib.ptr(self.t_base)))
        sb, side = C.c_uint64(), C.c_uint64()
        _lib.check(L.stride_batch_bytes(h, C.byref(sb), C.byref(side)))
        self.stream_bytes, self.side_bytes = int(sb.value), int(side.value)
        self.d_verts = torch.empty((max(self.n_v_total, 1), 3), dtype=torch.float32, device=self.device)
        self.d_tris = torch.empty((max(self.n_t_total, 1), 3), dtype=torch.int32, device=self.device)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.stride_batch_destroy(h)
            self._h = None

    def get_tune(self) -> dict:
        t = _lib.StrideTune()
        _lib.check(self._L.stride_batch_get_tune(self._h, C.byref(t)))
        return t.as_dict()

    def set_tune(self, **kw) -> None:
        t = _lib.make_tune(self.get_tune(), kw)
        _lib.check(self._L.stride_batch_set_tune(self._h, C.byref(t)))

    def decode(self, mode: str = "launch", stream=None):
        """mode: "launch" (one launch), "graph" (hipGraph replay) or
        "per_mesh" (one launch per stream, the baseline)."""
        fn = {"launch": self._L.stride_batch_decode,
              "graph": self._L.stride_batch_decode_graph,
              "per_mesh": self._L.stride_batch_decode_per_mesh}[mode]
        _lib.check(fn(self._h, _lib.ptr(self.d_verts), _lib.ptr(self.d_tris), _lib.stream_ptr(stream)))
        return self.d_verts, self.d_tris

    def split(self, verts, tris):
        """Per-mesh (verts, tris) views of concatenated outputs."""
        out = []
        for i in range(self.n_meshes):
            v0, v1 = int(self.v_base[i]), int(self.v_base[i + 1])
            t0, t1 = int(self.t_base[i]), int(self.t_base[i + 1])
            out.append((verts[v0:v1], tris[t0:t1]))
        return out

    @property
    def bytes_read(self) -> int:
        return self.stream_bytes + self.side_bytes

    @property
    def bytes_written(self) -> int:
        return 12 * self.n_v_total + 12 * self.n_t_total
