"""DPRB v4 side-table blobs (scripts/dump_stride_dup_blob.py:88-152).

`to_dprb(dpr4_bytes)` returns the blob the reference's native decode bench
reads: the DPR4 meshlet body plus per-meshlet tables, derived on the host
by the same parse as the GPU index pre-pass.  `ParaDeltaV5GpuDecoder`
accepts either format (it dispatches on the magic) and checks a DPRB
blob's tables against its own pre-pass before decoding.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib

DPRB_MAGIC = b"DPRB"


def is_dprb(data: bytes) -> bool:
    return bytes(data[:4]) == DPRB_MAGIC


def to_dprb(data: bytes) -> bytes:
    L = _lib.lib()
    buf = np.frombuffer(bytes(data), np.uint8)
    n = C.c_size_t()
    _lib.check(L.stride_dprb_from_dpr4(_lib.ptr(buf) if len(buf) else None, len(buf), None, 0,
                                       C.byref(n)))
    out = np.zeros(int(n.value), np.uint8)
    _lib.check(L.stride_dprb_from_dpr4(_lib.ptr(buf) if len(buf) else None, len(buf),
                                       _lib.ptr(out), len(out), C.byref(n)))
    return out.tobytes()
