"""SBLB v1 blobs for PDLT v5 streams (scripts/dump_stride_blob.py:1-114).

`to_sblb(pdlt_bytes)` returns the blob the reference's native PDLT v5 decode
bench reads (bench_cpp/stride_decode_bench.cu:430-484): the stream plus the
per-meshlet tables and the float32 boundary table, derived on the host by
the decoder's own pre-pass.  `ParaDeltaV5GpuDecoder` accepts it (it
dispatches on the magic) and checks its tables against that pre-pass
before decoding.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib

SBLB_MAGIC = b"SBLB"


def is_sblb(data: bytes) -> bool:
    return bytes(data[:4]) == SBLB_MAGIC


def to_sblb(data: bytes) -> bytes:
    L = _lib.lib()
    buf = np.frombuffer(bytes(data), np.uint8)
    n = C.c_size_t()
    _lib.check(L.stride_sblb_from_pdlt(_lib.ptr(buf) if len(buf) else None, len(buf), None, 0,
                                       C.byref(n)))
    out = np.zeros(int(n.value), np.uint8)
    _lib.check(L.stride_sblb_from_pdlt(_lib.ptr(buf) if len(buf) else None, len(buf),
                                       _lib.ptr(out), len(out), C.byref(n)))
    return out.tobytes()
