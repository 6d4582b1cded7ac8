"""ctypes binding of libstride_hip.so (declared in include/stride_hip.h).

The library is built in-tree by ``python -m meshpress_amd.build``.  There is
no fallback: if the shared object is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libstride_hip.so"

STRIDE_OK = 0
E_ARG, E_HIP = -1, -2
E_MAGIC, E_VERSION, E_TRUNCATED, E_CORRUPT, E_UNSUPPORTED = -10, -11, -12, -13, -14
E_PREP = -20

_lock = threading.Lock()
_lib = None

p = C.c_void_p
i32, i64, u32, u64, sz = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_size_t
P = C.POINTER


class StrideTune(C.Structure):
    _fields_ = [("waves", i32), ("stage_words", i32), ("grid", i32), ("use_queue", i32),
                ("mode", i32), ("group_size", i32), ("lds_bytes", i32), ("reserved", i32),
                ("tok_group", i32), ("tok_chains", i32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def make_tune(cur: dict, kw: dict) -> StrideTune:
    names = [n for n, _ in StrideTune._fields_]
    bad = set(kw) - set(names)
    if bad:
        raise ValueError(f"unknown tune fields {sorted(bad)}")
    vals = {**cur, **{k: int(v) for k, v in kw.items() if v is not None}}
    return StrideTune(*[vals[n] for n in names])


# name -> (restype, argtypes)
SIGNATURES = {
    "stride_last_error": (C.c_int, [C.c_char_p, sz]),
    "stride_init": (C.c_int, [C.c_int]),
    "stride_version": (C.c_int, []),
    "stride_debug_pyset_order": (C.c_int, [p, i32, p]),
    "stride_decoder_create": (C.c_int, [p, sz, P(p)]),
    "stride_decode_normalized": (C.c_int, [p, p, p, p]),
    "stride_decoder_format": (C.c_int, [p, P(i32), p, P(u32)]),
    "stride_decoder_set_precision": (C.c_int, [p, i32]),
    "stride_sblb_from_pdlt": (C.c_int, [p, sz, p, sz, P(sz)]),
    "stride_decoder_create_sblb": (C.c_int, [p, sz, P(p)]),
    "stride_decoder_destroy": (None, [p]),
    "stride_dprb_from_dpr4": (C.c_int, [p, sz, p, sz, P(sz)]),
    "stride_decoder_create_dprb": (C.c_int, [p, sz, P(p)]),
    "stride_decoder_sizes": (C.c_int, [p, P(u64), P(u64), P(u32)]),
    "stride_decode": (C.c_int, [p, p, p, p]),
    "stride_decode_to_host": (C.c_int, [p, p, p]),
    "stride_decoder_outputs": (C.c_int, [p, P(p), P(p)]),
    "stride_decode_noalloc": (C.c_int, [p, p, p, p]),
    "stride_decoder_bytes": (C.c_int, [p, P(u64), P(u64)]),
    "stride_decoder_header": (C.c_int, [p, p, p]),
    "stride_decoder_set_tune": (C.c_int, [p, P(StrideTune)]),
    "stride_decoder_get_tune": (C.c_int, [p, P(StrideTune)]),
    "stride_batch_create": (C.c_int, [p, p, i32, P(p)]),
    "stride_batch_create_tuned": (C.c_int, [p, p, i32, P(StrideTune), P(p)]),
    "stride_batch_destroy": (None, [p]),
    "stride_batch_outputs": (C.c_int, [p, P(p), P(p)]),
    "stride_batch_sizes": (C.c_int, [p, P(u64), P(u64), P(u32)]),
    "stride_batch_offsets": (C.c_int, [p, p, p]),
    "stride_batch_bytes": (C.c_int, [p, P(u64), P(u64)]),
    "stride_batch_set_tune": (C.c_int, [p, P(StrideTune)]),
    "stride_batch_get_tune": (C.c_int, [p, P(StrideTune)]),
    "stride_batch_decode": (C.c_int, [p, p, p, p]),
    "stride_batch_decode_graph": (C.c_int, [p, p, p, p]),
    "stride_batch_decode_per_mesh": (C.c_int, [p, p, p, p]),
    "stride_encoder_create": (C.c_int, [p, i64, p, i64, p, p, i32, i32, p, p, p, i32, i32, P(p)]),
    "stride_encoder_destroy": (None, [p]),
    "stride_encoder_counts": (C.c_int, [p, P(i64), P(i64), P(i64), P(i64), P(i64)]),
    "stride_encoder_export_plans": (C.c_int, [p, p, p, p, p, p, p, p, p, p]),
    "stride_encoder_para_samples": (C.c_int, [p, p, p, p, p, p]),
    "stride_encoder_residuals": (C.c_int, [p, p, p, P(i64), P(i64)]),
    "stride_encoder_export_residuals": (C.c_int, [p, p, p]),
    "stride_encoder_write": (C.c_int, [p, p, p, p, p, p, p, sz, P(sz)]),
    "stride_encoder_v5_prepare": (C.c_int, [p, p, p, p, P(i64), P(i64)]),
    "stride_encoder_v5_samples": (C.c_int, [p, p, p, p, p]),
    "stride_encoder_write_v5": (C.c_int, [p, p, p, sz, P(sz)]),
    "stride_select_codes": (C.c_int, [p, p, i64, p, p, p]),
    "stride_stream_costs": (C.c_int, [p, p, i64, p, p]),
    "stride_irlp_cost": (C.c_int, [p, p, p, p, p, i64, p, i32, i32, p, p]),
    "stride_irlp_cost_meshlets": (C.c_int, [p, p, p, p, p, i64, p, i32, i32, p, p]),
    "stride_encoder_features": (C.c_int, [p, i32, p, i64, p]),
    "stride_encoder_para_columns": (C.c_int, [p, p, i64, p, p]),
    "stride_mlp_sizes": (C.c_int, [P(i32), P(i32), P(i32), P(i32), P(i32)]),
    "stride_mlp_gemm": (C.c_int, [p, i32, i32, p, i32, i32, i32, i32, i32, i32, p, p, i32, p, p,
                                  i32, p]),
    "stride_mlp_forward": (C.c_int, [p, i64, p, p, p, p, p, p]),
    "stride_mlp_xent": (C.c_int, [p, p, i64, C.c_float, p, p, p]),
    "stride_mlp_backward": (C.c_int, [p, i64, p, p, p, p, p, p, p, p]),
    "stride_mlp_adam": (C.c_int, [p, p, p, p, p, i64, C.c_float, C.c_float, C.c_float, C.c_float,
                                  i32, p]),
    "stride_mlp_to_bf16": (C.c_int, [p, i64, p, p]),
    "stride_mlp_heads": (C.c_int, [p, i64, p, p, p, p]),
}


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m meshpress_amd.build`; "
                    "the STRIDE codec has no CPU fallback")
            # torch carries its own HIP runtime: load it first, so that the
            # library binds to the runtime torch's tensors and streams live
            # in (loaded the other way round, the library's first device call
            # finds no device)
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            L = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().stride_last_error(buf, 1024)
    return buf.value.decode("utf-8", "replace")


def check(rc: int) -> None:
    if rc == STRIDE_OK:
        return
    msg = last_error()
    if rc in (E_ARG, E_MAGIC, E_VERSION, E_CORRUPT, E_UNSUPPORTED):
        raise ValueError(msg)
    if rc == E_TRUNCATED:
        raise IndexError(msg)
    raise RuntimeError(msg)


def ptr(a) -> int:
    """Address of a numpy array's data or a torch tensor's storage."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def stream_ptr(stream=None):
    """hipStream_t of a torch stream (None = torch's current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
