"""ctypes wrapper of oracle/liboracle.so (the C oracle; test infrastructure)."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "oracle" / "liboracle.so"
_lib = None


class COracle:
    def __init__(self, L):
        self.L = L

    def sizes(self, data: bytes):
        nm, nv, nt = C.c_uint32(), C.c_uint64(), C.c_uint64()
        buf = np.frombuffer(data, np.uint8)
        rc = self.L.oracle_sizes(buf.ctypes.data, len(buf), C.byref(nm), C.byref(nv), C.byref(nt))
        if rc:
            raise ValueError(f"oracle_sizes failed ({rc})")
        return int(nm.value), int(nv.value), int(nt.value)

    def decode(self, data: bytes, n_threads: int = 1, with_codes: bool = False):
        nm, nv, nt = self.sizes(data)
        buf = np.frombuffer(data, np.uint8)
        V = np.zeros((nv, 3), np.float32)
        T = np.zeros((nt, 3), np.int64)
        codes = np.zeros((nv, 3), np.int64) if with_codes else None
        rc = self.L.oracle_decode(buf.ctypes.data, len(buf), V.ctypes.data, T.ctypes.data,
                                  codes.ctypes.data if codes is not None else None, n_threads)
        if rc:
            raise ValueError(f"oracle_decode failed ({rc})")
        return (V, T, codes) if with_codes else (V, T)

    def stream_costs(self, vals) -> np.ndarray:
        v = np.ascontiguousarray(vals, dtype=np.int64)
        out = np.zeros(20, np.int64)
        self.L.oracle_stream_costs(v.ctypes.data, len(v), out.ctypes.data)
        return out

    def pick(self, vals):
        v = np.ascontiguousarray(vals, dtype=np.int64)
        tag, k, bits = C.c_int(), C.c_int(), C.c_int64()
        self.L.oracle_pick(v.ctypes.data, len(v), C.byref(tag), C.byref(k), C.byref(bits))
        return tag.value, k.value, bits.value

    def irlp_cost(self, a, b, c, v, off, cands, K) -> np.ndarray:
        a, b, c, v = (np.ascontiguousarray(x, dtype=np.int64) for x in (a, b, c, v))
        off = np.ascontiguousarray(off, dtype=np.int64)
        cands = np.ascontiguousarray(np.asarray(cands).reshape(-1, 3), dtype=np.int64)
        out = np.zeros(len(cands), np.int64)
        self.L.oracle_irlp_cost(a.ctypes.data, b.ctypes.data, c.ctypes.data, v.ctypes.data,
                                off.ctypes.data, len(off) - 1, cands.ctypes.data, len(cands),
                                int(K), out.ctypes.data)
        return out


    def irlp_cost_mt(self, a, b, c, v, off, cands, K, n_threads) -> np.ndarray:
        """irlp_cost over n_threads host threads (int64 inputs, no copies
        when they are already contiguous int64)."""
        a, b, c, v = (np.ascontiguousarray(x, dtype=np.int64) for x in (a, b, c, v))
        off = np.ascontiguousarray(off, dtype=np.int64)
        cands = np.ascontiguousarray(np.asarray(cands).reshape(-1, 3), dtype=np.int64)
        out = np.zeros(len(cands), np.int64)
        self.L.oracle_irlp_cost_mt(a.ctypes.data, b.ctypes.data, c.ctypes.data, v.ctypes.data,
                                   off.ctypes.data, len(off) - 1, cands.ctypes.data, len(cands),
                                   int(K), out.ctypes.data, int(n_threads))
        return out

    def pick_streams(self, vals, off, n_threads):
        """(tag, k) per stream (n_streams x 2 uint8) and body bits."""
        v = np.ascontiguousarray(vals, dtype=np.int64)
        off = np.ascontiguousarray(off, dtype=np.int64)
        n = len(off) - 1
        tagk = np.zeros((n, 2), np.uint8)
        bits = np.zeros(n, np.int64)
        self.L.oracle_pick_streams(v.ctypes.data, off.ctypes.data, n, tagk.ctypes.data,
                                   bits.ctypes.data, int(n_threads))
        return tagk, bits


def build():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


def load() -> COracle:
    global _lib
    if _lib is None:
        if not SO.exists():
            build()
        L = C.CDLL(str(SO))
        vp = C.c_void_p
        L.oracle_sizes.argtypes = [vp, C.c_size_t, vp, vp, vp]
        L.oracle_decode.argtypes = [vp, C.c_size_t, vp, vp, vp, C.c_int]
        L.oracle_stream_costs.argtypes = [vp, C.c_int64, vp]
        L.oracle_stream_costs.restype = None
        L.oracle_pick.argtypes = [vp, C.c_int64, vp, vp, vp]
        L.oracle_pick.restype = None
        L.oracle_irlp_cost.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp, C.c_int, C.c_int, vp]
        L.oracle_irlp_cost.restype = None
        L.oracle_irlp_cost_mt.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp, C.c_int, C.c_int, vp,
                                          C.c_int]
        L.oracle_irlp_cost_mt.restype = None
        L.oracle_pick_streams.argtypes = [vp, vp, C.c_int64, vp, vp, C.c_int]
        L.oracle_pick_streams.restype = None
        _lib = L
    return COracle(_lib)
