"""Mesh container and compressed result (reference: utils/types.py:140-235).

Only the numpy-backed storage the STRIDE path uses is mirrored: a Model
holds float64 (N, 3) positions and int64 (M, 3) triangles; CompressedModel
is the encoder's return value.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


class Model:
    """Triangle mesh: `vertices_np` float64 (N, 3), `triangles_np` int64 (M, 3)."""

    __slots__ = ("vertices_np", "triangles_np")

    def __init__(self, vertices=None, triangles=None):
        self.vertices_np = self._coerce(vertices, np.float64, "vertices")
        self.triangles_np = self._coerce(triangles, np.int64, "triangles")

    @staticmethod
    def _coerce(a, dtype, what):
        if a is None or (isinstance(a, (list, tuple)) and len(a) == 0):
            return np.zeros((0, 3), dtype=dtype)
        arr = np.ascontiguousarray(np.asarray(a), dtype=dtype)
        if arr.ndim != 2 or arr.shape[1] != 3:
            raise ValueError(f"{what} ndarray must be (N, 3); got {arr.shape}")
        return arr

    @classmethod
    def from_arrays(cls, verts, tris) -> "Model":
        return cls(verts, tris)

    def copy(self) -> "Model":
        return Model(self.vertices_np.copy(), self.triangles_np.copy())

    def __repr__(self):
        return f"Model(n_v={len(self.vertices_np)}, n_t={len(self.triangles_np)})"


@dataclass
class CompressedModel:
    data: bytes
    bits_per_vertex: float
    bits_per_triangle: float
