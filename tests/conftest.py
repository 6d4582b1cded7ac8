import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and the built HIP library")
    config.addinivalue_line("markers", "slow: large workloads")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import stride_oracle
    return stride_oracle


@pytest.fixture(scope="session")
def coracle():
    from tests import oracle_c
    return oracle_c.load()


def small_meshes():
    """A spread of small seeded meshes (name, V, T)."""
    from meshpress_amd import workloads as w
    out = []
    V, T = w.displaced_uv_sphere(12, 24, 0.05, 4, 3)
    out.append(("uv_12x24", V, T))
    V, T = w.displaced_torus(30, 10, 1.0, 0.3, 0.05, 4, 5)
    out.append(("torus_30x10", V, T))
    V, T = w.heightfield(24, 0.1, 6, 7)
    out.append(("hf_24", V, T))
    V, T = w.displaced_icosphere(2, 0.08, 6, 1)
    out.append(("ico_2", V, T))
    # single triangle
    out.append(("one_tri", np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0.5]]), np.array([[0, 1, 2]])))
    # two disjoint triangles (two strips, shared nothing)
    out.append(("two_tris", np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [3, 3, 3], [4, 3, 3],
                                      [3, 4, 3.5]]), np.array([[0, 1, 2], [3, 4, 5]])))
    # a two-row ribbon whose meshlets reach the maximum of 256 locals
    # (local id 255, 9-bit raw ids)
    out.append(("ribbon_2x1000",) + ribbon(1000))
    return out


def ribbon(n, seed=0):
    """A strip of 2 x n vertices (2 (n - 1) triangles) with a little z noise:
    every vertex lies on the mesh border, so each meshlet of max_verts=256 and
    max_tris=256 holds 256 vertices and 254 triangles."""
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.linspace(0.0, 10.0, n), np.array([0.0, 0.1]))
    V = np.stack([x.ravel(), y.ravel(), 0.05 * rng.standard_normal(x.size)], 1)
    c = np.arange(n - 1)
    T = np.concatenate([np.stack([c, c + 1, c + n], 1), np.stack([c + 1, c + n + 1, c + n], 1)])
    return V, T.astype(np.int64)


def encode_with_oracle(V, T, *, pred="canonical", precision_error=0.0005,
                       precision_mode="world", max_verts=256, max_tris=256):
    """(prep with plans, DPR4 bytes) using the native front-end and the oracle writer."""
    from meshpress_amd.encoder.prep import prepare
    from oracle import stride_oracle as O
    p = prepare(V, T, precision_error=precision_error, precision_mode=precision_mode,
                max_verts=max_verts, max_tris=max_tris)
    prep = dict(p)
    prep["plans"] = p["native"].export_plans()
    if pred == "canonical":
        pn, pk = O.CANON_N, O.CANON_K
    elif pred == "fit":
        A, B, C, Vv, off = p["native"].para_samples()
        pn, pk = O.fit_predictor(A, B, C, Vv, off)
    else:
        pn, pk = pred
    prep["pred_n"], prep["pred_K"] = np.asarray(pn), np.asarray(pk)
    return prep, O.encode_dup(prep, pn, pk)
