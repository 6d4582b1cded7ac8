"""Seeded procedural meshes for the STRIDE decode/encode workloads.

No mesh assets ship with this repository, so every workload is generated
from a seed:

* ``uv_sphere``     lat-long sphere, ``n_rings`` rings of ``n_segs`` vertices
                    plus two poles (``n_rings=128, n_segs=256`` gives
                    32,770 vertices / 65,536 triangles).
* ``icosphere``     icosahedron subdivided ``level`` times
                    (level 8: 655,362 vertices / 1,310,720 triangles).
* ``torus``         ring torus, major radius R, minor radius r.
* ``heightfield``   ``n x n`` grid over the unit square
                    (n=1900: 3,610,000 vertices / 7,212,402 triangles).

Displacements use lattice value noise (3-D, trilinear, smoothstep) and
lattice gradient noise (2-D, quintic fade), summed as fBm with lacunarity 2
and gain 0.5.  Lattice values come from a 64-bit integer hash of
(lattice coords, seed, octave), so results are bit-reproducible on any
machine with IEEE doubles.  Positions are rounded to float32 (the
workloads are float32 meshes) and handed back as float64 arrays, the
storage type of ``Model.vertices_np``.
"""
from __future__ import annotations

import math

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def _mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 arrays."""
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * _M2
        x = (x ^ (x >> np.uint64(27))) * _M3
        x = x ^ (x >> np.uint64(31))
    return x


def _lattice_hash(seed: int, *coords: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        h = np.full(np.broadcast(*coords).shape,
                    np.uint64((seed * 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFFFFFF))
        for c in coords:
            h = _mix64(h + c.astype(np.int64).view(np.uint64) * _M1)
    return h


def _unit(h: np.ndarray) -> np.ndarray:
    """uint64 hash -> float64 uniform in [-1, 1)."""
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise3(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise with smoothstep weights, range about [-1, 1]."""
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    w = t * t * (3.0 - 2.0 * t)
    out = np.zeros(p.shape[0], dtype=np.float64)
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                h = _lattice_hash(seed, i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
                out += wx * wy * wz * _unit(h)
    return out


def gradient_noise2(p: np.ndarray, seed: int) -> np.ndarray:
    """2-D lattice gradient noise with quintic fade, range about [-0.7, 0.7]."""
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    w = t * t * t * (t * (t * 6.0 - 15.0) + 10.0)
    out = np.zeros(p.shape[0], dtype=np.float64)
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            h = _lattice_hash(seed, i[:, 0] + dx, i[:, 1] + dy)
            ang = _unit(h) * math.pi
            gx, gy = np.cos(ang), np.sin(ang)
            out += wx * wy * (gx * (t[:, 0] - dx) + gy * (t[:, 1] - dy))
    return out


def fbm3(p: np.ndarray, octaves: int, seed: int, base_freq: float = 2.0) -> np.ndarray:
    out = np.zeros(p.shape[0], dtype=np.float64)
    amp, freq = 1.0, base_freq
    for o in range(octaves):
        out += amp * value_noise3(p * freq, seed * 131 + o)
        amp *= 0.5
        freq *= 2.0
    return out


def fbm2(p: np.ndarray, octaves: int, seed: int, base_freq: float = 4.0) -> np.ndarray:
    out = np.zeros(p.shape[0], dtype=np.float64)
    amp, freq = 1.0, base_freq
    for o in range(octaves):
        out += amp * gradient_noise2(p * freq, seed * 131 + o)
        amp *= 0.5
        freq *= 2.0
    return out


def _as_f32(v: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(v.astype(np.float32).astype(np.float64))


# --------------------------------------------------------------------------
# Base surfaces
# --------------------------------------------------------------------------

def uv_sphere(n_rings: int, n_segs: int):
    """Unit lat-long sphere: north pole, n_rings rings, south pole."""
    th = np.pi * np.arange(1, n_rings + 1) / (n_rings + 1)
    ph = 2.0 * np.pi * np.arange(n_segs) / n_segs
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    ring = np.stack([st * np.cos(ph)[None, :], st * np.sin(ph)[None, :],
                     np.broadcast_to(ct, (n_rings, n_segs))], axis=-1).reshape(-1, 3)
    V = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]])
    north, south = 0, 1 + n_rings * n_segs
    s = np.arange(n_segs)
    s1 = (s + 1) % n_segs
    tris = [np.stack([np.full(n_segs, north), 1 + s, 1 + s1], axis=1)]
    for r in range(n_rings - 1):
        a = 1 + r * n_segs + s
        b = 1 + r * n_segs + s1
        c = 1 + (r + 1) * n_segs + s
        d = 1 + (r + 1) * n_segs + s1
        tris.append(np.stack([a, c, b], axis=1))
        tris.append(np.stack([b, c, d], axis=1))
    base = 1 + (n_rings - 1) * n_segs
    tris.append(np.stack([base + s, np.full(n_segs, south), base + s1], axis=1))
    return V, np.concatenate(tris).astype(np.int64)


def icosphere(level: int):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    T = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
                 dtype=np.int64)
    for _ in range(level):
        nv = V.shape[0]
        e = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        V = np.concatenate([V, mid])
        n = T.shape[0]
        m01, m12, m20 = nv + inv[:n], nv + inv[n:2 * n], nv + inv[2 * n:]
        T = np.concatenate([
            np.stack([T[:, 0], m01, m20], axis=1),
            np.stack([T[:, 1], m12, m01], axis=1),
            np.stack([T[:, 2], m20, m12], axis=1),
            np.stack([m01, m12, m20], axis=1)])
    return V, T.astype(np.int64)


def torus(n_u: int, n_v: int, R: float = 1.0, r: float = 0.3):
    u = 2.0 * np.pi * np.arange(n_u) / n_u
    v = 2.0 * np.pi * np.arange(n_v) / n_v
    uu, vv = np.meshgrid(u, v, indexing="ij")
    x = (R + r * np.cos(vv)) * np.cos(uu)
    y = (R + r * np.cos(vv)) * np.sin(uu)
    z = r * np.sin(vv)
    V = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    i = np.arange(n_u)[:, None]
    j = np.arange(n_v)[None, :]
    a = (i * n_v + j)
    b = (((i + 1) % n_u) * n_v + j)
    c = (i * n_v + (j + 1) % n_v)
    d = (((i + 1) % n_u) * n_v + (j + 1) % n_v)
    T = np.concatenate([np.stack([a, b, c], axis=-1).reshape(-1, 3),
                        np.stack([c, b, d], axis=-1).reshape(-1, 3)])
    return V, T.astype(np.int64)


def grid(n: int):
    """n x n vertex grid over the unit square, z = 0."""
    xs = np.arange(n, dtype=np.float64) / (n - 1)
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    V = np.stack([gx.ravel(), gy.ravel(), np.zeros(n * n)], axis=1)
    i = np.arange(n - 1)[:, None]
    j = np.arange(n - 1)[None, :]
    a = i * n + j
    b = (i + 1) * n + j
    c = i * n + j + 1
    d = (i + 1) * n + j + 1
    T = np.concatenate([np.stack([a, b, c], axis=-1).reshape(-1, 3),
                        np.stack([c, b, d], axis=-1).reshape(-1, 3)])
    return V, T.astype(np.int64)


# --------------------------------------------------------------------------
# Displaced families
# --------------------------------------------------------------------------

def displaced_uv_sphere(n_rings=128, n_segs=256, amplitude=0.05, octaves=4, seed=0):
    V, T = uv_sphere(n_rings, n_segs)
    d = fbm3(V, octaves, seed)
    return _as_f32(V * (1.0 + amplitude * d)[:, None]), T


def displaced_icosphere(level=8, amplitude=0.08, octaves=6, seed=0):
    V, T = icosphere(level)
    d = fbm3(V, octaves, seed)
    return _as_f32(V * (1.0 + amplitude * d)[:, None]), T


def displaced_torus(n_u, n_v, R=1.0, r=0.3, amplitude=0.05, octaves=4, seed=0):
    V, T = torus(n_u, n_v, R, r)
    d = fbm3(V, octaves, seed)
    # displace along the tube normal
    c = np.stack([R * np.cos(np.arctan2(V[:, 1], V[:, 0])),
                  R * np.sin(np.arctan2(V[:, 1], V[:, 0])), np.zeros(len(V))], axis=1)
    nrm = V - c
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return _as_f32(V + (amplitude * r * d)[:, None] * nrm), T


def heightfield(n=1900, amplitude=0.1, octaves=6, seed=0):
    V, T = grid(n)
    V[:, 2] = amplitude * fbm2(V[:, :2], octaves, seed)
    return _as_f32(V), T


# --------------------------------------------------------------------------
# Named workloads
# --------------------------------------------------------------------------

def uv_sphere_c1():
    """UV sphere 128 rings x 256 segments, 4-octave value noise amp 0.05, seed 0."""
    return displaced_uv_sphere(128, 256, 0.05, 4, 0)


def icosphere_l8(seed=0):
    """Icosphere level 8, 6-octave fBm amp 0.08."""
    return displaced_icosphere(8, 0.08, 6, seed)


def heightfield_1900(seed=0):
    """1900 x 1900 heightfield, z = 0.1 x 6-octave gradient noise."""
    return heightfield(1900, 0.1, 6, seed)


FAMILIES = ("uv_sphere", "torus", "heightfield")


def mesh_with_tris(family: str, n_tris: int, seed: int):
    """A member of `family` with about `n_tris` triangles."""
    if family == "uv_sphere":
        # tris = 2 * n_segs * n_rings, n_segs = 2 * n_rings
        n_rings = max(2, int(round(math.sqrt(n_tris / 4.0))))
        return displaced_uv_sphere(n_rings, 2 * n_rings, 0.05, 4, seed)
    if family == "torus":
        # tris = 2 * n_u * n_v, n_u = 3 * n_v
        n_v = max(3, int(round(math.sqrt(n_tris / 6.0))))
        return displaced_torus(3 * n_v, n_v, 1.0, 0.3, 0.05, 4, seed)
    if family == "heightfield":
        n = max(2, int(round(math.sqrt(n_tris / 2.0))) + 1)
        return heightfield(n, 0.1, 6, seed)
    raise ValueError(f"unknown family {family!r}")


def batch_specs(n_meshes: int = 2048, t_min: int = 1000, t_max: int = 200000,
                seed: int = 0):
    """(family, n_tris, seed) triples, n_tris log-uniform in [t_min, t_max]."""
    rng = np.random.default_rng(seed)
    logt = rng.uniform(math.log(t_min), math.log(t_max), size=n_meshes)
    fam = rng.integers(0, len(FAMILIES), size=n_meshes)
    return [(FAMILIES[int(f)], int(round(math.exp(l))), seed * 100003 + i)
            for i, (f, l) in enumerate(zip(fam, logt))]


def learned_specs(n_meshes: int, seed: int, t_min: int = 20000, t_max: int = 500000):
    """(family, n_tris, seed) triples for the learned-proposal meshes:
    families uniform, n_tris log-uniform in [t_min, t_max].  Training uses
    seed 0 and held-out evaluation seed 1, so the two sets share no mesh."""
    rng = np.random.default_rng(1000 + seed)
    logt = rng.uniform(math.log(t_min), math.log(t_max), size=n_meshes)
    fam = rng.integers(0, len(FAMILIES), size=n_meshes)
    return [(FAMILIES[int(f)], int(round(math.exp(l))), (seed + 1) * 1000003 + i)
            for i, (f, l) in enumerate(zip(fam, logt))]
