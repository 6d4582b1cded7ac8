"""CPU oracle for the PDLT v5 (ParaDelta v5) format.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it
as the checker.  It restates, in plain Python + numpy, the reference's
algorithm (maletsden/meshpress):

  boundary table   utils/boundary_split.py:46-70        identify_boundary_verts,
                                                       build_boundary_table
                   utils/boundary_bvh.py:41-53          morton_permute_boundary
  f32 recon chain  encoder/paradelta_v5_nb.py:20-212    interior_pass_strip_nb
  strip traversal  encoder/paradelta_v5.py:34-83        _strip_traversal
  LIN3 fit         encoder/paradelta_codec.py:626-639   _fit_linear3
  meshlet writer   encoder/paradelta_v5_nb.py:328-539   write_meshlet_nb
  stream writer    encoder/paradelta_v5.py:144-283      encode_from_prepared_v5
  decoder          encoder/paradelta_v5.py:286-462      decode_paradelta_v5

How each part is pinned (tests/test_pdlt_oracle.py):
  * decode_v5: golden vectors written by the reference's own
    decode_paradelta_v5 on streams of this writer (tests/golden/pdlt_*.npz),
    bit-exact in V and T; the boundary-centroid summation order
    (np.mean(axis=0) of the boundary rows) against numpy on random rows;
  * encode_v5: the reference decoder accepts every stream (golden), decoded
    positions stay within the quantisation bound of the source mesh, the
    per-axis coder choice is checked by brute force over the three coders,
    and hand-derived byte fixtures pin the field layout;
  * boundary table: the Morton key against a bit-by-bit interleave, the
    stable order against sorted() of (key, id) pairs;
  * LIN3 fit: exact weights on samples generated from known weights.
The reference's v5 writer itself is Numba code (numba is absent), so the
writer's byte stream is pinned only through the decoder and the invariants
above.
"""
from __future__ import annotations

import struct
from collections import deque

import numpy as np

from .stride_oracle import (FIFO_SIZE, REUSE_BITS, BitReader, BitWriter, emit_vert,
                            idx_bits_for, read_vert, root_orient, strip_step, zigzag)

MAGIC_PDLT = 0x504C4454
VERSION_V5 = 5
RICE_K_MAX = 11
EG_K_MAX = 7
f32 = np.float32


# ---------------------------------------------------------------------------
# Boundary table
# ---------------------------------------------------------------------------
def spread3_21(v: int) -> int:
    """Bits 0..20 of v moved to bit positions 0, 3, 6, ... (utils/boundary_split.py:148-157)."""
    v &= 0x1FFFFF
    v = (v | (v << 32)) & 0x001F00000000FFFF
    v = (v | (v << 16)) & 0x001F0000FF0000FF
    v = (v | (v << 8)) & 0x100F00F00F00F00F
    v = (v | (v << 4)) & 0x10C30C30C30C30C3
    v = (v | (v << 2)) & 0x1249249249249249
    return v


def morton3(code_xyz) -> int:
    x, y, z = (int(c) for c in code_xyz)
    return spread3_21(x) | (spread3_21(y) << 1) | (spread3_21(z) << 2)


def boundary_list(plans, global_codes) -> list[int]:
    """Global ids used by >= 2 meshlets, sorted, then stably ordered by
    Morton code of their global codes (utils/boundary_split.py:46-70,
    utils/boundary_bvh.py:41-53)."""
    count: dict[int, int] = {}
    for pl in plans:
        for g in set(int(x) for x in pl["local_to_global"]):
            count[g] = count.get(g, 0) + 1
    ids = sorted(g for g, c in count.items() if c >= 2)
    keys = [morton3(global_codes[g]) for g in ids]
    order = sorted(range(len(ids)), key=lambda i: (keys[i], i))
    return [ids[i] for i in order]


def plan_refs(plan, gv_to_ref) -> list[int]:
    """Boundary-table refs of a plan's boundary locals (0..n_bnd-1); they
    must be strictly increasing (paradelta_v5_nb.py:345-358 codes ref
    differences minus one)."""
    n_bnd = int(plan["n_bnd"])
    l2g = plan["local_to_global"]
    refs = [gv_to_ref[int(g)] for g in l2g[:n_bnd]]
    if any(b <= a for a, b in zip(refs, refs[1:])):
        raise ValueError("boundary refs are not strictly increasing")
    for g in l2g[n_bnd:]:
        if int(g) in gv_to_ref:
            raise ValueError("interior local is a boundary vertex")
    return refs


def dequant(codes, g_min, g_range, g_bits) -> np.ndarray:
    """g_min + code / (2^bits - 1) * g_range in f64 (paradelta_codec.py:167-175)."""
    codes = np.asarray(codes, dtype=np.int64).reshape(-1, 3)
    out = np.zeros(codes.shape, dtype=np.float64)
    for d in range(3):
        mx = (1 << int(g_bits[d])) - 1
        if mx == 0:
            out[:, d] = float(g_min[d])
        else:
            out[:, d] = float(g_min[d]) + codes[:, d].astype(np.float64) / mx * float(g_range[d])
    return out


# ---------------------------------------------------------------------------
# Strip traversal and the f32 reconstruction chain
# ---------------------------------------------------------------------------
def strip_traversal(tris_local, strips, n_bnd):
    """[(v, kind, refs)] for each interior local in first-use order: kind
    'none' for strip-root vertices, 'para' with (a, b, c) otherwise
    (encoder/paradelta_v5.py:34-83).  Boundary locals start as decoded."""
    decoded = set(range(n_bnd))
    order = []
    for strip in strips:
        root = root_orient(tris_local, strip[0], strip[1] if len(strip) > 1 else None)
        for v in root:
            if v >= n_bnd and v not in decoded:
                order.append((v, "none", ()))
            decoded.add(v)
        prev = list(root)
        for ti in strip[1:]:
            st = strip_step(prev, tris_local[ti])
            if st is None:
                raise RuntimeError("oldest-edge share in strip")
            _, refs, nv, prev = st
            if nv >= n_bnd and nv not in decoded:
                order.append((nv, "para", refs))
            decoded.add(nv)
    return order


def interior_pass(plan, vn, bnd_recon_norm, delta, w3=None, recon_out=None):
    """(codes int64 (n_int, 3), samples) of one meshlet with the encoder's
    float32 chain (encoder/paradelta_v5_nb.py:20-212).  w3 None is the
    sample-collecting pass with the plain parallelogram a + b - c.
    recon_out (dict): receives global id -> float32 reconstruction of
    every interior vertex."""
    n_bnd, n_int = int(plan["n_bnd"]), int(plan["n_int"])
    if n_int == 0:
        return np.zeros((0, 3), dtype=np.int64), []
    l2g = [int(g) for g in plan["local_to_global"]]
    n_local = n_bnd + n_int
    recon = np.zeros((n_local, 3), dtype=np.float32)
    for lid in range(n_bnd):
        for d in range(3):
            recon[lid, d] = f32(bnd_recon_norm[l2g[lid], d])
    fb = [f32(0.0)] * 3
    if n_bnd > 0:
        s = [f32(0.0)] * 3
        for lid in range(n_bnd):
            s = [f32(s[d] + recon[lid, d]) for d in range(3)]
        nb = f32(n_bnd)
        fb = [f32(s[d] / nb) for d in range(3)]
    dlt = f32(delta)
    w = None if w3 is None else [f32(x) for x in w3]
    codes = np.zeros((n_int, 3), dtype=np.int64)
    samples = []
    tl = np.asarray(plan["ml_tris_local"])
    for i, (v, kind, refs) in enumerate(strip_traversal(tl, plan["strips"], n_bnd)):
        g = l2g[v]
        if kind == "none":
            pred = list(fb)
        else:
            a, b, c = refs
            if w is None:
                pred = [f32(f32(recon[a, d] + recon[b, d]) - recon[c, d]) for d in range(3)]
            else:
                pred = [f32(f32(f32(w[0] * recon[a, d]) + f32(w[1] * recon[b, d]))
                            + f32(w[2] * recon[c, d])) for d in range(3)]
        for d in range(3):
            true = f32(vn[g, d])
            cd = int(np.round(f32(f32(true - pred[d]) / dlt)))
            codes[i, d] = cd
            recon[v, d] = f32(pred[d] + f32(f32(cd) * dlt))
        if recon_out is not None:
            recon_out[g] = recon[v].copy()
        if kind == "para" and w is None:
            a, b, c = refs
            samples.append((recon[a].astype(np.float64), recon[b].astype(np.float64),
                            recon[c].astype(np.float64), np.asarray(vn[g], dtype=np.float64)))
    return codes, samples


def fit_linear3(samples) -> np.ndarray:
    """Least-squares (w_a, w_b, w_c) over all axes, as float32
    (encoder/paradelta_codec.py:626-639)."""
    if len(samples) == 0:
        return np.array([1.0, 1.0, -1.0], dtype=np.float32)
    A = np.stack([s[0] for s in samples])
    B = np.stack([s[1] for s in samples])
    C = np.stack([s[2] for s in samples])
    T = np.stack([s[3] for s in samples])
    M = np.stack([A.ravel(), B.ravel(), C.ravel()], axis=1)
    w, *_ = np.linalg.lstsq(M, T.ravel(), rcond=None)
    return w.astype(np.float32)


# ---------------------------------------------------------------------------
# Writer
# ---------------------------------------------------------------------------
def best_rice_k(u, k_max=RICE_K_MAX) -> tuple[int, int]:
    """(k, bits) with the lowest Rice cost; lowest k on ties."""
    u = np.asarray(u, dtype=np.int64)
    best = (0, None)
    for k in range(k_max + 1):
        b = int((u >> k).sum()) + len(u) * (1 + k)
        if best[1] is None or b < best[1]:
            best = (k, b)
    return best


def best_eg_k(u, k_max=EG_K_MAX) -> tuple[int, int]:
    u = np.asarray(u, dtype=np.int64)
    best = (0, None)
    for k in range(k_max + 1):
        b = sum(2 * (int((x >> k) + 1).bit_length() - 1) + 1 + k for x in u.tolist())
        if best[1] is None or b < best[1]:
            best = (k, b)
    return best


def axis_coder(vals) -> tuple[int, dict, int]:
    """(tag, params, total bits incl. the tag byte) of one residual axis:
    fixed width (tag 0), Rice (1) or Exp-Golomb (2); strictly smaller total
    wins, in that order (paradelta_v5_nb.py:449-486)."""
    vals = [int(x) for x in vals]
    mn, mx = min(vals), max(vals)
    rng = mx - mn
    bw = max(1, (rng + 1).bit_length()) if rng > 0 else 1
    u = zigzag(np.array(vals, dtype=np.int64))
    fixed = 8 + 16 + 8 + len(vals) * bw
    if mn < -32768 or mn > 32767:
        fixed = None
    rk, rb = best_rice_k(u)
    ek, eb = best_eg_k(u)
    tag, best = 0, fixed
    if best is None or 16 + rb < best:
        tag, best = 1, 16 + rb
    if 16 + eb < best:
        tag, best = 2, 16 + eb
    params = {0: {"mn": mn, "bw": bw}, 1: {"k": rk}, 2: {"k": ek}}[tag]
    return tag, params, best


def write_meshlet(w: BitWriter, plan, refs, codes, force_tags=None) -> None:
    """One meshlet (paradelta_v5_nb.py:328-539).  force_tags (test streams
    only): a tag per axis overriding the cost choice (tag 0 only where the
    minimum fits int16)."""
    n_bnd, n_int = int(plan["n_bnd"]), int(plan["n_int"])
    tl = np.asarray(plan["ml_tris_local"])
    strips = plan["strips"]
    w.put(n_bnd, 16)
    w.put(n_int, 16)
    w.put(len(tl), 16)
    w.put(len(strips), 16)
    if n_bnd > 0:
        w.put(refs[0], 32)
        if n_bnd > 1:
            diffs = np.array([refs[i + 1] - refs[i] - 1 for i in range(n_bnd - 1)], dtype=np.int64)
            k, _ = best_rice_k(diffs)
            w.put(k, 8)
            for x in diffs.tolist():
                w.put_rice(int(x), k)
    ib = idx_bits_for(n_bnd + n_int)
    fifo: deque = deque(maxlen=FIFO_SIZE)
    for strip in strips:
        w.put(len(strip), 16)
        root = root_orient(tl, strip[0], strip[1] if len(strip) > 1 else None)
        for v in root:
            emit_vert(w, v, fifo, ib)
        prev = list(root)
        for ti in strip[1:]:
            st = strip_step(prev, tl[ti])
            if st is None:
                raise RuntimeError("oldest-edge share")
            ec, _, nv, prev = st
            w.put(ec, 1)
            emit_vert(w, nv, fifo, ib)
    if n_int > 0:
        for d in range(3):
            col = [int(x) for x in codes[:, d]]
            tag, p, _ = axis_coder(col)
            if force_tags is not None:
                ft = int(force_tags[d])
                mn, mx = min(col), max(col)
                if ft == 0 and -32768 <= mn <= 32767:
                    tag, p = 0, {"mn": mn, "bw": max(1, (mx - mn + 1).bit_length()) if mx > mn else 1}
                elif ft in (1, 2):
                    u = zigzag(np.array(col, dtype=np.int64))
                    tag, p = ft, {"k": (best_rice_k(u) if ft == 1 else best_eg_k(u))[0]}
            w.put(tag, 8)
            if tag == 0:
                w.put(p["mn"] & 0xFFFF, 16)
                w.put(p["bw"], 8)
                for x in col:
                    w.put(x - p["mn"], p["bw"])
            else:
                w.put(p["k"], 8)
                for x in zigzag(np.array(col, dtype=np.int64)).tolist():
                    (w.put_rice if tag == 1 else w.put_eg)(int(x), p["k"])


def plans_from_partition(tris, parts, global_codes):
    """Reference-shaped plans for an explicit partition (lists of triangle
    ids), one strip per triangle: boundary locals first in boundary-table
    order, interior locals in first-use order.  Builds small test streams
    with meshlets the partitioner never makes (no interior or no boundary
    vertices)."""
    tris = np.asarray(tris, dtype=np.int64)
    stub = [{"local_to_global": sorted(set(int(x) for x in tris[p].ravel()))} for p in parts]
    bl = boundary_list(stub, np.asarray(global_codes, dtype=np.int64))
    ref = {g: i for i, g in enumerate(bl)}
    plans = []
    for p in parts:
        used = []
        for t in p:
            for g in tris[t]:
                if int(g) not in used:
                    used.append(int(g))
        bnd = sorted((g for g in used if g in ref), key=lambda g: ref[g])
        intr = [g for g in used if g not in ref]
        l2g = bnd + intr
        loc = {g: i for i, g in enumerate(l2g)}
        tl = np.array([[loc[int(g)] for g in tris[t]] for t in p], dtype=np.int64).reshape(-1, 3)
        plans.append({"local_to_global": l2g, "ml_tris_local": tl,
                      "strips": [[i] for i in range(len(p))], "n_bnd": len(bnd),
                      "n_int": len(intr), "n_tris_m": len(p), "n_strips": len(p)})
    return plans


def prepare_v5(prep: dict) -> dict:
    """Adds the v5 fields to a prep dict with plans: boundary_list,
    per-plan refs and bnd_recon_norm (paradelta_codec.py:834-852)."""
    codes = np.asarray(prep["global_codes"], dtype=np.int64)
    bl = boundary_list(prep["plans"], codes)
    gv_to_ref = {g: i for i, g in enumerate(bl)}
    out = dict(prep)
    out["boundary_list"] = bl
    out["n_boundary"] = len(bl)
    out["refs"] = [plan_refs(pl, gv_to_ref) for pl in prep["plans"]]
    out["bnd_recon_norm"] = dequant(codes, prep["g_min"], prep["g_range"], prep["g_bits"])
    return out


def encode_v5(prep: dict, return_info: bool = False, force_tags=None):
    """PDLT v5 bytes for a prep dict with plans (encode_from_prepared_v5,
    encoder/paradelta_v5.py:144-283).  force_tags(m) -> 3 tags: test
    streams that exercise every residual coder."""
    p = prep if "boundary_list" in prep else prepare_v5(prep)
    vn = np.asarray(p["vn"], dtype=np.float64)
    delta = 2.0 * float(p["per_coord_err"])
    samples = []
    for pl in p["plans"]:
        samples.extend(interior_pass(pl, vn, p["bnd_recon_norm"], delta, None)[1])
    lin3 = fit_linear3(samples)
    w3 = lin3.astype(np.float64)
    body = BitWriter()
    rel, max_abs = [], 0
    for pl, refs in zip(p["plans"], p["refs"]):
        codes, _ = interior_pass(pl, vn, p["bnd_recon_norm"], delta, w3)
        if len(codes):
            max_abs = max(max_abs, int(np.abs(codes).max()))
        rel.append(body.bit_pos())
        write_meshlet(body, pl, refs, codes,
                      None if force_tags is None else force_tags(len(rel) - 1))
    g_bits = [int(b) for b in p["g_bits"]]
    w = BitWriter()
    w.put(MAGIC_PDLT, 32)
    w.put(VERSION_V5, 8)
    w.put(1 if max_abs > 32767 else 0, 8)
    for c in p["center"]:
        w.put_f32(float(c))
    w.put_f32(float(p["scale"]))
    w.put_f32(float(p["per_coord_err"]))
    for v in p["g_min"]:
        w.put_f32(float(v))
    for v in p["g_range"]:
        w.put_f32(float(v))
    for b in g_bits:
        w.put(b, 8)
    n_b = len(p["boundary_list"])
    w.put(int(p["n_v"]), 32)
    w.put(int(p["n_t"]), 32)
    w.put(n_b, 32)
    w.put(len(p["plans"]), 32)
    for x in lin3:
        w.put_f32(float(x))
    if n_b > 0:
        bc = np.asarray(p["global_codes"], dtype=np.int64)[np.asarray(p["boundary_list"])]
        for d in range(3):
            arr = bc[:, d]
            w.put(int(arr[0]), g_bits[d])
            if n_b > 1:
                u = zigzag(arr[1:] - arr[:-1])
                k, _ = best_rice_k(u)
                w.put(k, 8)
                for x in u.tolist():
                    w.put_rice(int(x), k)
    w.put(0, (-w.bit_pos()) & 7)
    region = w.bit_pos() + 32 * len(rel)
    for r in rel:
        w.put(region + r, 32)
    data = w.tobytes() + body.tobytes()
    if return_info:
        return data, {"lin3": lin3, "max_abs": max_abs, "n_boundary": n_b,
                      "n_samples": len(samples)}
    return data


# ---------------------------------------------------------------------------
# Decoder
# ---------------------------------------------------------------------------
def read_header(data: bytes) -> dict:
    """Global header and boundary table (paradelta_v5.py:290-335)."""
    r = BitReader(data)
    magic = r.get(32)
    if magic != MAGIC_PDLT:
        raise ValueError(f"bad magic 0x{magic:08X}")
    version = r.get(8)
    if version != VERSION_V5:
        raise ValueError(f"expected v5, got v{version}")
    h = {"code_width": r.get(8)}
    h["center"] = [r.f32() for _ in range(3)]
    h["scale"] = r.f32()
    h["per_coord_err"] = r.f32()
    h["g_min"] = [r.f32() for _ in range(3)]
    h["g_range"] = [r.f32() for _ in range(3)]
    h["g_bits"] = [r.get(8) for _ in range(3)]
    h["n_v"], h["n_t"], h["n_boundary"], h["n_meshlets"] = (r.get(32) for _ in range(4))
    h["lin3"] = [r.f32() for _ in range(3)]
    n_b = h["n_boundary"]
    bc = np.zeros((n_b, 3), dtype=np.int64)
    for d in range(3):
        if n_b == 0:
            continue
        bc[0, d] = r.get(h["g_bits"][d])
        if n_b > 1:
            k = r.get(8)
            for i in range(n_b - 1):
                u = r.rice(k)
                bc[i + 1, d] = bc[i, d] + ((u >> 1) ^ -(u & 1))
    h["bnd_codes"] = bc
    r.pos += (-r.pos) & 7
    h["offsets"] = [r.get(32) for _ in range(h["n_meshlets"])]
    h["meshlets_bit"] = r.pos
    return h


def decode_v5(data: bytes, return_info: bool = False, world: bool = True, chain: str = "f64"):
    """(V float32 (n, 3), T int64 (m, 3)) as decode_paradelta_v5
    (encoder/paradelta_v5.py:286-462): f64 positions, the boundary centroid
    as fallback prediction, LIN3 weights as stored (f32 widened to f64).
    world=False: the normalised positions (before scale and center) as
    float32.  chain="f32": the reconstruction chain in float32 exactly as
    the encoder ran it (interior_pass with the stored weights), so positions
    are the encoder's own reconstructions; the output transform stays f64."""
    if chain not in ("f64", "f32"):
        raise ValueError("chain must be 'f64' or 'f32'")
    c32 = chain == "f32"
    h = read_header(data)
    g_bits = h["g_bits"]
    bnd_norm = dequant(h["bnd_codes"], h["g_min"], h["g_range"], g_bits) \
        if h["n_boundary"] else np.zeros((0, 3))
    delta = 2.0 * h["per_coord_err"]
    w0, w1, w2 = h["lin3"]
    if c32:
        bnd_norm = bnd_norm.astype(np.float32).astype(np.float64)
        w32 = [f32(x) for x in h["lin3"]]
        d32 = f32(delta)
    verts = [row.copy() for row in bnd_norm]
    tris = []
    r = BitReader(data, h["meshlets_bit"])
    info = {"codes": [], "emit": []}
    for _ in range(h["n_meshlets"]):
        n_bnd, n_int, _n_tris, n_strips = (r.get(16) for _ in range(4))
        n_local = n_bnd + n_int
        refs = []
        if n_bnd > 0:
            refs.append(r.get(32))
            if n_bnd > 1:
                k = r.get(8)
                for _ in range(n_bnd - 1):
                    refs.append(refs[-1] + r.rice(k) + 1)
        l2g = list(refs)
        first_int = len(verts)
        l2g += [first_int + i for i in range(n_int)]
        verts += [np.zeros(3) for _ in range(n_int)]
        ib = idx_bits_for(n_local)
        fifo: deque = deque(maxlen=FIFO_SIZE)
        local_tris, emit = [], []
        seen = set(range(n_bnd))
        for _s in range(n_strips):
            L = r.get(16)
            v0, v1, v2 = (read_vert(r, fifo, ib) for _ in range(3))
            local_tris.append((v0, v1, v2))
            for v in (v0, v1, v2):
                if v >= n_bnd and v not in seen:
                    emit.append((v, None))
                seen.add(v)
            prev = [v0, v1, v2]
            for _ in range(L - 1):
                ec = r.get(1)
                nv = read_vert(r, fifo, ib)
                if ec == 0:
                    a, b, c = prev[1], prev[2], prev[0]
                    prev = [prev[1], prev[2], nv]
                else:
                    a, b, c = prev[0], prev[2], prev[1]
                    prev = [prev[0], prev[2], nv]
                local_tris.append((a, b, nv))
                if nv >= n_bnd and nv not in seen:
                    emit.append((nv, (a, b, c)))
                seen.add(nv)
        if n_int > 0:
            codes = np.zeros((n_int, 3), dtype=np.int64)
            for d in range(3):
                tag = r.get(8)
                if tag == 0:
                    mn = r.get(16)
                    mn = mn - 0x10000 if mn & 0x8000 else mn
                    bw = r.get(8)
                    for i in range(n_int):
                        codes[i, d] = mn + r.get(bw)
                elif tag in (1, 2):
                    k = r.get(8)
                    for i in range(n_int):
                        u = r.rice(k) if tag == 1 else r.eg(k)
                        codes[i, d] = (u >> 1) ^ -(u & 1)
                else:
                    raise ValueError(f"bad tag {tag}")
            recon = {lid: np.asarray(verts[l2g[lid]], dtype=np.float64).copy()
                     for lid in range(n_bnd)}
            # boundary centroid: rows summed in order, then divided by n_bnd
            fb = np.zeros(3)
            if n_bnd > 0 and not c32:
                s = np.zeros(3)
                for lid in range(n_bnd):
                    s = s + recon[lid]
                fb = s / n_bnd
            elif n_bnd > 0:
                s = [f32(0.0)] * 3
                for lid in range(n_bnd):
                    s = [f32(s[d] + f32(recon[lid][d])) for d in range(3)]
                fb = np.array([float(f32(s[d] / f32(n_bnd))) for d in range(3)])
            for i, (v, abc) in enumerate(emit):
                if c32:
                    if abc is None:
                        pred = [f32(x) for x in fb]
                    else:
                        a, b, c = abc
                        pred = [f32(f32(f32(w32[0] * f32(recon[a][d])) + f32(w32[1] * f32(recon[b][d])))
                                    + f32(w32[2] * f32(recon[c][d]))) for d in range(3)]
                    rec = np.array([float(f32(pred[d] + f32(f32(int(codes[i, d])) * d32)))
                                    for d in range(3)])
                elif abc is None:
                    rec = fb.copy() + codes[i].astype(np.float64) * delta
                else:
                    a, b, c = abc
                    pred = (w0 * recon[a] + w1 * recon[b]) + w2 * recon[c]
                    rec = pred + codes[i].astype(np.float64) * delta
                recon[v] = rec
                verts[l2g[v]] = rec
            info["codes"].append(codes)
        info["emit"].append(emit)
        for (a, b, c) in local_tris:
            tris.append((l2g[a], l2g[b], l2g[c]))
    Vn = np.array(verts, dtype=np.float64).reshape(-1, 3)
    V = (Vn * h["scale"] + np.asarray(h["center"], dtype=np.float64) if world else Vn).astype(np.float32)
    T = np.array(tris, dtype=np.int64).reshape(-1, 3)
    if return_info:
        info["header"] = h
        return V, T, info
    return V, T


def f32_bits(x: float) -> int:
    return struct.unpack(">I", struct.pack(">f", x))[0]
