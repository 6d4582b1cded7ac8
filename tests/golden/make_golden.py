"""Writes golden vectors from the reference's own Python codec functions.

Run in the development container only (it reads /root/reference, which never
reaches a GPU machine):

    python tests/golden/make_golden.py

The reference functions are called on this repository's own plans (the
reference front-end needs numba, which is absent), so the fixtures pin the
parts no closed form pins:

  * encode_dup(prep, predictor="canonical") -> exact DPR4 bytes
    (encoder/paradelta_v5_dup.py:316-464);
  * decode_dup(bytes) -> V float32, T int64, including the triangle corner
    order (encoder/paradelta_v5_dup.py:471-618), also on streams with a
    fitted IRLP predictor;
  * the IRLP fit: _gather + subsample + _do_fit + per-axis canonical fallback
    exactly as fit_predictor (encoder/_irlp_fit.py:364-394) but without its
    sample cache (the cache would be written into the read-only reference
    tree);
  * _pick_best_k (paradelta_v5_dup.py:159-171) and encoder_cost_batch
    (_irlp_fit.py:186-226) on seeded random residuals;
  * decode_paradelta_v5(bytes) -> V float32, T int64 on PDLT v5 streams of
    this repository's writer (encoder/paradelta_v5.py:286-462; the
    reference's v5 writer is Numba code and cannot run here), including
    streams that force each residual coder and explicit partitions with
    meshlets that have no interior or no boundary vertices
    (`python tests/golden/make_golden.py pdlt` writes only these);
  * the DPRB v4 side-table blob of scripts/dump_stride_dup_blob.py:88-152,
    built from encode_dup(return_meta=True)'s own per-meshlet meta (anchor
    offsets, n_kind0, substream offsets) with the layout that script packs
    (`python tests/golden/make_golden.py dprb` writes only these).

Fixtures are data only: inputs (bytes, sample arrays) and expected outputs.
Large outputs are stored as SHA-256 digests.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = Path("/root/reference")
sys.path.insert(0, str(ROOT))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def meshes():
    from meshpress_amd import workloads as w
    out = [("uv_12x24",) + w.displaced_uv_sphere(12, 24, 0.05, 4, 3),
           ("torus_30x10",) + w.displaced_torus(30, 10, 1.0, 0.3, 0.05, 4, 5),
           ("hf_24",) + w.heightfield(24, 0.1, 6, 7),
           ("ico_2",) + w.displaced_icosphere(2, 0.08, 6, 1),
           ("one_tri", np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0.5]]), np.array([[0, 1, 2]])),
           ("uv_c1",) + w.uv_sphere_c1(),
           ("torus_q16",) + w.displaced_torus(60, 20, 1.0, 0.3, 0.05, 4, 11)]
    return out


def main():
    from meshpress_amd.encoder.prep import prepare
    sys.path.insert(0, str(REF))
    from encoder.paradelta_v5_dup import (encode_dup, decode_dup, _pick_best_k,  # noqa: E402
                                          _walk_meshlet)
    import encoder._irlp_fit as F  # noqa: E402

    summary = {}
    for name, V, T in meshes():
        kw = {}
        if name == "torus_q16":
            kw = dict(precision_error=1.0 / 65534.5, precision_mode="bbox_frac")
        p = prepare(V, T, **kw)
        prep = dict(p)
        prep.pop("native")
        plans = p["native"].export_plans()
        for pl in plans:
            pl.pop("walk")
        prep["plans"] = plans
        data = encode_dup(prep, predictor="canonical")
        Vd, Td = decode_dup(data)
        rec = {"data": np.frombuffer(data, np.uint8)}
        # IRLP fit with the reference's functions (fit_predictor minus cache)
        A, B, C, Vs, off = F._gather(prep["global_codes"], plans, _walk_meshlet, use_cache=False)
        n_m = len(off) - 1
        if n_m == 0 or len(A) == 0:
            n_fit, K_fit = F.CANON_N.copy(), F.CANON_K.copy()
        else:
            if n_m > F.SUBSAMPLE_MESHLETS:
                rng = np.random.default_rng(0)
                sel = np.sort(rng.choice(n_m, size=F.SUBSAMPLE_MESHLETS, replace=False))
                rows = np.concatenate([np.arange(off[i], off[i + 1]) for i in sel])
                As, Bs, Cs, Vss = A[rows], B[rows], C[rows], Vs[rows]
                off_s = np.concatenate(([0], np.cumsum(np.diff(off)[sel])))
            else:
                As, Bs, Cs, Vss, off_s = A, B, C, Vs, off
            n_fit, K_fit = F._do_fit(As, Bs, Cs, Vss, off_s)
            for d in range(3):
                b_fit = F.encoder_cost_one(A[:, d], B[:, d], C[:, d], Vs[:, d], off,
                                           n_fit[d], int(K_fit[d]))
                b_can = F.encoder_cost_one(A[:, d], B[:, d], C[:, d], Vs[:, d], off,
                                           F.CANON_N[d], int(F.CANON_K[d]))
                if b_can < b_fit:
                    n_fit[d] = F.CANON_N[d]
                    K_fit[d] = F.CANON_K[d]
        rec["fit_n"] = np.asarray(n_fit, np.int64)
        rec["fit_K"] = np.asarray(K_fit, np.int64)
        small = len(V) <= 5000
        if small:
            rec["V"] = Vd
            rec["T"] = Td.astype(np.int64)
            rec["samples_A"] = A
            rec["samples_B"] = B
            rec["samples_C"] = C
            rec["samples_V"] = Vs
            rec["samples_off"] = off
        rec["V_sha"] = np.frombuffer(bytes.fromhex(sha(Vd)), np.uint8)
        rec["T_sha"] = np.frombuffer(bytes.fromhex(sha(Td.astype(np.int64))), np.uint8)
        np.savez_compressed(HERE / f"dup_{name}.npz", **rec)
        summary[name] = {"bytes": len(data), "n_v": int(len(Vd)), "n_t": int(len(Td)),
                         "fit_n": np.asarray(n_fit).tolist(), "fit_K": np.asarray(K_fit).tolist()}
        print(name, summary[name], flush=True)

    # Generalized-predictor streams: our oracle writes them, the reference decodes.
    from oracle import stride_oracle as O
    for name in ("uv_12x24", "ico_2", "uv_c1"):
        V, T = {n: (v, t) for n, v, t in meshes()}[name]
        p = prepare(V, T)
        prep = dict(p)
        prep["plans"] = p["native"].export_plans()
        z = np.load(HERE / f"dup_{name}.npz")
        n_fit, K_fit = z["fit_n"], z["fit_K"]
        if name == "uv_12x24":   # force a non-canonical predictor with K > 0
            n_fit = np.array([[33, 32, -33], [31, 34, -33], [32, 32, -32]], np.int64)
            K_fit = np.array([5, 5, 5], np.int64)
        data = O.encode_dup(prep, n_fit, K_fit)
        Vd, Td = decode_dup(data)
        rec = {"data": np.frombuffer(data, np.uint8), "pred_n": n_fit, "pred_K": K_fit,
               "V_sha": np.frombuffer(bytes.fromhex(sha(Vd)), np.uint8),
               "T_sha": np.frombuffer(bytes.fromhex(sha(Td.astype(np.int64))), np.uint8)}
        if len(V) <= 5000:
            rec["V"] = Vd
            rec["T"] = Td.astype(np.int64)
        np.savez_compressed(HERE / f"irlp_{name}.npz", **rec)
        print("irlp", name, len(data), flush=True)

    # _pick_best_k and encoder_cost_batch on seeded residuals.
    rng = np.random.default_rng(1234)
    streams = []
    for i in range(64):
        n = int(rng.integers(0, 200))
        scale = float(rng.choice([0.3, 2, 20, 300, 5000]))
        kind = i % 3
        if kind == 0:
            x = np.round(rng.laplace(0, scale, n))
        elif kind == 1:
            x = np.round(rng.normal(0, scale, n))
        else:   # heavy tail: a few outliers make Exp-Golomb win
            x = np.round(rng.laplace(0, 1, n))
            if n:
                x[rng.integers(0, n, max(1, n // 20))] = rng.integers(-50000, 50000, max(1, n // 20))
        streams.append(x.astype(np.int64))
    lens = np.array([len(s) for s in streams], np.int64)
    flat = np.concatenate(streams)
    tags, ks = [], []
    for s in streams:
        t, k = _pick_best_k(s.reshape(-1, 1).repeat(3, axis=1))
        tags.append(t[0])
        ks.append(k[0])
    m_off = np.concatenate(([0], np.cumsum(rng.integers(1, 60, 80))))
    ns = int(m_off[-1])
    a, b, c = (rng.integers(0, 4096, ns) for _ in range(3))
    v = np.clip(a + b - c + np.round(rng.laplace(0, 3, ns)).astype(np.int64), 0, 4095)
    cands = np.array([[1, 1, -1], [32, 32, -32], [33, 31, -32], [30, 35, -33], [64, 64, -64],
                      [70, 60, -66]], np.int64)
    Ks = np.array([0, 5, 5, 5, 6, 6], np.int64)
    costs = np.array([int(F.encoder_cost_batch(a, b, c, v, m_off, cands[i:i + 1], int(Ks[i]))[0])
                      for i in range(len(cands))], np.int64)
    np.savez_compressed(HERE / "coder_choice.npz", stream_lens=lens, stream_vals=flat,
                        tags=np.array(tags, np.int64), ks=np.array(ks, np.int64),
                        irlp_a=a, irlp_b=b, irlp_c=c, irlp_v=v, irlp_off=m_off,
                        irlp_cands=cands, irlp_K=Ks, irlp_costs=costs)
    (HERE / "summary.json").write_text(json.dumps(summary, indent=1))
    print("coder_choice", list(zip(tags, ks))[:10])


def dprb_goldens():
    """DPRB blobs from the reference encoder's meta, for two small meshes."""
    import struct
    from meshpress_amd.encoder.prep import prepare
    sys.path.insert(0, str(REF))
    from encoder.paradelta_v5_dup import encode_dup  # noqa: E402
    for name, V, T in meshes():
        if name not in ("torus_30x10", "uv_c1"):
            continue
        p = prepare(V, T)
        prep = dict(p)
        prep.pop("native")
        plans = p["native"].export_plans()
        for pl in plans:
            pl.pop("walk")
        prep["plans"] = plans
        data, meta = encode_dup(prep, predictor="canonical", return_meta=True)
        n_m = int(prep["n_meshlets"])
        head = 86 + 4 * n_m                       # DPR4 header + offset table, bytes
        body = data[head:]
        offs = [struct.unpack(">I", data[86 + 4 * i:90 + 4 * i])[0] - 8 * head for i in range(n_m)]
        hdr = {k: data[a:b] for k, a, b in (("f", 6, 50),)}
        f11 = struct.unpack(">11f", hdr["f"])
        g_bits = list(data[50:53])
        pred_n = [struct.unpack(">h", data[53 + 2 * i:55 + 2 * i])[0] for i in range(9)]
        pred_K = list(data[71:74])
        n_local = np.array([len(pl["local_to_global"]) for pl in plans], np.uint32)
        n_tris = np.array([pl["n_tris_m"] for pl in plans], np.uint32)
        n_strips = np.array([pl["n_strips"] for pl in plans], np.uint32)
        v_off = np.concatenate([[0], np.cumsum(n_local[:-1])]).astype(np.uint32)
        t_off = np.concatenate([[0], np.cumsum(n_tris[:-1])]).astype(np.uint32)
        out = bytearray()
        out += struct.pack("<II", 0x42525044, 4)
        out += struct.pack("<III", n_m, int(n_local.sum()), int(n_tris.sum()))
        out += struct.pack("<I", len(body)) + struct.pack("<II", 0, 0)
        out += struct.pack("<11f", *f11) + struct.pack("<III", *g_bits)
        out += struct.pack("<9h", *pred_n) + struct.pack("<3B", *pred_K) + b"\x00\x00\x00"
        out += body
        out += np.array(offs, np.uint64).tobytes()
        for a in (n_local, n_tris, n_strips, v_off, t_off):
            out += a.tobytes()
        out += np.asarray(meta["resid_offs_rel"], np.uint64).tobytes()
        out += np.asarray(meta["n_kind0"], np.uint32).tobytes()
        out += np.asarray(meta["axis_sub_offs"], np.uint16).tobytes()
        np.savez_compressed(HERE / f"dprb_{name}.npz", dpr4=np.frombuffer(data, np.uint8),
                            dprb=np.frombuffer(bytes(out), np.uint8))
        print("dprb", name, len(data), len(out), flush=True)


def pdlt_cases():
    """(name, data) PDLT v5 streams of the oracle writer."""
    from meshpress_amd.encoder.prep import prepare
    from oracle import pdlt_oracle as PO
    out = []
    ms = {n: (v, t) for n, v, t in meshes()}
    for name in ("uv_12x24", "torus_q16", "hf_24", "one_tri", "uv_c1"):
        V, T = ms[name]
        kw = dict(precision_error=1.0 / 65534.5, precision_mode="bbox_frac") \
            if name == "torus_q16" else {}
        p = prepare(V, T, **kw)
        prep = dict(p)
        prep["plans"] = p["native"].export_plans()
        out.append((name, PO.encode_v5(prep)))
        if name in ("hf_24", "torus_q16"):
            out.append((name + "_tags", PO.encode_v5(prep, force_tags=lambda m: [
                m % 3, (m + 1) % 3, (m + 2) % 3])))
    # explicit partitions: a fan whose middle meshlet is all boundary, plus a
    # separate triangle (a meshlet without boundary vertices)
    V = np.array([[0.0, 0, 0], [1, 0, 0.1], [1, 1, 0.2], [0, 1, 0.3], [-1, 1, 0.1], [-1, 0, 0.2],
                  [5, 5, 5], [6, 5, 5.5], [5, 6, 4.5]])
    T = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 5], [6, 7, 8]])
    p = prepare(V, T)
    prep = dict(p)
    prep["plans"] = PO.plans_from_partition(T, [[0, 1], [2], [3], [4]], p["global_codes"])
    out.append(("fan_parts", PO.encode_v5(prep)))
    return out


def pdlt_goldens():
    sys.path.insert(0, str(REF))
    from encoder.paradelta_v5 import decode_paradelta_v5  # noqa: E402
    for name, data in pdlt_cases():
        Vd, Td = decode_paradelta_v5(data)
        rec = {"data": np.frombuffer(data, np.uint8),
               "V_sha": np.frombuffer(bytes.fromhex(sha(Vd)), np.uint8),
               "T_sha": np.frombuffer(bytes.fromhex(sha(Td.astype(np.int64))), np.uint8)}
        if len(Vd) <= 5000:
            rec["V"] = Vd
            rec["T"] = Td.astype(np.int64)
        np.savez_compressed(HERE / f"pdlt_{name}.npz", **rec)
        print("pdlt", name, len(data), Vd.shape, Td.shape, flush=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["dprb"]:
        dprb_goldens()
    elif sys.argv[1:] == ["pdlt"]:
        pdlt_goldens()
    else:
        main()
        dprb_goldens()
        pdlt_goldens()
