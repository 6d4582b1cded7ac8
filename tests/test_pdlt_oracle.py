"""CPU tests of the PDLT v5 oracle (oracle/pdlt_oracle.py) and of the
native v5 writer against it.  Golden vectors come from the reference's own
decode_paradelta_v5 (tests/golden/make_golden.py pdlt)."""
import hashlib
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle import pdlt_oracle as PO
from tests.conftest import small_meshes

GOLD = Path(__file__).resolve().parent / "golden"
PDLT_GOLDEN = sorted(GOLD.glob("pdlt_*.npz"))


def _prep(V, T, **kw):
    from meshpress_amd.encoder.prep import prepare
    p = prepare(V, T, **kw)
    prep = dict(p)
    prep["plans"] = p["native"].export_plans()
    return p, prep


def _v5_order(prep):
    """Global vertex id of each decoded vertex: the boundary table, then the
    interior locals of each meshlet (decode_paradelta_v5's output order)."""
    pv = PO.prepare_v5(prep)
    ids = list(pv["boundary_list"])
    for pl in prep["plans"]:
        ids += [int(g) for g in pl["local_to_global"][pl["n_bnd"]:]]
    return np.array(ids, dtype=np.int64)


@pytest.mark.parametrize("path", PDLT_GOLDEN, ids=lambda p: p.stem)
def test_golden_decode_v5(path):
    z = np.load(path)
    data = z["data"].tobytes()
    V, T = PO.decode_v5(data)
    assert hashlib.sha256(np.ascontiguousarray(V).tobytes()).digest() == z["V_sha"].tobytes()
    assert hashlib.sha256(np.ascontiguousarray(T.astype(np.int64)).tobytes()).digest() == \
        z["T_sha"].tobytes()
    if "V" in z:
        assert np.array_equal(V.view(np.uint32), z["V"].view(np.uint32))
        assert np.array_equal(T, z["T"])


def test_golden_covers_every_coder_and_edge_meshlet():
    tags, n_bnd0, n_int0 = set(), 0, 0
    for path in PDLT_GOLDEN:
        data = np.load(path)["data"].tobytes()
        _, _, info = PO.decode_v5(data, return_info=True)
        r = PO.BitReader(data, info["header"]["meshlets_bit"])
        for m, off in enumerate(info["header"]["offsets"]):
            r.pos = off
            n_bnd, n_int = r.get(16), r.get(16)
            n_bnd0 += n_bnd == 0
            n_int0 += n_int == 0
        # the residual tags: re-walk through the oracle's decode of each meshlet
        tags |= _tags_of(data, info)
    assert tags == {0, 1, 2}
    assert n_bnd0 > 0 and n_int0 > 0


def _tags_of(data, info):
    """Residual tags used by a stream (the byte after each meshlet's strips)."""
    h = info["header"]
    out = set()
    n_m = h["n_meshlets"]
    ends = list(h["offsets"][1:]) + [8 * len(data)]
    for m in range(n_m):
        # decode the meshlet with a fresh reader to find where its residuals start
        r = PO.BitReader(data, h["offsets"][m])
        n_bnd, n_int, _nt, n_strips = (r.get(16) for _ in range(4))
        if n_bnd > 0:
            r.get(32)
            if n_bnd > 1:
                k = r.get(8)
                for _ in range(n_bnd - 1):
                    r.rice(k)
        from collections import deque
        fifo = deque(maxlen=PO.FIFO_SIZE)
        ib = PO.idx_bits_for(n_bnd + n_int)
        for _ in range(n_strips):
            L = r.get(16)
            for _ in range(3):
                PO.read_vert(r, fifo, ib)
            for _ in range(L - 1):
                r.get(1)
                PO.read_vert(r, fifo, ib)
        if n_int == 0:
            continue
        for d in range(3):
            tag = r.get(8)
            out.add(tag)
            if tag == 0:
                r.get(16)
                bw = r.get(8)
                r.pos += n_int * bw
            else:
                k = r.get(8)
                for _ in range(n_int):
                    r.rice(k) if tag == 1 else r.eg(k)
        assert r.pos <= ends[m]
    return out


@pytest.mark.parametrize("name,V,T", small_meshes(), ids=[m[0] for m in small_meshes()])
def test_roundtrip_error_bound(name, V, T):
    p, prep = _prep(V, T)
    data = PO.encode_v5(prep)
    Vd, Td = PO.decode_v5(data)                    # decode_paradelta_v5's float64 chain
    V32, T32 = PO.decode_v5(data, chain="f32")     # the encoder's own float32 chain
    ids = _v5_order(prep)
    assert len(ids) == len(Vd) == len(V)
    assert sorted(ids.tolist()) == list(range(len(V)))
    # per-axis bound: half a step of the world grid plus float32 slack.  The
    # encoder quantises each residual against its float32 chain, so that
    # chain's positions meet it; the float64 chain differs from them by the
    # rounding drift along the LIN3 chains (0.03 of the bound on the ribbon,
    # whose 256-vertex meshlets make the longest chains here)
    half = p["per_coord_err"] * p["scale"]
    bound = half * (1 + 1e-3) + 4 * np.finfo(np.float32).eps * (np.abs(V).max() + 1)
    err = np.abs(V32.astype(np.float64) - V[ids]).max()
    assert err <= bound, (err, bound)
    drift = np.abs(Vd.astype(np.float64) - V32.astype(np.float64)).max()
    assert drift <= 0.1 * half, (drift, half)
    assert np.array_equal(Td, T32)
    got = sorted(map(tuple, np.sort(ids[Td], axis=1).tolist()))
    assert got == sorted(map(tuple, np.sort(T, axis=1).tolist()))


def test_native_writer_matches_oracle_bytes():
    from meshpress_amd.encoder.paradelta_v5 import encode_from_prepared_v5
    from meshpress_amd import workloads as w
    cases = [(m[1], m[2], {}) for m in small_meshes()]
    cases.append(w.displaced_torus(60, 20, 1.0, 0.3, 0.05, 4, 11) +
                 (dict(precision_error=1.0 / 65534.5, precision_mode="bbox_frac"),))
    for V, T, kw in cases:
        p, prep = _prep(V, T, **kw)
        d_nat, info = encode_from_prepared_v5(p, return_info=True)
        d_or, oi = PO.encode_v5(prep, return_info=True)
        assert d_nat == d_or
        assert np.array_equal(info["lin3"], oi["lin3"])


def test_fallback_centroid_order_matches_numpy_mean():
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 8, 9, 16, 31, 64, 129, 200):
        rows = rng.standard_normal((n, 3)) * rng.choice([1e-3, 1.0, 1e3], (n, 3))
        s = np.zeros(3)
        for r in rows:
            s = s + r
        seq = s / n
        ref = np.mean([r for r in rows], axis=0)
        assert np.array_equal(seq, ref)


def test_morton_key_bit_interleave():
    rng = np.random.default_rng(2)
    for _ in range(200):
        c = rng.integers(0, 1 << 21, 3)
        want = 0
        for b in range(21):
            for d in range(3):
                want |= ((int(c[d]) >> b) & 1) << (3 * b + d)
        assert PO.morton3(c) == want
    # bits above 21 are ignored
    assert PO.morton3([1 << 21, 0, 0]) == 0


def test_boundary_list_stable_morton_order():
    codes = np.array([[1, 1, 1], [0, 0, 0], [1, 1, 1], [0, 0, 0], [5, 0, 0], [0, 0, 0]])
    plans = [{"local_to_global": [0, 1, 2, 3, 4]}, {"local_to_global": [0, 1, 2, 3, 5]},
             {"local_to_global": [4, 5]}]
    bl = PO.boundary_list(plans, codes)
    # shared ids {0,1,2,3,4,5}; keys: 0->7, 1->0, 2->7, 3->0, 4->key(5), 5->0
    key = {g: PO.morton3(codes[g]) for g in range(6)}
    assert bl == sorted(range(6), key=lambda g: (key[g], g))


def test_fit_linear3_exact():
    rng = np.random.default_rng(3)
    w = np.array([0.75, 0.5, -0.25])
    A, B, C = (rng.standard_normal((50, 3)) for _ in range(3))
    T = w[0] * A + w[1] * B + w[2] * C
    got = PO.fit_linear3(list(zip(A, B, C, T)))
    assert np.allclose(got, w.astype(np.float32), atol=1e-6)
    assert np.array_equal(PO.fit_linear3([]), np.array([1, 1, -1], np.float32))


@pytest.mark.parametrize("seed", range(6))
def test_axis_coder_is_the_cheapest(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 120))
    kind = seed % 3
    if kind == 0:
        vals = rng.integers(-3, 4, n)           # narrow uniform: fixed width
    elif kind == 1:
        vals = np.round(rng.laplace(0, 6, n)).astype(np.int64)
    else:
        vals = np.round(rng.laplace(0, 1, n)).astype(np.int64)
        vals[:: max(1, n // 7)] = rng.integers(-40000, 40000, len(vals[:: max(1, n // 7)]))
    tag, p, total = PO.axis_coder(vals)
    # brute force: every coder's emitted size
    sizes = {}
    mn, mx = int(vals.min()), int(vals.max())
    bw = max(1, (mx - mn + 1).bit_length()) if mx > mn else 1
    if -32768 <= mn <= 32767:
        sizes[0] = 32 + n * bw
    u = PO.zigzag(vals)
    for t, kmax, put in ((1, 11, "put_rice"), (2, 7, "put_eg")):
        best = None
        for k in range(kmax + 1):
            w = PO.BitWriter()
            for x in u.tolist():
                getattr(w, put)(int(x), k)
            if best is None or w.bit_pos() < best:
                best = w.bit_pos()
        sizes[t] = 16 + best
    low = min(sizes.values())
    assert total == low
    assert tag == min(t for t, s in sizes.items() if s == low)


def test_header_layout_bytes():
    V, T = small_meshes()[0][1:]
    p, prep = _prep(V, T)
    data, info = PO.encode_v5(prep, return_info=True)
    assert struct.unpack(">I", data[0:4])[0] == 0x504C4454
    assert data[4] == 5 and data[5] == 0
    center = struct.unpack(">3f", data[6:18])
    assert np.array_equal(np.float32(center), np.asarray(p["center"]).astype(np.float32))
    scale, pce = struct.unpack(">2f", data[18:26])
    assert scale == np.float32(p["scale"]) and pce == np.float32(p["per_coord_err"])
    assert list(data[50:53]) == [int(b) for b in p["g_bits"]]
    n_v, n_t, n_b, n_m = struct.unpack(">4I", data[53:69])
    assert (n_v, n_t, n_m) == (len(V), len(T), len(prep["plans"]))
    assert n_b == info["n_boundary"]
    assert np.array_equal(np.float32(struct.unpack(">3f", data[69:81])), info["lin3"])
    h = PO.read_header(data)
    # offsets are absolute bit positions of byte-aligned meshlet region
    assert h["offsets"][0] == h["meshlets_bit"] and h["meshlets_bit"] % 8 == 0


def test_decode_errors():
    V, T = small_meshes()[0][1:]
    _, prep = _prep(V, T)
    data = bytearray(PO.encode_v5(prep))
    bad = bytearray(data)
    bad[0] ^= 1
    with pytest.raises(ValueError, match="bad magic"):
        PO.decode_v5(bytes(bad))
    bad = bytearray(data)
    bad[4] = 4
    with pytest.raises(ValueError, match="expected v5, got v4"):
        PO.decode_v5(bytes(bad))
    with pytest.raises(IndexError):
        PO.decode_v5(bytes(data[: len(data) // 2]))


@pytest.mark.parametrize("q16", [False, True], ids=["q12", "q16"])
def test_f32_chain_reproduces_the_encoder_reconstruction(q16):
    """decode_v5(chain="f32") rebuilds exactly the float32 positions the
    encoder's chain produced (interior_pass, paradelta_v5_nb.py:20-212), so
    it stays within the quantisation bound; the float64 chain of
    decode_paradelta_v5 drifts from them by float32 rounding."""
    from meshpress_amd import workloads as w
    kw = dict(precision_error=1.0 / 65534.5, precision_mode="bbox_frac") if q16 else {}
    V, T = w.displaced_torus(60, 20, 1.0, 0.3, 0.05, 4, 11)
    p, prep = _prep(V, T, **kw)
    data, info = PO.encode_v5(prep, return_info=True)
    pv = PO.prepare_v5(prep)
    rec = {}
    for pl in pv["plans"]:
        PO.interior_pass(pl, pv["vn"], pv["bnd_recon_norm"], 2.0 * float(pv["per_coord_err"]),
                         info["lin3"].astype(np.float64), rec)
    Vn, _ = PO.decode_v5(data, world=False, chain="f32")
    ids = _v5_order(prep)
    nb = pv["n_boundary"]
    want = np.array([rec[g] for g in ids[nb:]], np.float32).reshape(-1, 3)
    assert np.array_equal(Vn[nb:].view(np.uint32), want.view(np.uint32))
    assert np.array_equal(Vn[:nb], pv["bnd_recon_norm"][ids[:nb]].astype(np.float32))
    Vw, _ = PO.decode_v5(data, chain="f32")
    step = p["per_coord_err"] * p["scale"]
    # the encoder rounds to the grid against its float32 reconstruction:
    # half a step (2 pce) plus float32 rounding of coordinates of size ~1
    assert np.abs(Vw.astype(np.float64) - V[ids]).max() <= step + 4 * 2.0 ** -24 * p["scale"] * 4


def _meshlet_counts(data):
    h = PO.read_header(data)
    rows = []
    for off in h["offsets"]:
        r = PO.BitReader(data, off)
        rows.append([r.get(16) for _ in range(4)])
    return h, np.array(rows, dtype=np.int64).reshape(-1, 4)


@pytest.mark.parametrize("path", PDLT_GOLDEN, ids=lambda p: p.stem)
def test_sblb_writer_layout(path):
    """The SBLB v1 blob (scripts/dump_stride_blob.py:1-114 layout) holds the
    stream and the tables the reference bench uploads, each recomputed here
    from the oracle's parse."""
    from meshpress_amd.utils.sblb import to_sblb
    data = np.load(path)["data"].tobytes()
    blob = to_sblb(data)
    h, cnt = _meshlet_counts(data)
    n_m, n_b = h["n_meshlets"], h["n_boundary"]
    V, T = PO.decode_v5(data)
    hdr = struct.unpack("<8I4f4I", blob[:64])
    assert hdr[0] == 0x424C4253 and hdr[1] == 1
    assert hdr[2] == h["code_width"] and hdr[3] == n_m and hdr[4] == n_b
    assert hdr[5] == len(V) == n_b + cnt[:, 1].sum() and hdr[6] == len(T) and hdr[7] == len(data)
    assert np.array_equal(np.float32(hdr[8:11]), np.float32(h["lin3"]))
    assert np.float32(hdr[11]) == np.float32(2.0 * h["per_coord_err"])
    code_bytes = 4 if h["code_width"] else 2
    nl, nb, ni, nt = ((cnt[:, 0] + cnt[:, 1]).max(), cnt[:, 0].max(), cnt[:, 1].max(), cnt[:, 2].max()) \
        if n_m else (0, 0, 0, 0)
    warp = int(nl) * 16 + int(nb) * 4 + 32 + int(ni) * 3 * code_bytes + int(ni) * 8 + int(nt) * 3
    assert hdr[12] == (warp + 15) & ~15 and hdr[13] == 1
    p = 64
    assert blob[p:p + len(data)] == data
    p += len(data)
    off = np.frombuffer(blob, np.uint64, n_m, p)
    assert np.array_equal(off, np.array(h["offsets"], np.uint64))
    p += 8 * n_m
    for j in range(4):
        assert np.array_equal(np.frombuffer(blob, np.int32, n_m, p), cnt[:, j])
        p += 4 * n_m
    tris_off = np.frombuffer(blob, np.int32, n_m + 1, p)
    assert np.array_equal(tris_off, np.concatenate([[0], np.cumsum(cnt[:, 2])]))
    p += 4 * (n_m + 1)
    cur = np.frombuffer(blob, np.int32, n_m, p)
    assert np.array_equal(cur, np.concatenate([[0], np.cumsum(cnt[:, 1])[:-1]]) if n_m else cur)
    p += 4 * n_m
    bnd = np.frombuffer(blob, np.float32, 3 * n_b, p).reshape(-1, 3)
    want = PO.dequant(h["bnd_codes"], h["g_min"], h["g_range"], h["g_bits"]).astype(np.float32)
    assert np.array_equal(bnd.view(np.uint32), want.view(np.uint32))
    assert p + 12 * n_b == len(blob)


def test_sblb_rejects_inconsistent_tables():
    """Checked on the host before any device work: a table that disagrees
    with the embedded stream is STRIDE_E_CORRUPT."""
    import ctypes as C
    from meshpress_amd import _lib
    from meshpress_amd.utils.sblb import to_sblb
    data = np.load(PDLT_GOLDEN[0])["data"].tobytes()
    blob = bytearray(to_sblb(data))
    L = _lib.lib()
    for pos in (8, 20, 44, 64 + len(data) + 3, len(blob) - 2):
        bad = np.frombuffer(bytes(blob[:pos]) + bytes([blob[pos] ^ 1]) + bytes(blob[pos + 1:]), np.uint8)
        h = C.c_void_p()
        rc = L.stride_decoder_create_sblb(_lib.ptr(bad), len(bad), C.byref(h))
        assert rc == _lib.E_CORRUPT, (pos, rc, _lib.last_error())
    short = np.frombuffer(bytes(blob[:-1]), np.uint8)
    h = C.c_void_p()
    assert L.stride_decoder_create_sblb(_lib.ptr(short), len(short), C.byref(h)) == -12
