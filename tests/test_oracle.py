"""Pins the CPU oracle (oracle/stride_oracle.py, oracle/stride_oracle.c).

Checks, in order of strength:
  * golden vectors written by the reference's own encode_dup / decode_dup /
    fit search / _pick_best_k (tests/golden/make_golden.py);
  * the reference's bit-codec known-answer test (utils/bit_codec.py:262-288);
  * hand-derived meshlet bitstreams (FIFO hit/miss/eviction, both edge
    codes, delta/para classification, Rice and Exp-Golomb substreams);
  * brute force (coder costs against emitted bit lengths, IRLP costs);
  * invariants (round trip codes == global_codes[l2g], triangle sets).
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import encode_with_oracle, small_meshes

GOLD = Path(__file__).resolve().parent / "golden"
DUP = sorted(GOLD.glob("dup_*.npz"))
IRLP = sorted(GOLD.glob("irlp_*.npz"))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---------------------------------------------------------------- bit codec
def test_bitcodec_known_answer(oracle):
    """utils/bit_codec.py:262-288 payload, round-tripped."""
    w = oracle.BitWriter()
    w.put(0xABCD, 16)
    w.put_rice(0, 3)
    w.put_rice(1, 0)
    w.put_rice(42, 4)
    w.put_eg(0, 0)
    w.put_eg(7, 2)
    w.put_zigzag(-12345, 16)
    w.put_zigzag(12345, 16)
    w.put_f32(3.14159)
    r = oracle.BitReader(w.tobytes())
    assert r.get(16) == 0xABCD
    assert r.rice(3) == 0
    assert r.rice(0) == 1
    assert r.rice(4) == 42
    assert r.eg(0) == 0
    assert r.eg(2) == 7
    assert r.zigzag(16) == -12345
    assert r.zigzag(16) == 12345
    assert abs(r.f32() - 3.14159) < 1e-5


def test_bitcodec_layouts(oracle):
    """Closed-form codeword layouts, MSB first, bit 7 of byte 0 first."""
    w = oracle.BitWriter()
    w.put_rice(42, 4)          # q = 2 -> 0 0 1, low 1010
    assert w.bits == [0, 0, 1, 1, 0, 1, 0]
    w = oracle.BitWriter()
    w.put_eg(7, 2)             # shifted 1 -> x=2, lb=1: 0 1 0, low 11
    assert w.bits == [0, 1, 0, 1, 1]
    w = oracle.BitWriter()
    w.put_eg(6, 0)             # x = 7, lb = 2: 0 0 1 1 1
    assert w.bits == [0, 0, 1, 1, 1]
    w = oracle.BitWriter()
    w.put(0b101, 3)
    assert w.tobytes() == bytes([0b10100000])


@pytest.mark.parametrize("seed", range(5))
def test_bitcodec_roundtrip_random(oracle, seed):
    rng = np.random.default_rng(seed)
    ops = []
    w = oracle.BitWriter()
    for _ in range(300):
        kind = int(rng.integers(0, 3))
        u = int(rng.integers(0, 1 << int(rng.integers(1, 20))))
        k = int(rng.integers(0, 12))
        if kind == 0:
            n = max(1, u.bit_length())
            w.put(u, n)
        elif kind == 1:
            w.put_rice(u, k)
        else:
            w.put_eg(u, k)
        ops.append((kind, u, k))
    r = oracle.BitReader(w.tobytes())
    for kind, u, k in ops:
        got = r.get(max(1, u.bit_length())) if kind == 0 else (r.rice(k) if kind == 1 else r.eg(k))
        assert got == u


# ------------------------------------------------------------- coder costs
def test_stream_costs_match_emitted_bits(oracle, coracle):
    rng = np.random.default_rng(3)
    for _ in range(40):
        vals = np.round(rng.laplace(0, float(rng.choice([1, 10, 400])), int(rng.integers(1, 80))))
        vals = vals.astype(np.int64)
        costs = oracle.stream_costs(vals)
        for j in range(20):
            w = oracle.BitWriter()
            for u in oracle.zigzag(vals):
                (w.put_rice if j < 12 else w.put_eg)(int(u), j if j < 12 else j - 12)
            assert w.bit_pos() == costs[j]
        assert list(coracle.stream_costs(vals)) == costs


def test_pick_best_k_golden(oracle, coracle):
    z = np.load(GOLD / "coder_choice.npz")
    off = np.concatenate(([0], np.cumsum(z["stream_lens"])))
    for i in range(len(z["stream_lens"])):
        vals = z["stream_vals"][off[i]:off[i + 1]]
        t, k, _ = oracle.pick_tag_k(vals)
        assert (t, k) == (int(z["tags"][i]), int(z["ks"][i]))
        tc, kc, _ = coracle.pick(vals)
        assert (tc, kc) == (t, k)
    assert set(z["tags"].tolist()) == {1, 2}, "golden set must exercise both coders"


def test_irlp_cost_golden(oracle, coracle):
    z = np.load(GOLD / "coder_choice.npz")
    for cand, K, want in zip(z["irlp_cands"], z["irlp_K"], z["irlp_costs"]):
        got = oracle.irlp_cost(z["irlp_a"], z["irlp_b"], z["irlp_c"], z["irlp_v"], z["irlp_off"],
                               cand, int(K))
        assert got == int(want)
        gc = coracle.irlp_cost(z["irlp_a"], z["irlp_b"], z["irlp_c"], z["irlp_v"], z["irlp_off"],
                               cand[None], int(K))
        assert int(gc[0]) == int(want)


def test_irlp_pred_closed_forms(oracle):
    a, b, c = np.array([10, 20, 30]), np.array([7, 9, 11]), np.array([3, 4, 5])
    assert list(oracle.irlp_pred(a, b, c, oracle.CANON_N, oracle.CANON_K)) == [14, 25, 36]
    # K > 0 rounds half up then floors: (s + 2^(K-1)) >> K, negative sums floor
    n = np.array([[1, 0, 0], [1, 1, 0], [0, 0, -1]])
    K = np.array([1, 2, 1])
    assert list(oracle.irlp_pred(np.array([3, 1, 0]), np.array([0, 1, 0]), np.array([0, 0, 3]),
                                 n, K)) == [2, 1, -1]


# ------------------------------------------------------ hand-built streams
def _header_bits(w, g_bits, n_v, n_t, n_m, pred_n=((1, 1, -1),) * 3, pred_K=(0, 0, 0)):
    """DPR4 v3 header written field by field from the layout."""
    w.put(0x44505234, 32)
    w.put(3, 8)
    w.put(0, 8)
    for f in (0.0, 0.0, 0.0, 1.0, 0.001, 0.0, 0.0, 0.0, 2.0, 3.0, 4.0):
        w.put_f32(f)
    for b in g_bits:
        w.put(b, 8)
    for row in pred_n:
        for x in row:
            w.put(x & 0xFFFF, 16)
    for k in pred_K:
        w.put(k, 8)
    w.put(n_v, 32)
    w.put(n_t, 32)
    w.put(n_m, 32)


def _bits(w, s):
    for ch in s.replace(" ", ""):
        w.put(int(ch), 1)


def hand_stream_small():
    """5 vertices, 3 triangles, 2 strips (derivation in the comments)."""
    from oracle.stride_oracle import BitWriter
    body = BitWriter()
    body.put(5, 16); body.put(3, 16); body.put(2, 16)
    body.put(2, 16)                        # strip 1, two triangles
    _bits(body, "1 000  1 001  1 010")     # raw 0, 1, 2 (idx_bits = 3)
    _bits(body, "1  1 011")                # edge 1, raw 3 -> tri (0, 2, 3), refs (0, 2, 1)
    body.put(1, 16)                        # strip 2, one triangle
    _bits(body, "0 00001")                 # FIFO [0,1,2,3] hit 1 -> 1; FIFO [0,2,3,1]
    _bits(body, "0 00010")                 # hit 2 -> 3; FIFO [0,2,1,3]
    _bits(body, "1 100")                   # raw 4 -> tri (1, 3, 4)
    body.put(5, 4); body.put(6, 4); body.put(7, 4)        # anchor (5, 6, 7)
    for tag, k in ((1, 0), (1, 1), (2, 0), (1, 0), (2, 1), (1, 2)):
        body.put(tag, 8); body.put(k, 8)
    _bits(body, "001 01 1")                # dx Rice0 zz(1,-1,0) = 2,1,0
    _bits(body, "0010 10 011")             # dy Rice1 zz(2,0,-2) = 4,0,3
    _bits(body, "1 00111 010")             # dz EG0  zz(0,3,-1) = 0,6,1
    _bits(body, "0000001")                 # px Rice0 zz(3) = 6
    _bits(body, "0101")                    # py EG1  zz(-2) = 3
    _bits(body, "00110")                   # pz Rice2 zz(5) = 10
    w = BitWriter()
    _header_bits(w, (4, 4, 4), 5, 3, 1)
    w.put(688 + 32, 32)
    data = w.tobytes() + body.tobytes()
    codes = np.array([[5, 6, 7], [6, 8, 7], [5, 8, 10], [7, 4, 15], [7, 2, 14]])
    tris = np.array([[0, 1, 2], [0, 2, 3], [1, 3, 4]])
    return data, codes, tris


def hand_stream_fifo():
    """17 raw ids evict id 0 from the 16-entry FIFO; a second strip hits the
    new oldest entry (id 1), re-sends id 0 raw and adds id 17."""
    from oracle.stride_oracle import BitWriter
    body = BitWriter()
    body.put(18, 16); body.put(16, 16); body.put(2, 16)
    body.put(15, 16)
    for v in (0, 1, 2):
        body.put(1, 1); body.put(v, 5)     # idx_bits(18) = 5
    for v in range(3, 17):
        body.put(0, 1)                     # edge 0 -> (p1, p2, new)
        body.put(1, 1); body.put(v, 5)
    body.put(1, 16)
    body.put(0, 1); body.put(0, 5)         # FIFO [1..16]: hit 0 -> 1
    body.put(1, 1); body.put(0, 5)         # raw 0 (evicted earlier)
    body.put(1, 1); body.put(17, 5)        # raw 17
    for _ in range(3):
        body.put(16, 5)                    # anchor 16 per axis
    for _ in range(6):
        body.put(1, 8); body.put(0, 8)
    # n_kind0 = 4 (0, 1, 2, 17): 3 deltas per axis; 14 para per axis; all 0
    for _ in range(3 * 3 + 3 * 14):
        body.put(1, 1)
    w = BitWriter()
    _header_bits(w, (5, 5, 5), 18, 16, 1)
    w.put(688 + 32, 32)
    data = w.tobytes() + body.tobytes()
    codes = np.full((18, 3), 16)
    tris = [[0, 1, 2]] + [[k - 2, k - 1, k] for k in range(3, 17)] + [[1, 0, 17]]
    return data, codes, np.array(tris)


@pytest.mark.parametrize("make", [hand_stream_small, hand_stream_fifo])
def test_hand_streams(oracle, coracle, make):
    data, codes, tris = make()
    V, T, C = oracle.decode_dup(data, return_codes=True)
    assert np.array_equal(C, codes)
    assert np.array_equal(T, tris)
    Vc, Tc, Cc = coracle.decode(data, with_codes=True)
    assert np.array_equal(Cc, codes) and np.array_equal(Tc, tris)
    assert np.array_equal(Vc.view(np.uint32), V.view(np.uint32))
    # dequantisation end points: code 0 -> g_min*scale+center, code mx -> g_range side
    h = oracle.read_header(data)
    mx = (1 << h["g_bits"][0]) - 1
    q = oracle.dequantize(np.array([[0, 0, 0], [mx, mx, mx]]), h)
    assert np.array_equal(q[0], np.float32([0, 0, 0]))
    assert np.array_equal(q[1], np.float32([2, 3, 4]))


def test_empty_stream_decodes_empty(oracle, coracle):
    """A header with n_meshlets = 0 decodes to empty (0, 3) outputs in both
    restatements, as in the GPU decoder.  The reference's decode_dup raises
    ValueError there instead (np.concatenate of no arrays,
    paradelta_v5_dup.py:608), a by-product of its stacking, not a check."""
    h = bytearray(hand_stream_small()[0][:86])
    h[74:86] = bytes(12)
    for dec in (oracle.decode_dup, coracle.decode):
        V, T = dec(bytes(h))
        assert V.shape == (0, 3) and T.shape == (0, 3) and V.dtype == np.float32


def test_decode_errors(oracle):
    data, _, _ = hand_stream_small()
    with pytest.raises(ValueError, match="bad magic"):
        oracle.decode_dup(b"\x00" * 100)
    bad = bytearray(data)
    bad[4] = 2
    with pytest.raises(ValueError, match="version 2"):
        oracle.decode_dup(bytes(bad))
    with pytest.raises(IndexError):
        oracle.decode_dup(data[:len(data) - 3])


# ------------------------------------------------------------ golden files
@pytest.mark.parametrize("path", DUP, ids=[p.stem for p in DUP])
def test_golden_decode(oracle, coracle, path):
    z = np.load(path)
    data = z["data"].tobytes()
    V, T = oracle.decode_dup(data) if len(data) < 20000 else coracle.decode(data)
    assert sha(V) == z["V_sha"].tobytes().hex()
    assert sha(T.astype(np.int64)) == z["T_sha"].tobytes().hex()
    if "V" in z:
        assert np.array_equal(V.view(np.uint32), z["V"].view(np.uint32))
        assert np.array_equal(T, z["T"])
    Vc, Tc = coracle.decode(data, n_threads=3)
    assert sha(Vc) == z["V_sha"].tobytes().hex()
    assert sha(Tc) == z["T_sha"].tobytes().hex()


@pytest.mark.parametrize("path", IRLP, ids=[p.stem for p in IRLP])
def test_golden_decode_irlp(coracle, oracle, path):
    z = np.load(path)
    data = z["data"].tobytes()
    h = oracle.read_header(data)
    assert np.array_equal(h["pred_n"], z["pred_n"]) and np.array_equal(h["pred_K"], z["pred_K"])
    V, T = coracle.decode(data)
    assert sha(V) == z["V_sha"].tobytes().hex()
    assert sha(T) == z["T_sha"].tobytes().hex()
    if "V" in z:
        Vp, Tp = oracle.decode_dup(data)
        assert np.array_equal(Vp.view(np.uint32), z["V"].view(np.uint32))
        assert np.array_equal(Tp, z["T"])


GOLD_MESHES = {"uv_12x24", "torus_30x10", "hf_24", "ico_2", "one_tri"}


@pytest.mark.parametrize("name", sorted(GOLD_MESHES))
def test_golden_encode_bytes(oracle, name):
    """The oracle writer reproduces the reference's encode_dup bytes on the
    same plans, and the fit search reproduces the reference's (n, K)."""
    meshes = {n: (v, t) for n, v, t in small_meshes()}
    z = np.load(GOLD / f"dup_{name}.npz")
    V, T = meshes[name]
    prep, data = encode_with_oracle(V, T)
    assert data == z["data"].tobytes()
    n, K = oracle.fit_predictor(z["samples_A"], z["samples_B"], z["samples_C"], z["samples_V"],
                                z["samples_off"])
    assert np.array_equal(n, z["fit_n"]) and np.array_equal(K, z["fit_K"])


def test_golden_encode_bytes_c1(oracle):
    from meshpress_amd import workloads as w
    z = np.load(GOLD / "dup_uv_c1.npz")
    V, T = w.uv_sphere_c1()
    prep, data = encode_with_oracle(V, T)
    assert data == z["data"].tobytes()


def test_golden_encode_bytes_q16(oracle):
    from meshpress_amd import workloads as w
    z = np.load(GOLD / "dup_torus_q16.npz")
    V, T = w.displaced_torus(60, 20, 1.0, 0.3, 0.05, 4, 11)
    prep, data = encode_with_oracle(V, T, precision_error=1.0 / 65534.5, precision_mode="bbox_frac")
    assert data == z["data"].tobytes()
    assert max(prep["g_bits"]) == 16


# ------------------------------------------------------------- invariants
@pytest.mark.parametrize("pred", ["canonical", "fit"])
def test_roundtrip_invariant(oracle, coracle, pred):
    """decoded codes == global_codes[l2g] and triangle sets equal the source
    (scripts/verify_stride_dup.py:16 PASS criterion)."""
    for name, V, T in small_meshes():
        prep, data = encode_with_oracle(V, T, pred=pred)
        Vd, Td, C = oracle.decode_dup(data, return_codes=True)
        l2g = np.concatenate([pl["local_to_global"] for pl in prep["plans"]]).astype(np.int64)
        assert np.array_equal(C, prep["global_codes"][l2g]), name
        src = np.unique(np.sort(T, axis=1), axis=0)
        dec = np.unique(np.sort(l2g[Td], axis=1), axis=0)
        assert np.array_equal(src, dec), name
        Vc, Tc, Cc = coracle.decode(data, with_codes=True)
        assert np.array_equal(Cc, C) and np.array_equal(Tc, Td)
        assert np.array_equal(Vc.view(np.uint32), Vd.view(np.uint32))


from hypothesis import given, settings, strategies as st  # noqa: E402

_ops = st.lists(st.tuples(st.integers(0, 2), st.integers(0, (1 << 24) - 1), st.integers(0, 11)),
                min_size=1, max_size=60)


@settings(max_examples=200, deadline=None)
@given(_ops)
def test_bitcodec_roundtrip_hypothesis(ops):
    """Fixed-width, Rice(k) and Exp-Golomb(k) fields written back to back
    (MSB first, no alignment) read back exactly, and the stream length is
    the sum of the closed-form field lengths."""
    from oracle import stride_oracle as O
    w = O.BitWriter()
    bits = 0
    for kind, u, k in ops:
        if kind == 0:
            n = max(1, u.bit_length())
            w.put(u, n)
            bits += n
        elif kind == 1:
            w.put_rice(u, k)
            bits += (u >> k) + 1 + k
        else:
            w.put_eg(u, k)
            lb = ((u >> k) + 1).bit_length() - 1
            bits += 2 * lb + 1 + k
    data = w.tobytes()
    assert len(data) == (bits + 7) // 8
    r = O.BitReader(data)
    for kind, u, k in ops:
        got = r.get(max(1, u.bit_length())) if kind == 0 else (r.rice(k) if kind == 1 else r.eg(k))
        assert got == u


def test_threaded_search_costs_equal_serial(coracle):
    """The C oracle's threaded candidate cost and stream picks (the host side
    of the exhaustive-search baseline) equal the serial forms, for every
    thread count, including empty meshlets and streams."""
    rng = np.random.default_rng(3)
    cnt = rng.integers(0, 90, 37)
    cnt[[0, 5]] = 0
    off = np.concatenate(([0], np.cumsum(cnt))).astype(np.int64)
    n = int(off[-1])
    a, b, c = (rng.integers(0, 4096, n) for _ in range(3))
    v = a + b - c + rng.integers(-40, 40, n)
    cands = [[32, 32, -32], [33, 31, -32], [1, 1, -1], [0, 0, 0]]
    for K in (0, 5, 9):
        ser = coracle.irlp_cost(a, b, c, v, off, cands, K)
        for t in (1, 3, 8, 64):
            assert np.array_equal(coracle.irlp_cost_mt(a, b, c, v, off, cands, K, t), ser)
    vals = rng.integers(-300, 300, n)
    for t in (1, 4, 40):
        tagk, bits = coracle.pick_streams(vals, off, t)
        for s in range(len(off) - 1):
            assert (int(tagk[s, 0]), int(tagk[s, 1]), int(bits[s])) == \
                coracle.pick(vals[off[s]:off[s + 1]])
