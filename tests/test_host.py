"""Host logic on the CPU: the C-ABI library loads and exports what
include/stride_hip.h declares, the native front-end's invariants, the native
writer against the oracle writer, and host-side error paths (no GPU work)."""
import re
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import encode_with_oracle, small_meshes

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "stride_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|void)\s+(stride_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_declared_symbols():
    import ctypes
    from meshpress_amd import _lib
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_pyset_order_matches_cpython():
    import ctypes as C
    from meshpress_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(0)
    for _ in range(300):
        n = int(rng.integers(1, 9))
        hi = int(rng.choice([8, 40, 256, 70000]))
        ins = [int(x) for x in rng.integers(0, hi, n)]
        s = set()
        for x in ins:
            s.add(x)
        arr = np.array(ins, np.int32)
        out = np.zeros(16, np.int32)
        k = L.stride_debug_pyset_order(_lib.ptr(arr), n, _lib.ptr(out))
        assert list(out[:k]) == list(s), ins


def test_frontend_invariants():
    from meshpress_amd.encoder.prep import prepare
    for name, V, T in small_meshes():
        p = prepare(V, T)
        plans = p["native"].export_plans()
        seen_tris = []
        for pl in plans:
            n_local = len(pl["local_to_global"])
            assert 1 <= n_local <= 256 and 1 <= pl["n_tris_m"] <= 256
            assert len(set(pl["local_to_global"])) == n_local
            # every triangle appears in exactly one strip position
            flat = [t for s in pl["strips"] for t in s]
            assert sorted(flat) == list(range(pl["n_tris_m"]))
            # walk covers every local vertex once, roots first
            walk = pl["walk"]
            assert sorted(walk[:, 0].tolist()) == list(range(n_local))
            assert walk[0, 1] == 0
            l2g = np.asarray(pl["local_to_global"])
            seen_tris.append(np.sort(l2g[pl["ml_tris_local"]], axis=1))
        got = np.unique(np.concatenate(seen_tris), axis=0)
        assert np.array_equal(got, np.unique(np.sort(T, axis=1), axis=0)), name
        assert len(np.concatenate(seen_tris)) == len(T)


def test_native_writer_matches_oracle_writer(oracle):
    from meshpress_amd.encoder.prep import header_f32
    for name, V, T in small_meshes():
        for pred in ("canonical", ([[33, 32, -33], [31, 34, -33], [32, 32, -32]], [5, 5, 5])):
            prep, data = encode_with_oracle(V, T, pred=pred)
            nat = prep["native"]
            tagk = []
            for pl in prep["plans"]:
                _, d, pa = oracle.meshlet_residuals(pl, prep["global_codes"], prep["pred_n"],
                                                    prep["pred_K"])
                row = []
                for arr in (d, pa):
                    for ax in range(3):
                        t, k, _ = oracle.pick_tag_k(arr[:, ax])
                        row += [t, k]
                tagk.append(row)
            vals, off, _ = nat.residuals(prep["pred_n"], prep["pred_K"])
            got = nat.write(header_f32(prep), prep["g_bits"], prep["pred_n"], prep["pred_K"],
                            np.array(tagk, np.uint8))
            assert got == data, name


def test_native_residual_streams(oracle):
    """Residual streams exported for the GPU scorer equal the oracle's."""
    V, T = small_meshes()[0][1:]
    prep, _ = encode_with_oracle(V, T)
    vals, off, mx = prep["native"].residuals(oracle.CANON_N, oracle.CANON_K)
    for m, pl in enumerate(prep["plans"]):
        _, d, pa = oracle.meshlet_residuals(pl, prep["global_codes"], oracle.CANON_N, oracle.CANON_K)
        for s in range(6):
            want = (d if s < 3 else pa)[:, s % 3]
            got = vals[off[6 * m + s]:off[6 * m + s + 1]]
            assert np.array_equal(got, want)


def test_para_samples(oracle):
    V, T = small_meshes()[1][1:]
    prep, _ = encode_with_oracle(V, T)
    A, B, C, Vs, off = prep["native"].para_samples()
    z = np.load(ROOT / "tests" / "golden" / "dup_torus_30x10.npz")
    assert np.array_equal(A, z["samples_A"]) and np.array_equal(Vs, z["samples_V"])
    assert np.array_equal(B, z["samples_B"]) and np.array_equal(C, z["samples_C"])
    assert np.array_equal(off, z["samples_off"])
    # the same samples as axis-major columns, whole and for a meshlet selection
    from meshpress_amd.encoder import search
    nat = prep["native"]
    cols, off2 = nat.para_columns()
    assert np.array_equal(off2, off) and np.array_equal(cols, search.as_columns(A, B, C, Vs))
    sel = np.array([0, 2, 3, nat.n_meshlets - 1])
    sel = np.unique(sel[sel < nat.n_meshlets])
    cs, os_ = nat.para_columns(sel)
    rows = np.concatenate([np.arange(off[i], off[i + 1]) for i in sel])
    assert np.array_equal(os_, np.concatenate(([0], np.cumsum(np.diff(off)[sel]))))
    assert np.array_equal(cs, search.as_columns(A[rows], B[rows], C[rows], Vs[rows]))
    F = nat.features(0)
    assert np.array_equal(nat.features(0, sel), F[sel])
    with pytest.raises(ValueError):
        nat.para_columns(np.array([nat.n_meshlets]))


def test_subsample_selection_matches_row_subsample():
    from meshpress_amd.encoder import search
    rng = np.random.default_rng(3)
    counts = rng.integers(0, 9, 5000)
    off = np.concatenate(([0], np.cumsum(counts)))
    n = int(off[-1])
    A, B, C, V = (rng.integers(0, 4096, (n, 3)).astype(np.int32) for _ in range(4))
    As, Bs, Cs, Vs, offs = search.subsample(A, B, C, V, off)
    sel = search.subsample_sel(5000)
    assert len(sel) == search.SUBSAMPLE_MESHLETS and np.all(np.diff(sel) > 0)
    rows = np.concatenate([np.arange(off[i], off[i + 1]) for i in sel])
    assert np.array_equal(As, A[rows]) and np.array_equal(offs, np.concatenate(([0], np.cumsum(counts[sel]))))
    assert search.subsample_sel(4000) is None


def test_quantiser_rounds_half_even():
    from meshpress_amd.encoder.prep import quantize_global
    vn = np.array([[0.0, 0, 0], [1.0, 1, 1], [0.5 / 3, 0.5, 2.5 / 3]])
    codes, gmin, grng, gbits = quantize_global(vn, 1.0 / 3 / 2)   # step 1/3 -> 2 bits, mx 3
    assert list(gbits) == [2, 2, 2]
    # 0.5 -> 1.5 * ... x = vn * 3: 0.5, 1.5, 2.5 -> half-even 0, 2, 2
    assert codes[2].tolist() == [0, 2, 2]


def test_decoder_rejects_bad_headers_before_device_use():
    """Header validation runs on the host; these raise without a GPU."""
    import ctypes as C
    from meshpress_amd import _lib
    L = _lib.lib()
    h = C.c_void_p()
    bad = np.frombuffer(b"\x00" * 100, np.uint8)
    rc = L.stride_decoder_create(_lib.ptr(bad), len(bad), C.byref(h))
    assert rc == _lib.E_MAGIC and "bad magic 0x00000000" in _lib.last_error()
    from tests.test_oracle import hand_stream_small
    data = bytearray(hand_stream_small()[0])
    data[4] = 5
    arr = np.frombuffer(bytes(data), np.uint8)
    rc = L.stride_decoder_create(_lib.ptr(arr), len(arr), C.byref(h))
    assert rc == _lib.E_VERSION and _lib.last_error() == "expected v3, got v5"
    arr = np.frombuffer(bytes(hand_stream_small()[0][:50]), np.uint8)
    rc = L.stride_decoder_create(_lib.ptr(arr), len(arr), C.byref(h))
    assert rc == _lib.E_TRUNCATED
    with pytest.raises(IndexError):
        _lib.check(rc)


def test_error_messages_are_per_thread():
    """The last error is per thread: eight threads reject their own bad
    headers at once (ctypes releases the GIL around each call) and each
    reads back its own message."""
    import ctypes as C
    import threading
    from meshpress_amd import _lib
    L = _lib.lib()
    n_threads, reps = 8, 300
    start = threading.Barrier(n_threads)
    wrong = []

    def work(t):
        magic = 0x10203040 + 0x01010101 * t
        arr = np.frombuffer(magic.to_bytes(4, "big") + b"\x03" + b"\x00" * 95, np.uint8)
        want = "bad magic 0x%08X" % magic
        h = C.c_void_p()
        start.wait()
        for _ in range(reps):
            rc = L.stride_decoder_create(_lib.ptr(arr), len(arr), C.byref(h))
            msg = _lib.last_error()
            if rc != _lib.E_MAGIC or msg != want:
                wrong.append((t, rc, msg))

    threads = [threading.Thread(target=work, args=(t,)) for t in range(n_threads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not wrong, wrong[:5]


def test_encoder_argument_errors():
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.encoder.prep import prepare
    with pytest.raises(ValueError):
        STRIDEEncoder(selector="nope")
    V, T = small_meshes()[0][1:]
    with pytest.raises(ValueError, match="precision_mode"):
        prepare(V, T, precision_mode="cube")
    with pytest.raises(ValueError, match="gen_method"):
        prepare(V, T, gen_method="meshopt")
    with pytest.raises(IndexError):      # numpy indexing, as the reference front-end
        prepare(V, np.array([[0, 1, 99999]]))


# ------------------------------------------------- kernel logic on the CPU
LANE_SRC = ROOT / "tests" / "lane_check.cpp"


@pytest.fixture(scope="module")
def lane_check(tmp_path_factory):
    """Host build (g++, ASan + UBSan) of the kernels' per-lane logic."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("lane") / "lane_check"
    subprocess.run([gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-ffp-contract=off", str(LANE_SRC),
                    "-o", str(exe)], check=True)
    return exe


def test_codeword_fast_path_and_irlp24(tmp_path):
    """The axis lanes' 32-bit codeword fast path equals the general stream
    decoder (value and length, every q, k and coder, long codewords
    rejected), irlp24 equals irlp64 wherever irlp32_exact allows it, and
    quot_mx equals the float64 division for every 16-bit code and g_bits."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "codeword_check"
    subprocess.run([gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", str(ROOT / "tests" / "codeword_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fast path" in r.stdout and "quotients 1048576" in r.stdout


def run_lane_check(exe, data, tmp_path, split=False):
    """split: False (one lane per meshlet), True or "split" (int32 codes),
    "split16" (uint16 codes), "twopass" / "twopass16" (token pass, then the
    axis lanes), "chains<K>" / "chains<K>n16" (the chain token pass with K
    chains per meshlet, then the axis lanes)."""
    import subprocess
    src = tmp_path / "s.dpr4"
    src.write_bytes(data)
    fv, ft = tmp_path / "v.f32", tmp_path / "t.u32"
    mode = [] if not split else (["split"] if split is True else [split])
    r = subprocess.run([str(exe), str(src), str(fv), str(ft)] + mode, capture_output=True, text=True)
    return r.returncode, np.fromfile(fv, np.float32).reshape(-1, 3) if r.returncode in (0, 3) else None, \
        np.fromfile(ft, np.uint32).reshape(-1, 3) if r.returncode in (0, 3) else None


LANE_MODES = [False, "split", "split16", "twopass", "twopass16", "chains1", "chains2",
              "chains3n16", "chains16"]


@pytest.mark.parametrize("split", LANE_MODES, ids=["lane"] + LANE_MODES[1:])
def test_lane_decoder_matches_oracle(lane_check, coracle, tmp_path, split):
    """The gfx950 kernels' pre-pass and per-lane decoders (one lane per
    meshlet, the split token/axis lanes, and the chain token pass: chain
    starts, placeholder FIFOs and the per-meshlet record resolution),
    compiled for the host, reproduce the oracle bit-for-bit on every golden
    stream, the hand-built streams and freshly encoded small meshes."""
    from tests.test_oracle import hand_stream_fifo, hand_stream_small
    streams = [np.load(p)["data"].tobytes() for p in sorted((ROOT / "tests" / "golden").glob("*_*.npz"))
               if "data" in np.load(p) and not p.stem.startswith("pdlt_")]
    streams += [hand_stream_small()[0], hand_stream_fifo()[0]]
    for name, V, T in small_meshes():
        streams.append(encode_with_oracle(V, T, pred=([[33, 32, -33], [31, 34, -33], [32, 32, -32]],
                                                      [5, 5, 5]))[1])
    for data in streams:
        rc, Vl, Tl = run_lane_check(lane_check, data, tmp_path, split)
        assert rc == 0
        Vo, To = coracle.decode(data)
        assert np.array_equal(Tl.astype(np.int64), To)
        assert np.array_equal(Vl.view(np.uint32), Vo.view(np.uint32))


def test_lane_prepass_rejects_corruption(lane_check, coracle, tmp_path):
    """Bit flips: whenever the host-built pre-pass accepts a stream, the lane
    decoder's output equals the oracle's (or the oracle rejects it too)."""
    z = np.load(ROOT / "tests" / "golden" / "dup_hf_24.npz")
    base = bytearray(z["data"].tobytes())
    rng = np.random.default_rng(11)
    for trial in range(60):
        d = bytearray(base)
        pos = int(rng.integers(8 * (86 + 32) + 8, 8 * len(d)))
        d[pos >> 3] ^= 1 << (7 - (pos & 7))
        rc, Vl, Tl = run_lane_check(lane_check, bytes(d), tmp_path)
        if rc != 0:
            continue
        try:
            Vo, To = coracle.decode(bytes(d))
        except ValueError:
            raise AssertionError(f"trial {trial}: pre-pass accepted a stream the oracle rejects")
        assert np.array_equal(Tl.astype(np.int64), To)
        assert np.array_equal(Vl.view(np.uint32), Vo.view(np.uint32))


def test_fit_predictor_without_samples_is_canonical():
    """A mesh without parallelogram vertices keeps the canonical predictor
    (no GPU work is needed to decide that)."""
    from meshpress_amd.encoder import search
    z = np.zeros(0, np.int32)
    n, K = search.fit_predictor(z, z, z, z, np.zeros(2, np.int64))
    assert np.array_equal(n, search.CANON_N) and np.array_equal(K, search.CANON_K)


@pytest.mark.parametrize("name", ["torus_30x10", "uv_c1"])
def test_dprb_writer_matches_reference_blob(name):
    """The DPRB blob derived from the DPR4 bytes alone equals the blob built
    from the reference encoder's own per-meshlet meta (anchor offsets,
    n_kind0, substream offsets; scripts/dump_stride_dup_blob.py:88-152)."""
    from meshpress_amd.utils.dprb import to_dprb
    z = np.load(ROOT / "tests" / "golden" / f"dprb_{name}.npz")
    dpr4 = z["dpr4"].tobytes()
    assert to_dprb(dpr4) == z["dprb"].tobytes()
    # the stream is the reference encoder's output for the same mesh
    assert dpr4 == np.load(ROOT / "tests" / "golden" / f"dup_{name}.npz")["data"].tobytes()


def test_dprb_writer_rejects_bad_streams():
    from meshpress_amd.utils.dprb import to_dprb
    z = np.load(ROOT / "tests" / "golden" / "dprb_torus_30x10.npz")
    dpr4 = bytearray(z["dpr4"].tobytes())
    with pytest.raises(ValueError, match="magic"):
        to_dprb(b"\0" * 100)
    with pytest.raises(IndexError):
        to_dprb(bytes(dpr4[:60]))
    bad = bytearray(dpr4)
    bad[86 + 4 * 0 + 3] ^= 0x01          # first meshlet offset off by one bit
    with pytest.raises(ValueError):
        to_dprb(bytes(bad))
