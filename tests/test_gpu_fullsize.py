"""Full-size decode parity on the benchmark workloads.

* the 1900 x 1900 heightfield (7,212,402 triangles) at 12 and 16 bits per
  axis, with the decoder's default launch shape and with the autotuned one;
* the batch of 2,048 meshes with triangle counts log-uniform in
  [1K, 200K], decoded by one hipGraph replay and by one launch per mesh.

Every decode is compared with the C oracle (the restatement of decode_dup,
encoder/paradelta_v5_dup.py:471-618) bit for bit: indices, and float32
positions at 0 ulp.  The oracle's integer codes are also checked against the
quantised source codes of every meshlet's vertices (global_codes[l2g]), the
reference's own PASS rule (scripts/verify_stride_dup.py:176-179)."""
import functools
import os
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.slow,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def say(*a):
    print("[fullsize]", *a, file=sys.stderr, flush=True)


def host_threads() -> int:
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return max(1, min(16, os.cpu_count() or 1))


def bbox_error(bits: int) -> float:
    """bbox_frac precision giving exactly `bits` bits on the longest axis."""
    return 1.0 / ((1 << bits) - 1.5)


def assert_same(V, T, Vo, To, what=""):
    assert V.dtype == np.float32 and T.dtype == np.uint32, what
    assert V.shape == Vo.shape and T.shape == To.shape, what
    assert np.array_equal(T.astype(np.int64), To), what
    assert np.array_equal(V.view(np.uint32), Vo.view(np.uint32)), what


@functools.lru_cache(maxsize=None)
def _encoded(bits):
    from meshpress_amd import workloads
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.utils.types import Model
    V, T = workloads.heightfield_1900()
    enc = STRIDEEncoder(max_verts=256, max_tris=256, precision_error=bbox_error(bits),
                        precision_mode="bbox_frac")
    prep = enc.prepare(Model.from_arrays(V, T))
    data = enc.encode_prepared(prep)
    l2g = prep["native"].local_to_global()
    say(f"heightfield {bits}-bit: {len(data):,} bytes, {prep['n_meshlets']:,} meshlets, "
        f"g_bits {list(prep['g_bits'])}")
    return data, l2g, prep["global_codes"], [int(x) for x in prep["g_bits"]], len(T)


@functools.lru_cache(maxsize=None)
def _oracle_out(bits):
    from tests import oracle_c
    data = _encoded(bits)[0]
    return oracle_c.load().decode(data, n_threads=host_threads(), with_codes=True)


@pytest.mark.parametrize("bits", [12, 16])
def test_heightfield_1900_codes_roundtrip(bits):
    """The oracle's decoded codes are the quantised source codes."""
    data, l2g, gcodes, g_bits, n_t = _encoded(bits)
    Vo, To, codes = _oracle_out(bits)
    assert max(g_bits) == bits
    assert len(To) == n_t == 7212402
    assert np.array_equal(codes, gcodes[l2g])


@pytest.mark.parametrize("shape", ["default", "autotuned", "token_lanes", "chains2", "chains16"])
@pytest.mark.parametrize("bits", [12, 16])
def test_heightfield_1900_gpu_equals_oracle(bits, shape):
    from meshpress_amd.autotune import BENCH_CANDIDATES, tune
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    data = _encoded(bits)[0]
    Vo, To, _ = _oracle_out(bits)
    dec = ParaDeltaV5GpuDecoder(data)
    if shape == "autotuned":
        best, _ = tune(dec, steps=5, cands=BENCH_CANDIDATES)
        say(f"heightfield {bits}-bit: tuned {best}")
    elif shape == "token_lanes":
        dec.set_tune(mode=4, tok_chains=0, tok_group=32)
    elif shape.startswith("chains"):
        k = int(shape[6:])
        dec.set_tune(mode=4, tok_chains=k, tok_group=64 // k)
    say(f"heightfield {bits}-bit {shape}: decoding with {dec.get_tune()}")
    dec.d_verts.fill_(float("nan"))
    dec.d_tris.fill_(-1)
    V, T = dec.decode_to_host()
    assert_same(V, T, Vo, To, f"{bits}-bit {shape}")
    # a second decode into the same buffers gives the same bytes
    V2, T2 = dec.decode_to_host()
    assert_same(V2, T2, Vo, To, f"{bits}-bit {shape}, second decode")


@pytest.fixture(scope="module")
def batch2048():
    import time
    from meshpress_amd import workloads
    from meshpress_amd.encoder.many import encode_many
    specs = workloads.batch_specs(2048, 1000, 200000, 0)
    sources = [functools.partial(workloads.mesh_with_tris, f, n, s) for f, n, s in specs]
    t0 = time.perf_counter()

    def progress(i, n):
        if i % 256 == 0 or i == n:
            say(f"batch: encoded {i}/{n} meshes ({time.perf_counter() - t0:.0f} s)")

    blobs, _ = encode_many(sources, progress=progress, precision_error=bbox_error(12),
                           precision_mode="bbox_frac", predictor="generalized")
    return blobs


def test_batch_2048_graph_and_per_mesh_equal_oracle(batch2048):
    import torch
    from meshpress_amd.autotune import BENCH_CANDIDATES, tune
    from meshpress_amd.batch import BatchDecoder
    from tests import oracle_c
    co = oracle_c.load()
    blobs = batch2048
    want = []
    for blob in blobs:
        Vo, To = co.decode(blob)
        want.append((Vo, To.astype(np.uint32)))
    bd = BatchDecoder(blobs)
    assert bd.n_meshes == 2048
    say(f"batch: {bd.n_t_total:,} tris, {bd.n_meshlets:,} meshlets")
    default = bd.get_tune()
    best, _ = tune(bd, steps=3, cands=[c for c in BENCH_CANDIDATES if c["mode"] == 4],
                   mode_kw={"mode": "graph"})
    lanes = {**default, "tok_chains": 0, "tok_group": 32}
    for name, shape in (("default", default), ("autotuned", best), ("token_lanes", lanes)):
        bd.set_tune(**shape)
        for mode in ("graph", "graph", "per_mesh"):
            say(f"batch {name} {mode}: {bd.get_tune()}")
            bd.d_verts.fill_(float("nan"))
            bd.d_tris.fill_(-1)
            v, t = bd.decode(mode)
            torch.cuda.synchronize()
            vh, th = v.cpu().numpy(), t.cpu().numpy().view(np.uint32)
            for i, ((Vi, Ti), (Vo, To)) in enumerate(zip(bd.split(vh, th), want)):
                assert_same(Vi, Ti, Vo, To.astype(np.int64), f"mesh {i} ({name}, {mode})")
