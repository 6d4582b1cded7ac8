"""GPU encoder search parity: coder choice, coder costs and IRLP costs on the
GPU against the golden reference outputs and the oracle; STRIDEEncoder
bytes against the reference's encode_dup bytes."""
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import encode_with_oracle, gpu_available, small_meshes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

GOLD = Path(__file__).resolve().parent / "golden"


def test_select_codes_golden(oracle):
    from meshpress_amd.encoder import search
    z = np.load(GOLD / "coder_choice.npz")
    off = np.concatenate(([0], np.cumsum(z["stream_lens"]))).astype(np.int64)
    tagk, bits = search.select_codes(z["stream_vals"].astype(np.int32), off)
    assert np.array_equal(tagk[:, 0], z["tags"]) and np.array_equal(tagk[:, 1], z["ks"])
    for i in range(len(z["stream_lens"])):
        assert bits[i] == oracle.pick_tag_k(z["stream_vals"][off[i]:off[i + 1]])[2]
    costs = search.stream_costs(z["stream_vals"].astype(np.int32), off)
    for i in range(len(z["stream_lens"])):
        assert list(costs[i]) == oracle.stream_costs(z["stream_vals"][off[i]:off[i + 1]])


def test_irlp_cost_golden():
    from meshpress_amd.encoder import search
    z = np.load(GOLD / "coder_choice.npz")
    cost = search.IrlpCost(z["irlp_a"].astype(np.int32), z["irlp_b"].astype(np.int32),
                           z["irlp_c"].astype(np.int32), z["irlp_v"].astype(np.int32),
                           z["irlp_off"].astype(np.int64))
    for cand, K, want in zip(z["irlp_cands"], z["irlp_K"], z["irlp_costs"]):
        assert int(cost(cand[None], int(K))[0]) == int(want)
    # many candidates in one launch (more than one 64-lane group)
    cands = np.stack([z["irlp_cands"][i % 6] + np.array([i // 6, 0, 0]) for i in range(150)])
    got = cost(cands, 5)
    from tests.oracle_c import load
    want = load().irlp_cost(z["irlp_a"], z["irlp_b"], z["irlp_c"], z["irlp_v"], z["irlp_off"],
                            cands, 5)
    assert np.array_equal(got, want)


def test_irlp_cost_axes_equals_per_axis_costs():
    """The three-axis evaluator (one sample upload, one candidate upload and
    one read-back per batch) against the oracle's costs of every axis,
    including an empty candidate list."""
    from meshpress_amd.encoder import search
    from tests.oracle_c import load
    z = np.load(GOLD / "coder_choice.npz")
    rng = np.random.default_rng(3)
    n = len(z["irlp_a"])
    # three axes' samples: the golden axis and two permutations of it
    perms = [np.arange(n), rng.permutation(n), rng.permutation(n)]
    cols = np.stack([np.stack([z[f"irlp_{c}"].astype(np.int32)[p] for p in perms])
                     for c in "abcv"])
    off = z["irlp_off"].astype(np.int64)
    cost = search.IrlpCostAxes(cols, off)
    items = [(0, z["irlp_cands"][:3]), (2, z["irlp_cands"][2:6]), (1, np.zeros((0, 3), np.int32)),
             (1, z["irlp_cands"][1:2] + 1)]
    got = cost.run(items, 5)
    for (d, cands), g in zip(items, got):
        want = load().irlp_cost(cols[0, d], cols[1, d], cols[2, d], cols[3, d], off,
                                np.asarray(cands).reshape(-1, 3), 5)
        assert np.array_equal(g, want), d
    assert cost.evals == 8


@pytest.mark.parametrize("name", ["uv_12x24", "torus_30x10", "hf_24", "ico_2", "one_tri"])
def test_fit_predictor_golden(name):
    from meshpress_amd.encoder import search
    z = np.load(GOLD / f"dup_{name}.npz")
    n, K = search.fit_predictor(z["samples_A"], z["samples_B"], z["samples_C"], z["samples_V"],
                                z["samples_off"])
    assert np.array_equal(n, z["fit_n"]) and np.array_equal(K, z["fit_K"])


@pytest.mark.parametrize("name", ["uv_12x24", "torus_30x10", "hf_24", "ico_2", "one_tri"])
def test_encoder_canonical_bytes_golden(name):
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.utils.types import Model
    meshes = {n: (v, t) for n, v, t in small_meshes()}
    V, T = meshes[name]
    out = STRIDEEncoder(predictor="canonical").encode(Model.from_arrays(V, T))
    z = np.load(GOLD / f"dup_{name}.npz")
    assert out.data == z["data"].tobytes()
    assert out.bits_per_vertex == len(out.data) * 8 / len(V)
    assert out.bits_per_triangle == len(out.data) * 8 / len(T)


def test_encoder_c1_generalized(oracle, coracle):
    """Default STRIDEEncoder on the 65,536-triangle sphere: the GPU fit equals
    the reference fit, bytes equal the oracle writer with that predictor,
    and the stream decodes back to the quantised source."""
    from meshpress_amd import workloads as w
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.utils.types import Model
    V, T = w.uv_sphere_c1()
    enc = STRIDEEncoder(max_verts=256, precision_error=0.0005)
    out = enc.encode(Model.from_arrays(V, T))
    z = np.load(GOLD / "dup_uv_c1.npz")
    assert np.array_equal(np.array(enc.last_stats["pred_n"]), z["fit_n"])
    assert np.array_equal(np.array(enc.last_stats["pred_K"]), z["fit_K"])
    prep, data = encode_with_oracle(V, T, pred=(z["fit_n"], z["fit_K"]))
    assert out.data == data
    assert 20.0 < out.bits_per_vertex < 60.0
    _, _, codes = coracle.decode(out.data, with_codes=True)
    l2g = np.concatenate([pl["local_to_global"] for pl in prep["plans"]]).astype(np.int64)
    assert np.array_equal(codes, prep["global_codes"][l2g])


def test_compressed_model_formulas_and_size():
    """STRIDEEncoder.encode returns CompressedModel with the reference's
    bpv/bpt formulas (encoder/implementation/stride.py:45-53); its bytes
    equal encode_prepared on the same prep; the size lands in the reference
    test's 20 < bpv < 60 band (tests/test_stride_encoder.py:45-49) for the
    65,536-triangle sphere."""
    from meshpress_amd import workloads
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.utils.types import Model
    V, T = workloads.uv_sphere_c1()
    enc = STRIDEEncoder(max_verts=256, precision_error=0.0005)
    cm = enc.encode(Model(V, T))
    assert cm.bits_per_vertex == len(cm.data) * 8 / len(V)
    assert cm.bits_per_triangle == len(cm.data) * 8 / len(T)
    assert 20 < cm.bits_per_vertex < 60
    again = enc.encode_prepared(enc.prepare(Model(V, T)))
    assert again == cm.data
