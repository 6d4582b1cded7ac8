"""The learned encoder proposal with the packaged weights, on the 50
held-out meshes (seeds the network never saw): bit-rate regret against the
exhaustive search within 1%, every stream decoding to exactly what the
exhaustive stream decodes to, and the gain of the trained network over
the method without it (the closed-form coder choice with the canonical
predictor, and the same network untrained)."""
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]



def held_out():
    """All 50 held-out meshes (seeds 2,000,006+; training used 1,000,003 +
    i for i < 500), 20K-500K triangles."""
    from meshpress_amd import workloads
    return workloads.learned_specs(50, 1)


def _preps():
    from meshpress_amd import workloads
    from meshpress_amd.encoder.prep import prepare
    for fam, n, seed in held_out():
        V, T = workloads.mesh_with_tris(fam, n, seed)
        yield prepare(V, T, precision_error=1.0 / 4094.5, precision_mode="bbox_frac")


@pytest.fixture(scope="module")
def encodes():
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.learned.mlp import ProposalMLP
    from meshpress_amd.learned.proposal import BaselineProposal, LearnedProposal
    kw = dict(precision_error=1.0 / 4094.5, precision_mode="bbox_frac")
    encs = {
        "exhaustive": STRIDEEncoder(**kw),
        "learned": STRIDEEncoder(selector="learned", **kw),
        "learned+fallback": STRIDEEncoder(selector="learned+fallback", **kw),
        "baseline": STRIDEEncoder(selector="learned", proposal=BaselineProposal(), **kw),
        "untrained": STRIDEEncoder(selector="learned",
                                   proposal=LearnedProposal(mlp=ProposalMLP(seed=0)), **kw),
    }
    res = {k: [] for k in encs}
    for i, prep in enumerate(_preps()):
        print(f"[learned] held-out mesh {i}: {prep['n_t']:,} tris", file=sys.stderr, flush=True)
        for k, e in encs.items():
            res[k].append((e.encode_prepared(prep), dict(e.last_stats)))
    return res


def total(res, k):
    return sum(len(d) for d, _ in res[k])


def test_learned_regret_within_one_percent(encodes):
    ex = total(encodes, "exhaustive")
    for k in ("learned", "learned+fallback"):
        regret = (total(encodes, k) - ex) / ex
        assert regret <= 0.01, (k, regret)
        for (d, _), (de, _) in zip(encodes[k], encodes["exhaustive"]):
            assert (len(d) - len(de)) / len(de) <= 0.01
    # the fallback scores the streams the network is least sure of exactly
    assert total(encodes, "learned+fallback") <= total(encodes, "learned")


def test_learned_streams_decode_like_exhaustive(encodes):
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    for i, (de, _) in enumerate(encodes["exhaustive"]):
        Ve, Te = ParaDeltaV5GpuDecoder(de).decode_to_host()
        for k in ("learned", "learned+fallback", "baseline", "untrained"):
            V, T = ParaDeltaV5GpuDecoder(encodes[k][i][0]).decode_to_host()
            assert np.array_equal(V.view(np.uint32), Ve.view(np.uint32)) and np.array_equal(T, Te), (k, i)


def test_trained_network_gains_over_method_without_it(encodes):
    learned = total(encodes, "learned")
    baseline = total(encodes, "baseline")
    untrained = total(encodes, "untrained")
    # fewer bytes than the network-free coder rule and than the untrained net
    assert learned < baseline, (learned, baseline)
    assert learned < untrained, (learned, untrained)
    # and far fewer IRLP candidate evaluations than the exhaustive fit
    ev_l = sum(s["irlp_evals"] for _, s in encodes["learned"])
    ev_e = sum(s["irlp_evals"] for _, s in encodes["exhaustive"])
    assert ev_l * 5 <= ev_e, (ev_l, ev_e)


def test_cpu_exhaustive_search_chooses_like_gpu(encodes):
    """The host-thread exhaustive search that bench.py times as the
    encode-side CPU baseline (the oracle's fit_predictor with the threaded
    C cost, then the threaded C coder pick) makes the GPU search's choices:
    the same predictor and K per axis, the same code per substream."""
    from itertools import islice

    from meshpress_amd.encoder import search
    from oracle import stride_oracle as O
    from tests.oracle_c import load
    co = load()
    for i, prep in enumerate(islice(_preps(), 3)):
        nat = prep["native"]
        A, B, C, V, off = (np.ascontiguousarray(x, dtype=np.int64) for x in nat.para_samples())
        pn, pk = O.fit_predictor(
            A, B, C, V, off,
            cost_fn_factory=lambda a, b, c, v, o: (
                lambda cands, K: co.irlp_cost_mt(a, b, c, v, o, cands, K, 4)),
            full_cost_fn=lambda a, b, c, v, o, n3, K: int(
                co.irlp_cost_mt(a, b, c, v, o, [n3], K, 4)[0]))
        st = encodes["exhaustive"][i][1]
        assert np.array_equal(np.asarray(pn).reshape(-1), np.asarray(st["pred_n"]).reshape(-1)), i
        assert np.array_equal(np.asarray(pk).reshape(-1), np.asarray(st["pred_K"]).reshape(-1)), i
        vals, soff, _ = nat.residuals(pn, pk)
        tagk_cpu, _ = co.pick_streams(vals, soff, 4)
        tagk_gpu, _ = search.select_codes(vals.astype(np.int32), soff)
        assert np.array_equal(tagk_cpu, tagk_gpu), i
