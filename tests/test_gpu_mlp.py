"""The proposal network's HIP kernels against the CPU oracle: the MFMA GEMM
(all operand layouts and epilogues), forward, softmax cross-entropy,
backward (gradients), Adam, the inference heads, and a short training run
that must lower the loss."""
import numpy as np
import pytest

from oracle import learned_oracle as LO
from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _t():
    import torch
    return torch


def bf16_dev(a):
    """host float array -> (device bf16 tensor, float64 host copy of its values)"""
    torch = _t()
    q = LO.bf16_round(np.asarray(a, np.float32))
    d = torch.from_numpy(q).to("cuda").to(torch.bfloat16)
    return d, q.astype(np.float64)


def host(t):
    torch = _t()
    if t.dtype == torch.bfloat16:
        return t.float().cpu().numpy().astype(np.float64)
    return t.cpu().numpy()


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 16, 0, 0), (100, 70, 40, 0, 0), (33, 176, 128, 0, 1),
                                         (128, 176, 300, 1, 0), (32, 128, 517, 1, 0),
                                         (5, 3, 7, 1, 1)])
def test_gemm_layouts(M, N, K, ta, tb):
    from meshpress_amd import _lib
    torch = _t()
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    dA, hA = bf16_dev(A)
    dB, hB = bf16_dev(B)
    opA = hA.T if ta else hA
    opB = hB.T if tb else hB
    want = opA @ opB
    bound = np.abs(opA) @ np.abs(opB)
    bias = rng.standard_normal(N).astype(np.float32)
    dbias = torch.from_numpy(bias).to("cuda")
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    L = _lib.lib()
    lda = A.shape[1]
    ldb = B.shape[1]
    _lib.check(L.stride_mlp_gemm(_lib.ptr(dA), lda, ta, _lib.ptr(dB), ldb, tb, M, N, K, 0,
                                 _lib.ptr(dbias), None, 0, _lib.ptr(C), None, N, _lib.stream_ptr()))
    got = host(C)
    assert np.all(np.abs(got - (want + bias)) <= 1e-5 * bound + 1e-5), np.abs(got - want - bias).max()
    # relu epilogue, bf16 out
    Cb = torch.empty((M, N), dtype=torch.bfloat16, device="cuda")
    _lib.check(L.stride_mlp_gemm(_lib.ptr(dA), lda, ta, _lib.ptr(dB), ldb, tb, M, N, K, 1,
                                 _lib.ptr(dbias), None, 0, None, _lib.ptr(Cb), N, _lib.stream_ptr()))
    ref = np.maximum(want + bias, 0)
    gotb = host(Cb)
    assert np.all(np.abs(gotb - ref) <= np.abs(ref) * 2 ** -8 + 1e-5 * bound + 1e-5)
    # mask epilogue
    mask = rng.standard_normal((M, N))
    dmask, hmask = bf16_dev(mask)
    _lib.check(L.stride_mlp_gemm(_lib.ptr(dA), lda, ta, _lib.ptr(dB), ldb, tb, M, N, K, 2,
                                 None, _lib.ptr(dmask), N, None, _lib.ptr(Cb), N, _lib.stream_ptr()))
    ref = np.where(hmask > 0, want, 0)
    gotb = host(Cb)
    assert np.all(np.abs(gotb - ref) <= np.abs(ref) * 2 ** -8 + 1e-5 * bound + 1e-5)


def _net(seed, n):
    from meshpress_amd.learned.mlp import ProposalMLP, init_params
    p = init_params(seed)
    rng = np.random.default_rng(seed + 7)
    # non-zero biases so the bias paths are exercised
    p[LO.N_W:] = rng.standard_normal(LO.N_P - LO.N_W).astype(np.float32) * 0.1
    p[LO.N_W + 2 * LO.N_HID + 171:] = 0
    net = ProposalMLP(params=p)
    x = rng.uniform(0, 1.5, (n, LO.N_IN)).astype(np.float32)
    labels = np.zeros((n, LO.N_HEADS), np.uint8)
    for h in range(LO.N_HEADS):
        labels[:, h] = rng.integers(0, LO.HEAD_OFF[h + 1] - LO.HEAD_OFF[h], n)
    labels[rng.random((n, LO.N_HEADS)) < 0.1] = LO.IGNORE
    return net, p, x, labels


def close_bf16(got, want, frac=2e-3):
    """bf16 tensors: equal except rare one-ulp rounding flips."""
    d = np.abs(got - want)
    tol = np.abs(want) * 2 ** -7 + 1e-6
    assert np.all(d <= tol), d.max()
    assert np.mean(d > 1e-6 + np.abs(want) * 1e-6) <= frac


def test_forward_matches_oracle():
    net, p, x, _ = _net(1, 300)
    dx = net.features_to_device(x)
    h1, h2, z = net.forward(dx)
    o1, o2, oz = LO.forward(x, p, emulate=True)
    close_bf16(host(h1), o1)
    close_bf16(host(h2), o2)
    # logits from the GPU's own h2 are exact up to float32 summation
    W3 = LO.bf16_round(LO.split(p)[2]).astype(np.float64)
    zz = host(h2) @ W3 + LO.split(p)[5]
    assert np.allclose(host(z), zz, rtol=1e-5, atol=1e-5)
    assert np.abs(host(z) - oz).max() <= 1e-2 * np.abs(oz).max()


def test_xent_backward_adam_match_oracle():
    torch = _t()
    net, p, x, labels = _net(2, 517)
    n = len(x)
    dx = net.features_to_device(x)
    h1, h2, z = net.forward(dx)
    dy = torch.from_numpy(labels).to("cuda")
    dz, loss = net.xent(z, dy, 1.0 / n)
    odz, oloss = LO.xent(host(z), labels, 1.0 / n, emulate=True)
    close_bf16(host(dz), odz)
    assert np.allclose(host(loss), oloss, rtol=1e-4, atol=1e-4)
    dh2, dh1 = net.backward(dx, h1, h2, dz)
    hx = host(dx)
    g, odh2, odh1 = LO.backward(hx, host(h1), host(h2), host(dz), p, emulate=True,
                                return_hidden=True)
    close_bf16(host(dh2), odh2)
    close_bf16(host(dh1), odh1)
    # weight and bias gradients from the GPU's own hidden gradients: float32 sums
    g2 = np.zeros(LO.N_P)
    gW1, gW2, gW3, gb1, gb2, gb3 = LO.split(g2)
    gW3[:] = host(h2).T @ host(dz)
    gb3[:] = host(dz).sum(0)
    gW2[:] = host(h1).T @ host(dh2)
    gb2[:] = host(dh2).sum(0)
    gW1[:] = hx.T @ host(dh1)
    gb1[:] = host(dh1).sum(0)
    gg = host(net.g)
    for a, b in zip(LO.split(gg), LO.split(g2)):
        assert np.abs(a - b).max() <= 1e-3 * max(1e-3, np.abs(b).max()), np.abs(a - b).max()
    # and against the oracle's end-to-end gradient (rounding flips allowed)
    for a, b in zip(LO.split(gg), LO.split(g)):
        assert np.abs(a - b).max() <= 2e-2 * max(1e-6, np.abs(b).max())
    # Adam: two steps against the float32 oracle
    p0, m0, v0 = net.params_host(), np.zeros(LO.N_P, np.float32), np.zeros(LO.N_P, np.float32)
    net.adam(1e-3)
    po, mo, vo = LO.adam(p0, m0, v0, gg.astype(np.float32), 1, 1e-3)
    assert np.allclose(net.params_host(), po, rtol=1e-6, atol=1e-7)
    net.adam(1e-3)
    po, mo, vo = LO.adam(po, mo, vo, gg.astype(np.float32), 2, 1e-3)
    assert np.allclose(net.params_host(), po, rtol=1e-6, atol=1e-7)
    assert np.array_equal(host(net.wb).astype(np.float32), LO.bf16_round(net.params_host()[:LO.N_W]))


def test_heads_match_oracle():
    net, p, x, _ = _net(3, 200)
    _, _, z = net.forward(net.features_to_device(x))
    pp, cls, cp = net.heads(z)
    opp, ocls, ocp = LO.heads(host(z))
    assert np.allclose(host(pp), opp, rtol=1e-5, atol=1e-6)
    assert np.array_equal(host(cls), ocls)
    assert np.allclose(host(cp), ocp, rtol=1e-5, atol=1e-6)


def test_training_lowers_loss():
    """Labels that are a fixed function of the features must be learned."""
    torch = _t()
    from meshpress_amd.learned.mlp import ProposalMLP
    rng = np.random.default_rng(5)
    n = 4096
    x = rng.uniform(0, 1, (n, LO.N_IN)).astype(np.float32)
    Wt = rng.standard_normal((LO.N_IN, LO.HEAD_OFF[-1]))
    zt = x @ Wt
    labels = np.zeros((n, LO.N_HEADS), np.uint8)
    for h in range(LO.N_HEADS):
        labels[:, h] = zt[:, LO.HEAD_OFF[h]:LO.HEAD_OFF[h + 1]].argmax(1)
    net = ProposalMLP(seed=0)
    dx = net.features_to_device(x)
    dy = torch.from_numpy(labels).to("cuda")
    first = last = None
    for it in range(300):
        sel = torch.from_numpy(rng.integers(0, n, 512)).to("cuda")
        loss = net.train_step(dx.index_select(0, sel), dy.index_select(0, sel), 3e-3)
        v = float(loss.mean())
        first = v if first is None else first
        last = v
    assert last < 0.5 * first, (first, last)
    _, _, z = net.forward(dx)
    _, cls, _ = net.heads(z)
    acc = (host(cls) == labels[:, 3:]).mean()
    assert acc > 0.3, acc
