"""CPU checks of the learned-proposal oracle and of the native features:
bf16 rounding cases, gradients against finite differences, Adam's closed
form, the features on a hand-built meshlet, and the native features
against the oracle's restatement."""
import numpy as np
import pytest

from oracle import learned_oracle as LO


def f32(bits):
    return np.array([bits], dtype=np.uint32).view(np.float32)[0]


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    cases = [
        (one, one),
        (np.float32(1 + 2 ** -8), one),                     # tie -> even (mantissa 0)
        (np.float32(1 + 3 * 2 ** -8), np.float32(1 + 2 ** -6)),  # tie -> even (mantissa 2)
        (np.float32(1 + 2 ** -8 + 2 ** -20), np.float32(1 + 2 ** -7)),  # above the tie
        (np.float32(-2.5), np.float32(-2.5)),
        (f32(0x3F80FFFF), np.float32(1 + 2 ** -7)),
        (f32(0x7F7FFFFF), np.float32(np.inf)),              # rounds past the largest bf16
    ]
    for x, want in cases:
        got = LO.bf16_round(np.array([x], np.float32))[0]
        assert got == want, (x, got, want)
    assert np.isnan(LO.bf16_round(np.array([np.nan], np.float32))[0])


def _tiny_problem(seed=0, n=6):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(LO.N_P) * 0.2).astype(np.float64)
    x = rng.uniform(0, 1, (n, LO.N_IN))
    labels = np.zeros((n, LO.N_HEADS), np.uint8)
    for h in range(LO.N_HEADS):
        c = LO.HEAD_OFF[h + 1] - LO.HEAD_OFF[h]
        labels[:, h] = rng.integers(0, c, n)
    labels[0, 2] = LO.IGNORE
    labels[3, 5] = LO.IGNORE
    return p, x, labels


def test_gradient_matches_finite_differences():
    p, x, labels = _tiny_problem()
    h1, h2, z = LO.forward(x, p, emulate=False)
    dz, _ = LO.xent(z, labels, 1.0 / len(x), emulate=False)
    g = LO.backward(x, h1, h2, dz, p, emulate=False)
    rng = np.random.default_rng(1)
    # weights of every layer, biases of every layer
    W1, W2, W3 = 0, LO.N_IN * LO.N_HID, LO.N_IN * LO.N_HID + LO.N_HID * LO.N_HID
    idx = np.concatenate([rng.integers(W1, W2, 12), rng.integers(W2, W3, 12),
                          rng.integers(W3, LO.N_W, 12), rng.integers(LO.N_W, LO.N_P, 12)])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (LO.loss64(pp, x, labels) - LO.loss64(pm, x, labels)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_xent_gradient_and_ignored_heads():
    p, x, labels = _tiny_problem(3)
    _, _, z = LO.forward(x, p, emulate=False)
    dz, loss = LO.xent(z, labels, 1.0, emulate=False)
    for h in range(LO.N_HEADS):
        o, e = LO.HEAD_OFF[h], LO.HEAD_OFF[h + 1]
        assert np.allclose(dz[:, o:e].sum(axis=1), 0.0, atol=1e-12)
    assert np.all(dz[0, LO.HEAD_OFF[2]:LO.HEAD_OFF[3]] == 0)
    assert np.all(dz[:, LO.HEAD_OFF[-1]:] == 0)
    eps = 1e-6
    for (b, j) in [(1, 5), (2, 60), (4, 170), (5, 17)]:
        zp, zm = z.copy(), z.copy()
        zp[b, j] += eps
        zm[b, j] -= eps
        fd = (LO.xent(zp, labels, 1.0, False)[1].sum() - LO.xent(zm, labels, 1.0, False)[1].sum()) / (2 * eps)
        assert abs(fd - dz[b, j]) < 1e-6


def test_adam_closed_form():
    g = np.array([0.5, -2.0, 1e-3, 0.0], np.float32)
    p = np.array([1.0, 1.0, -1.0, 0.25], np.float32)
    m = np.zeros(4, np.float32)
    v = np.zeros(4, np.float32)
    lr = 0.01
    p1, m1, v1 = LO.adam(p, m, v, g, 1, lr)
    # step 1: m_hat = g, v_hat = g^2 -> update = lr * g / (|g| + eps)
    want = p - lr * g / (np.abs(g) + np.float32(1e-8))
    assert np.allclose(p1, want, rtol=1e-6, atol=1e-9)
    assert p1[3] == p[3]
    # step 2 by hand for element 0 with gradient 0.25
    g2 = np.array([0.25, 0, 0, 0], np.float32)
    p2, _, _ = LO.adam(p1, m1, v1, g2, 2, lr)
    mm = 0.9 * 0.9 * 0 + 0.9 * 0.1 * 0.5 + 0.1 * 0.25
    vv = 0.999 * 0.001 * 0.25 + 0.001 * 0.0625
    mh, vh = mm / (1 - 0.81), vv / (1 - 0.999 ** 2)
    assert abs(p2[0] - (p1[0] - lr * mh / (np.sqrt(vh) + 1e-8))) < 1e-6


def test_features_hand_built_meshlet():
    # 4 local vertices: three strip roots (anchor + 2 deltas) and one
    # parallelogram vertex v3 with refs (a, b, c) = (0, 1, 2)
    codes = np.array([[10, 20, 30], [13, 18, 30], [11, 25, 31], [15, 24, 28]], np.int64)
    walk = [(0, 0, 0, 0, 0), (1, 0, 0, 0, 0), (2, 0, 0, 0, 0), (3, 1, 0, 1, 2)]
    f = LO.features([walk], [[0, 1, 2, 3]], [[2]], codes)[0]
    assert f[0] == np.float32(4 / 256) and f[1] == np.float32(2 / 256)
    assert f[2] == np.float32(0.5) and f[3] == np.float32(0.75)
    assert f[4] == np.float32(np.log2(1 + 5) / 16)
    assert f[8] == 1.0 and f[7] == 0 and f[9] == 0 and f[10] == 0
    assert f[11] == np.float32(np.log2(3) / 8) and f[14] == np.float32(1 / 8)
    # dx = [3, -2] -> zigzag [6, 3] -> floor(log2(u+1)) = [2, 2]
    assert f[17] == np.float32(2 / 16)
    assert f[23] == np.float32(np.log2(1 + 4.5) / 16)
    # px = 15 - (10 + 13 - 11) = 3 -> zigzag 6 -> floor(log2 7) = 2
    assert f[20] == np.float32(2 / 16)
    assert f[26] == np.float32(np.log2(7) / 16)
    # shrink x: x = a + b - 2c = 1, y = v - c = 4 -> s = 4 -> clamp 2 -> 1.0
    assert f[29] == 1.0
    # y: x = 20 + 18 - 50 = -12, y = -1 -> s = 12/144
    assert f[30] == np.float32((12 / 144) / 2)
    # with predictor (128, 128, 0) >> 8 on x: (10*128 + 13*128 + 128) >> 8 = 12 -> px = 3
    pn = np.array([[128, 128, 0], [1, 1, -1], [1, 1, -1]])
    g = LO.features([walk], [[0, 1, 2, 3]], [[2]], codes, pred=(pn, [8, 0, 0]))[0]
    assert g[20] == np.float32(2 / 16) and np.array_equal(g[21:23], f[21:23])


@pytest.mark.parametrize("k", [0, 1, 2])
def test_native_features_match_oracle(k):
    from meshpress_amd import workloads as w
    from meshpress_amd.encoder.prep import prepare
    V, T = [w.displaced_uv_sphere(20, 40, 0.05, 4, 3), w.heightfield(40, 0.1, 6, 7),
            w.displaced_torus(45, 15, 1.0, 0.3, 0.05, 4, 5)][k]
    p = prepare(V, T, precision_error=1 / 4094.5, precision_mode="bbox_frac")
    nat = p["native"]
    F = nat.features(0)
    plans = nat.export_plans()
    walks = [[tuple(int(x) for x in r) for r in pl["walk"]] for pl in plans]
    l2g = [pl["local_to_global"] for pl in plans]
    sl = [[len(s) for s in pl["strips"]] for pl in plans]
    Fo = LO.features(walks, l2g, sl, p["global_codes"])
    assert F.shape == Fo.shape == (nat.n_meshlets, 32)
    assert np.allclose(F, Fo, rtol=1e-6, atol=1e-7), np.abs(F - Fo).max()
    assert np.all((F >= 0) & (F <= 2))
    # second pass: statistics of the residuals under a non-canonical predictor
    pn = np.array([[220, 220, -184], [1, 1, -1], [200, 210, -154]])
    pk = np.array([8, 0, 8])
    nat.residuals(pn, pk)
    F1 = nat.features(1)
    Fo1 = LO.features(walks, l2g, sl, p["global_codes"], pred=(pn, pk))
    assert np.allclose(F1, Fo1, rtol=1e-6, atol=1e-7), np.abs(F1 - Fo1).max()
    assert np.array_equal(F1[:, :17], F[:, :17]) and np.array_equal(F1[:, 29:], F[:, 29:])
