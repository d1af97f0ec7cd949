"""CPU checks that pin the oracle of the learned-correction network."""
import numpy as np
import pytest

from oracle import cnn as C


def test_bf16_known_answers():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, 0.0, 3e38, 1e-40],
                 dtype=np.float32)
    # halfway cases round to even; 1e-40 (subnormal) keeps its top bits
    want = np.array([1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -2.5, 0.0, 2.9979e38, 9.99e-41])
    got = C.bf16(x)
    np.testing.assert_allclose(got[:6], want[:6], rtol=0, atol=0)
    assert abs(got[6] / 3e38 - 1) < 2 ** -8
    assert got.view(np.uint32)[7] & 0xFFFF == 0


def test_conv_adjoint_and_weight_grad():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2, 5, 7, 3))
    D = rng.standard_normal((2, 5, 7, 4))
    W = rng.standard_normal((9, 4, 3))
    # <conv(X, W), D> = <X, conv_T(D, W)> = <W, weight_grad(D, X)>
    a = np.sum(C.conv(X, W) * D)
    b = np.sum(X * C.conv_T(D, W))
    c = np.sum(W * C.weight_grad(D, X))
    assert a == pytest.approx(b, rel=1e-12) and a == pytest.approx(c, rel=1e-12)
    # brute-force conv at one pixel of the border and one inside
    for (bb, y, x) in [(0, 0, 0), (1, 2, 3), (1, 4, 6)]:
        s = np.zeros(4)
        for t in range(9):
            dy, dx = t // 3 - 1, t % 3 - 1
            if 0 <= y + dy < 5 and 0 <= x + dx < 7:
                s += W[t] @ X[bb, y + dy, x + dx]
        np.testing.assert_allclose(C.conv(X, W)[bb, y, x], s, rtol=1e-13)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = C.init_params(3).astype(np.float64)
    params[C.NPARAM - 1] = 0.3
    X0 = rng.standard_normal((2, 6, 5, C.CIN))
    dC = rng.standard_normal((2, 6, 5))
    _, cache = C.forward(params, X0, exact=True)
    g = C.backward(cache, dC)
    idx = np.concatenate([rng.choice(C.NPARAM, 40, replace=False),
                          [0, 1152, 1184, 10432, 19680, C.NPARAM - 1]])
    h = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (np.sum(dC * C.forward(pp, X0, exact=True)[0]) -
              np.sum(dC * C.forward(pm, X0, exact=True)[0])) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8), i


def test_planted_mistake_is_caught():
    """A transposed W4 in the backward (a plausible slip) fails the same check."""
    rng = np.random.default_rng(4)
    params = C.init_params(5).astype(np.float64)
    X0 = rng.standard_normal((1, 4, 4, C.CIN))
    dC = rng.standard_normal((1, 4, 4))
    _, cache = C.forward(params, X0, exact=True)
    bad = dict(cache)
    bad["Wq"] = dict(cache["Wq"])
    bad["Wq"]["W4"] = cache["Wq"]["W4"][::-1]  # taps reversed
    g_bad = C.backward(bad, dC)
    g_good = C.backward(cache, dC)
    h = 1e-6
    worst = 0.0
    for i in range(1184, 1184 + 9216, 97):  # W2 entries
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (np.sum(dC * C.forward(pp, X0, exact=True)[0]) -
              np.sum(dC * C.forward(pm, X0, exact=True)[0])) / (2 * h)
        assert g_good[i] == pytest.approx(fd, rel=1e-5, abs=1e-8)
        worst = max(worst, abs(g_bad[i] - fd) / max(1e-3, abs(fd)))
    assert worst > 1e-2


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3])
    p1, m, v = C.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.1)
    # with zero moments the bias-corrected first step is lr * g / (|g| + eps)
    np.testing.assert_allclose(p1, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)
