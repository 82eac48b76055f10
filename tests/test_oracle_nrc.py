"""Pins for the FP64 neural-cache oracle (oracle/nrc.py): finite
differences for the backward pass, closed forms for the encoding, the
forward pass, bf16 rounding and Adam."""
import numpy as np
import pytest

from oracle import nrc


def _random_net(seed=3, n=6):
    rng = np.random.default_rng(seed)
    params = nrc.init_params(seed).astype(np.float64)
    p = rng.uniform(0, 512, (n, 3))
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    alb = rng.uniform(0, 1, (n, 3))
    X = nrc.encode(p, nn, alb, (0, 0, 0), 1.0 / 512)
    t = rng.uniform(0, 0.3, (n, 3))
    return params, X, t


def test_backward_matches_finite_differences():
    params, X, t = _random_net()

    def loss(pv):
        y, _ = nrc.forward(pv, X)
        return nrc.mse_grad(y, t)[0]

    # forward() takes float32 parameters; use a float64 copy of the same values
    p32 = params.astype(np.float32).astype(np.float64)
    y, cache = nrc.forward(p32, X)
    _, dy = nrc.mse_grad(y, t)
    g = nrc.backward(cache, dy)
    rng = np.random.default_rng(5)
    # indices in every layer (W0 used rows, W1..W4)
    offs = np.cumsum([0] + [a * b for a, b in nrc.SHAPES])
    idx = np.concatenate([rng.integers(offs[i], offs[i + 1], 12) for i in range(5)])
    idx = idx[(idx >= offs[1]) | ((idx - offs[0]) // 64 < nrc.N_FEAT)]
    h = 1e-6  # small enough that no ReLU changes side
    for i in idx:
        # finite differences in float64 on a float64 forward pass
        def f(delta):
            Ws = [w.astype(np.float64).copy() for w in nrc.split(p32.astype(np.float32))]
            flat = np.concatenate([w.ravel() for w in Ws])
            flat[i] += delta
            a = X
            o = 0
            for li, (r, c) in enumerate(nrc.SHAPES):
                a = a @ flat[o:o + r * c].reshape(r, c)
                o += r * c
                if li < 4:
                    a = np.maximum(a, 0.0)
            return nrc.mse_grad(a, t)[0]
        fd = (f(h) - f(-h)) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-4, abs=1e-9), i
    assert loss(p32) > 0


def test_zero_padded_inputs_get_zero_gradient():
    params, X, t = _random_net()
    y, cache = nrc.forward(params, X)
    g = nrc.backward(cache, nrc.mse_grad(y, t)[1])
    gw0 = g[: 96 * 64].reshape(96, 64)
    assert np.all(gw0[nrc.N_FEAT:] == 0.0)


def test_encoding_closed_form():
    p = np.array([[128.0, 0.0, 384.0]])
    X = nrc.encode(p, [[0, 1, 0]], [[0.25, 0.5, 0.75]], (0, 0, 0), 1.0 / 512)
    # p_hat = (0.25, 0, 0.75)
    assert X[0, 0] == pytest.approx(np.sin(np.pi / 4), abs=1e-12)   # sin(pi * 0.25)
    assert X[0, 1] == pytest.approx(np.cos(np.pi / 4), abs=1e-12)
    assert X[0, 2] == pytest.approx(1.0, abs=1e-12)                 # sin(2 pi * 0.25)
    assert X[0, 3] == pytest.approx(0.0, abs=1e-12)
    assert X[0, 4] == pytest.approx(0.0, abs=1e-12)                 # sin(4 pi * 0.25)
    assert X[0, 5] == pytest.approx(-1.0, abs=1e-12)
    assert X[0, 24] == 0.0 and X[0, 25] == 1.0                      # y axis: p_hat 0
    assert X[0, 48] == pytest.approx(np.sin(0.75 * np.pi), abs=1e-12)
    assert X[0, 48 + 23] == pytest.approx(np.cos(2 ** 11 * np.pi * 0.75), abs=1e-9)
    assert list(X[0, 72:79]) == [0, 1, 0, 0.25, 0.5, 0.75, 1]
    assert np.all(X[0, 79:] == 0)


def test_encoding_large_frequency_argument_reduction():
    rng = np.random.default_rng(1)
    ph = rng.uniform(-1, 2, 1000)
    X = nrc.encode(np.stack([ph, ph, ph], 1) * 512, np.zeros((1000, 3)), np.zeros((1000, 3)), (0, 0, 0), 1 / 512)
    phf = nrc.p_hat(np.stack([ph, ph, ph], 1) * 512, (0, 0, 0), 1 / 512)[:, 0].astype(np.float64)
    for k in range(12):
        assert np.allclose(X[:, 2 * k], np.sin(2.0 ** k * np.pi * phf), atol=1e-9)
        assert np.allclose(X[:, 2 * k + 1], np.cos(2.0 ** k * np.pi * phf), atol=1e-9)


def test_bf16_rounding():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, 3.0e-5], np.float32)
    r = nrc.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == 1 + 2 ** -7              # above the midpoint
    assert r[3] == 1 + 2 ** -6              # tie with odd lower neighbour -> even (up)
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-5) / 3.0e-5 < 2 ** -8


def test_forward_closed_form():
    params = np.zeros(nrc.N_PARAMS, np.float32)
    W = nrc.split(params)
    W[0][78, 0] = 2.0          # hidden0 = 2 * bias feature
    W[0][72, 1] = -1.0         # hidden1 = relu(-normal.x)
    for li in (1, 2, 3):
        W[li][0, 0] = 1.0
        W[li][1, 1] = 1.0
    W[4][0] = [1.0, 2.0, 3.0]
    W[4][1] = [10.0, 0.0, 0.0]
    X = nrc.encode(np.zeros((2, 3)), [[1, 0, 0], [-1, 0, 0]], np.zeros((2, 3)), (0, 0, 0), 1.0)
    y, _ = nrc.forward(params, X)
    assert np.allclose(y, [[2, 4, 6], [12, 4, 6]])
    yb, _ = nrc.forward(params, X, bf=True)
    assert np.allclose(yb, y)


def test_adam_closed_form():
    w0 = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-4, 0.0])
    opt = nrc.Adam(3, lr=1e-3, beta1=0.9, beta2=0.99, eps=1e-8)
    w1 = opt.step(w0, g)
    # first step: m_hat = g, v_hat = g^2
    assert np.allclose(w1, w0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    w2 = opt.step(w1, g)
    # constant gradient keeps m_hat = g and v_hat = g^2
    assert np.allclose(w2, w1 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


def test_training_reduces_loss():
    rng = np.random.default_rng(2)
    n = 256
    p = rng.uniform(0, 512, (n, 3))
    nn = np.tile([0.0, 1.0, 0.0], (n, 1))
    alb = rng.uniform(0, 1, (n, 3))
    X = nrc.encode(p, nn, alb, (0, 0, 0), 1 / 512)
    t = 0.1 + 0.05 * np.sin(p[:, :1] / 64.0) * np.ones((1, 3))
    w = nrc.init_params(1).astype(np.float64)
    opt = nrc.Adam(nrc.N_PARAMS, lr=1e-3)
    y, c = nrc.forward(w, X)
    l0, _ = nrc.mse_grad(y, t)
    for _ in range(200):
        y, c = nrc.forward(w, X)
        _, dy = nrc.mse_grad(y, t)
        w = opt.step(w, nrc.backward(c, dy))
    y, _ = nrc.forward(w, X)
    assert nrc.mse_grad(y, t)[0] < 0.05 * l0


def test_rel_mse():
    assert nrc.rel_mse([1.0, 2.0], [1.0, 2.0]) == 0.0
    assert nrc.rel_mse([0.0, 0.0], [3.0, 4.0]) == 1.0
    assert nrc.rel_mse([1.1, 2.0], [1.0, 2.0]) == pytest.approx(0.01 / 5)
