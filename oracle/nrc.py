"""FP64 oracle of the neural irradiance cache — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this module, and only as the checker; the product (libtgx.so)
never uses it.

The cache has no counterpart in tg: it is specified by the benchmark
configuration (position with a 12-frequency sin/cos encoding, normal and
albedo in; four hidden layers of width 64 with ReLU; RGB irradiance out;
BF16 weights and activations with FP32 accumulation; FP32 Adam). The
choices this restatement fixes, shared with the HIP kernels:

* Input features (96, zero padded): for axis a in x, y, z and frequency
  k in 0..11, sin and cos of 2^k pi p_hat[a] at feature 24a + 2k (+1 for
  cos), with p_hat = (p - origin) * scale evaluated in fp32; then the normal
  (72..74), the albedo (75..77) and a constant 1 (78) that gives the first
  layer its bias. sin(2^k pi x) is evaluated as sin(2 pi frac(2^(k-1) x)),
  which is exact argument reduction.
* Layers W0 [96 x 64], W1..W3 [64 x 64], W4 [64 x 3], no biases, ReLU after
  W0..W3, linear output. Parameters are stored row-major W[in][out], in
  layer order (18624 floats).
* Loss: mean squared error over samples and channels,
  L = sum((y - t)^2) / (3 N), so dL/dy = 2 (y - t) / (3 N).
* Adam with bias correction: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
  w -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).

With ``bf16=True`` the forward and backward passes round to bfloat16 (round
to nearest even) exactly where the GPU kernels do: the inputs, the weights,
each hidden activation after ReLU, and every back-propagated gradient
(dL/dy and each masked dL/dz) before it enters a matrix product. Products
accumulate in float64 here and in fp32 on the GPU.
"""
from __future__ import annotations

import numpy as np

N_IN = 96
N_FEAT = 79
HIDDEN = 64
N_OUT = 3
N_FREQ = 12
DIMS = (N_IN, HIDDEN, HIDDEN, HIDDEN, HIDDEN, N_OUT)
SHAPES = tuple((DIMS[i], DIMS[i + 1]) for i in range(5))
N_PARAMS = sum(a * b for a, b in SHAPES)  # 18624
FLOPS_FORWARD = 2 * (78 * 64 + 3 * 64 * 64 + 64 * 3)  # per sample, unpadded


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params: np.ndarray) -> list[np.ndarray]:
    """Flat parameter vector -> [W0..W4] views (row-major W[in][out])."""
    out, o = [], 0
    for a, b in SHAPES:
        out.append(params[o:o + a * b].reshape(a, b))
        o += a * b
    return out


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights; rows of W0 past the used features are zero."""
    rng = np.random.default_rng(seed)
    ws = []
    for i, (a, b) in enumerate(SHAPES):
        fan_in = N_FEAT if i == 0 else a
        lim = np.sqrt(6.0 / fan_in) * (0.5 if i == 4 else 1.0)
        w = rng.uniform(-lim, lim, (a, b))
        if i == 0:
            w[N_FEAT:] = 0.0
        ws.append(w.ravel())
    return np.concatenate(ws).astype(np.float32)


def p_hat(p: np.ndarray, origin, scale: float) -> np.ndarray:
    """(p - origin) * scale in fp32, as the kernels evaluate it."""
    p = np.asarray(p, np.float32)
    o = np.asarray(origin, np.float32)
    return ((p - o).astype(np.float32) * np.float32(scale)).astype(np.float32)


def encode(p: np.ndarray, n: np.ndarray, albedo: np.ndarray, origin, scale: float) -> np.ndarray:
    """[N, 96] float64 input features."""
    ph = p_hat(p, origin, scale).astype(np.float64)
    N = ph.shape[0]
    X = np.zeros((N, N_IN))
    for a in range(3):
        for k in range(N_FREQ):
            x = ph[:, a] * 2.0 ** (k - 1)
            r = x - np.floor(x)
            X[:, 24 * a + 2 * k] = np.sin(2.0 * np.pi * r)
            X[:, 24 * a + 2 * k + 1] = np.cos(2.0 * np.pi * r)
    X[:, 72:75] = np.asarray(n, np.float32)
    X[:, 75:78] = np.asarray(albedo, np.float32)
    X[:, 78] = 1.0
    return X


def forward(params: np.ndarray, X: np.ndarray, bf: bool = False):
    """Returns (y [N, 3], cache) where cache holds the layer inputs."""
    Ws = [w.astype(np.float64) for w in split(np.asarray(params, np.float32))]
    if bf:
        Ws = [bf16(w) for w in Ws]
        X = bf16(X)
    acts = [X]
    a = X
    for i, w in enumerate(Ws):
        z = a @ w
        if i < 4:
            z = np.maximum(z, 0.0)
            if bf:
                z = bf16(z)
        a = z
        acts.append(a)
    return a, (Ws, acts, bf)


def mse_grad(y: np.ndarray, t: np.ndarray):
    """(loss, dL/dy) of the mean squared error over samples and channels."""
    n = y.shape[0] * N_OUT
    d = y - t
    return float(np.sum(d * d) / n), 2.0 * d / n


def backward(cache, dy: np.ndarray) -> np.ndarray:
    """Flat dL/dparams for the forward pass in cache."""
    Ws, acts, bf = cache
    g = bf16(dy) if bf else dy
    grads = [None] * 5
    for i in range(4, -1, -1):
        grads[i] = acts[i].T @ g
        if i > 0:
            da = g @ Ws[i].T
            da = da * (acts[i] > 0.0)
            g = bf16(da) if bf else da
    return np.concatenate([gw.ravel() for gw in grads])


class Adam:
    """FP64 Adam with bias correction."""

    def __init__(self, n: int, lr=1e-3, beta1=0.9, beta2=0.99, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps

    def step(self, w: np.ndarray, g: np.ndarray) -> np.ndarray:
        self.t += 1
        self.m = self.b1 * self.m + (1.0 - self.b1) * g
        self.v = self.b2 * self.v + (1.0 - self.b2) * g * g
        mh = self.m / (1.0 - self.b1 ** self.t)
        vh = self.v / (1.0 - self.b2 ** self.t)
        return w - self.lr * mh / (np.sqrt(vh) + self.eps)


def rel_mse(x: np.ndarray, ref: np.ndarray) -> float:
    """sum((x - ref)^2) / sum(ref^2) over all pixels and channels."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sum((x - ref) ** 2) / np.sum(ref ** 2))
