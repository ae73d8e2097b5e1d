"""CPU restatement (test infrastructure only) of the learned closure's
network and of its loss gradient, in float64.

The network has no counterpart in the reference (its LES has only the
constant-Smagorinsky model of les.f:196-227); this restates the device
kernels' definition so that their forward pass and gradients can be
checked: 6 inputs -> 64 -> 64 -> 64 -> 1, GELU (tanh form) hidden layers,
softplus output, loss sum_n (softplus(y_n) - t_n)^2 / N.

With ``bf16=True`` every value the device rounds to BF16 before an MFMA
(the weights, the inputs of each layer, dL/dy and the pre-activation
gradients that the weight and bias sums read) is rounded the same way here,
and everything else is float64; with ``bf16=False`` it is plain float64
mathematics.  The float64 gradient is pinned by central differences of the
float64 loss (tests/test_oracle_mlp.py).

Only tests may import this module.
"""
import numpy as np

IN, INPAD, H = 6, 32, 64
OW1 = 0
OB1 = OW1 + INPAD * H
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H * H
OW4 = OB3 + H
OB4 = OW4 + H
NP = OB4 + 1


def bf16(x):
    """Round to the nearest BF16 (ties to even), as float64 values."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    W1 = P[OW1:OB1].reshape(INPAD, H)[:IN]
    return (W1, P[OB1:OW2], P[OW2:OB2].reshape(H, H), P[OB2:OW3], P[OW3:OB3].reshape(H, H), P[OB3:OW4],
            P[OW4:OB4], P[OB4])


def gelu_t(z):
    a = np.sqrt(2.0 / np.pi) * (z + 0.044715 * z ** 3)
    return np.tanh(a)


def gelu(z):
    return 0.5 * z * (1.0 + gelu_t(z))


def gelu_grad(z):
    t = gelu_t(z)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * np.sqrt(2.0 / np.pi) * (1.0 + 3 * 0.044715 * z * z)


def softplus(y):
    return np.maximum(y, 0.0) + np.log1p(np.exp(-np.abs(y)))


def sigmoid(y):
    return 1.0 / (1.0 + np.exp(-y))


def forward(P, X, bf=False):
    """X [n][6] -> (y [n], cache)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(P)
    x = r(X)
    z1 = x @ r(W1) + b1
    z2 = r(gelu(z1)) @ r(W2) + b2
    z3 = r(gelu(z2)) @ r(W3) + b3
    y = r(gelu(z3)) @ r(w4) + b4
    return y, (x, z1, z2, z3)


def predict(P, X, scale=1.0, bf=False):
    return scale * softplus(forward(P, X, bf)[0])


def loss(P, X, t, bf=False):
    y, _ = forward(P, X, bf)
    return float(np.sum((softplus(y) - t) ** 2) / len(t))


def loss_grad(P, X, t, bf=False):
    """(loss, gradient [NP]) in the parameter layout of the device kernels."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(P)
    n = len(t)
    y, (x, z1, z2, z3) = forward(P, X, bf)
    res = softplus(y) - t
    L = float(np.sum(res * res) / n)
    dy = 2.0 * res * sigmoid(y) / n
    h1, h2, h3 = r(gelu(z1)), r(gelu(z2)), r(gelu(z3))
    d3 = dy[:, None] * r(w4)[None, :] * gelu_grad(z3)
    d2 = (r(d3) @ r(W3).T) * gelu_grad(z2)
    d1 = (r(d2) @ r(W2).T) * gelu_grad(z1)
    g = np.zeros(NP)
    g[OW1:OB1].reshape(INPAD, H)[:IN] = x.T @ r(d1)
    g[OB1:OW2] = r(d1).sum(0)
    g[OW2:OB2] = (h1.T @ r(d2)).ravel()
    g[OB2:OW3] = r(d2).sum(0)
    g[OW3:OB3] = (h2.T @ r(d3)).ravel()
    g[OB3:OW4] = r(d3).sum(0)
    g[OW4:OB4] = h3.T @ r(dy)
    g[OB4] = r(dy).sum()
    return L, g


def init_params(seed=0):
    """He-style initial weights (normal, std sqrt(2/fan_in)), zero biases, and
    an output bias that starts softplus near 1."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NP)
    P[OW1:OB1].reshape(INPAD, H)[:IN] = rng.standard_normal((IN, H)) * np.sqrt(2.0 / IN)
    P[OW2:OB2] = (rng.standard_normal((H, H)) * np.sqrt(2.0 / H)).ravel()
    P[OW3:OB3] = (rng.standard_normal((H, H)) * np.sqrt(2.0 / H)).ravel()
    P[OW4:OB4] = rng.standard_normal(H) * np.sqrt(1.0 / H)
    P[OB4] = 0.5
    return P
