"""Pins the CPU restatement of the learned closure's network (oracle/mlp_ref.py,
CPU only): its float64 gradient against central differences of its float64
loss, on every parameter block, and its BF16 rounding against exact cases."""
import numpy as np

from oracle import mlp_ref as M


def _batch(n=48, seed=3):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.IN))
    t = np.abs(rng.standard_normal(n)) * 2.0
    return X, t


def test_gradient_matches_central_differences():
    P = M.init_params(1)
    X, t = _batch()
    L, g = M.loss_grad(P, X, t, bf=False)
    assert abs(L - M.loss(P, X, t)) < 1e-15 * max(1.0, L)
    rng = np.random.default_rng(5)
    blocks = [(M.OW1, M.OW1 + M.IN * M.H), (M.OB1, M.OW2), (M.OW2, M.OB2), (M.OB2, M.OW3), (M.OW3, M.OB3),
              (M.OB3, M.OW4), (M.OW4, M.OB4), (M.OB4, M.NP)]
    worst = 0.0
    for lo, hi in blocks:
        for p in rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False):
            h = 1e-6 * max(1.0, abs(P[p]))
            Pp, Pm = P.copy(), P.copy()
            Pp[p] += h
            Pm[p] -= h
            fd = (M.loss(Pp, X, t) - M.loss(Pm, X, t)) / (2 * h)
            worst = max(worst, abs(fd - g[p]) / max(1e-8, abs(g).max()))
    assert worst < 1e-7, worst
    # padded inputs 6..31 of layer 1 have no gradient
    assert np.all(g[M.OW1 + M.IN * M.H: M.OB1] == 0.0)


def test_activation_derivatives():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(z)).max() < 1e-8
    assert np.abs(M.softplus(z) - np.log1p(np.exp(z))).max() < 1e-12
    fd = (M.softplus(z + h) - M.softplus(z - h)) / (2 * h)
    assert np.abs(fd - M.sigmoid(z)).max() < 1e-8
    # GELU's tanh form: 0 at 0, identity for large z, ~0 for very negative z
    assert M.gelu(0.0) == 0.0 and abs(M.gelu(8.0) - 8.0) < 1e-12 and abs(M.gelu(-8.0)) < 1e-12


def test_bf16_rounding():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 keeps 8 significant bits
    assert M.bf16(one) == one
    assert M.bf16(one + ulp / 2) == one            # tie -> even
    assert M.bf16(one + 3 * ulp / 2) == one + 2 * ulp  # tie -> even (up)
    assert M.bf16(one + ulp * 0.51) == one + ulp
    assert M.bf16(-3.140625) == -3.140625
    x = np.random.default_rng(0).standard_normal(1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
