"""The learned LES closure (dh_set_les_closure; no counterpart in the
reference, which has only the constant-Smagorinsky model it replaces) on
an MI355X.

* Random weights: NU_T after one LES substep equals a numpy evaluation of
  the same network on the device's own features (dh_field_dev 301), with
  the kernel's roundings restated (features, weights and the first hidden
  layer rounded to BF16 with ties to even, FP32 sums).  FP32 sums in
  another order move the result by ~1e-7 of max|NU_T|; a first-layer unit
  whose FP32 sum falls on the other side of a BF16 rounding boundary (about
  one unit in 2^16) moves its point by up to ~1e-3.  So: at most 0.1 % of
  the points beyond 1e-5 of max|NU_T|, none beyond 1e-2.  A wrong lane or
  point mapping puts almost every point out by O(1).
* The network that returns its first input (|S| at GY) is the Smagorinsky
  model with |S| rounded to BF16: its NU_T is the model's to 2^-8 relative,
  point by point (a scrambled lane or point mapping fails this at once),
  and two steps later the velocity is within 1e-3 of the model's run."""
import numpy as np
import pytest

from diablo_amd import channel
from tests import devmem, helpers

pytestmark = pytest.mark.gpu

F, H = 8, 32
C_SMAG = 0.13


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


def _bf16(x):
    """float32 -> bf16 (round to nearest even) as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32)


def _get(ch, which, n):
    out = np.empty(n)
    ch.synchronize()
    assert devmem.hip().hipMemcpy(out.ctypes.data, ch.field_ptr(which), n * 8, 2) == 0
    return out


def _widths(p, gy):
    _, _, DY, _ = channel.grid_metrics(gy)
    dx, dz = p["lx"] / p["nx"], p["lz"] / p["nz"]
    return np.array([-2.0 * C_SMAG ** 2 * (dx * d * 2.0 * dz) ** (2.0 / 3.0) if j >= 1 else 0.0
                     for j, d in enumerate(DY)])


def _pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]]).astype(np.float32)


def _ctx(dor, name):
    o, p, gy, ic = helpers.make_oracle(dor, name)
    ch = channel.Channel(p, gy)
    ch.upload(helpers.oracle_state(dor, o))
    ch.set_clock(0.0, p["delta_t"], 0)
    return ch, p, gy


def test_closure_matches_its_restatement(dor):
    _say("[closure] random weights, les_channel, one substep")
    rng = np.random.default_rng(11)
    W1 = rng.standard_normal((F, H)) * 0.5
    b1 = rng.standard_normal(H) * 0.1
    W2 = rng.standard_normal((H, H)) * 0.3
    b2 = rng.standard_normal(H) * 0.1
    w3 = rng.standard_normal(H) * 0.3
    b3 = 0.05
    ch, p, gy = _ctx(dor, "les_channel")
    ch.set_les_closure(_pack(W1, b1, W2, b2, w3, b3))
    ch.rk_substep(1)
    ny, nz, nx = p["ny"], p["nz"], p["nx"]
    n = (ny + 2) * nz * nx
    feat = _get(ch, 301, F * n).reshape(F, ny + 2, nz * nx)
    nut = _get(ch, 300, n).reshape(ny + 2, nz * nx)
    # the restatement
    x = _bf16(feat.transpose(1, 2, 0).astype(np.float32))  # [J][col][f]
    h1 = _bf16(np.maximum(x @ _bf16(W1.astype(np.float32)) + b1.astype(np.float32), 0.0))
    h2 = np.maximum(h1 @ _bf16(W2.astype(np.float32)) + b2.astype(np.float32), 0.0)
    y = (h2 @ w3.astype(np.float32) + np.float32(b3)).astype(np.float64)
    dyw = _widths(p, gy)
    want = np.zeros_like(nut)
    for J in range(3, ny):
        want[J] = -0.5 * dyw[J] * np.maximum(y[J], 0.0)
    scale = np.abs(want).max()
    err = np.abs(nut - want) / scale
    frac = float((err[3:ny] > 1e-5).mean())
    _say(f"  max|NU_T| {scale:.3e}; error / max|NU_T|: max {err.max():.2e}, median {np.median(err[3:ny]):.1e}, "
         f"share beyond 1e-5 {frac:.1e}; nonzero rows {int((np.abs(nut).max(axis=1) > 0).sum())} of {ny - 3}")
    assert scale > 0 and (np.abs(want[3:ny]).max(axis=1) > 0).sum() > (ny - 3) // 2
    assert np.abs(nut[:3]).max() == 0 and np.abs(nut[ny:]).max() == 0
    assert frac <= 1e-3
    assert err.max() <= 1e-2
    ch.close()


def test_closure_reproduces_smagorinsky(dor):
    _say("[closure] |S| network vs the Smagorinsky model, les_small")
    W1 = np.zeros((F, H)); W1[0, 0] = 1.0
    W2 = np.zeros((H, H)); W2[0, 0] = 1.0
    w3 = np.zeros(H); w3[0] = 1.0
    weights = _pack(W1, np.zeros(H), W2, np.zeros(H), w3, 0.0)
    a, p, gy = _ctx(dor, "les_small")
    b, _, _ = _ctx(dor, "les_small")
    b.set_les_closure(weights)
    for ch in (a, b):
        ch.rk_substep(1)
    n = (p["ny"] + 2) * p["nz"] * p["nx"]
    na, nb = _get(a, 300, n), _get(b, 300, n)
    assert np.abs(na).max() > 0
    rel = np.abs(nb - na) / np.maximum(np.abs(na), 1e-300)
    _say(f"  max pointwise relative difference {rel[na != 0].max():.2e} (bound 2^-8 = {2.0 ** -8:.2e})")
    assert np.array_equal(na == 0, nb == 0)
    assert rel[na != 0].max() <= 2.0 ** -8
    for ch in (a, b):
        ch.rk_substep(2)
        ch.rk_substep(3)
        ch.time_step(2)
    sa, sb = a.download(), b.download()
    d = {k: helpers.rel_l2(sb[k], sa[k]) for k in ("cu1", "cu2", "cu3")}
    _say("  velocity after the run, relative to the model's:", {k: f"{v:.1e}" for k, v in d.items()})
    assert max(d.values()) <= 1e-3
    b.set_les_closure(None)  # back to the model
    assert not np.isnan(b.download()["cu1"]).any()
    a.close()
    b.close()
