"""Golden fixtures of the channel timestep (tests/golden/, written by
tests/golden/make_golden.py from the oracle; the reference ships no outputs
and cannot be built here).  The CPU test pins the oracle against them (any
change to its arithmetic shows); the GPU test checks the HIP path from the
device-built initial condition against them, with no oracle run."""
import numpy as np
import pytest

from tests import helpers
from tests.golden import make_golden as mg


def _load(name):
    return dict(np.load(f"{mg.HERE}/{name}.npz"))


@pytest.mark.parametrize("name", list(mg.STEPS))
def test_oracle_reproduces_golden(name):
    g = _load(name)
    o, p, ic_state, st, clk = mg.run(name)
    assert clk[2] == int(g["time_step"]) and clk[0] == g["time"] and clk[1] == g["dt"]
    if name == "small_case":
        for k in mg.FIELDS:
            np.testing.assert_allclose(helpers.dealiased(st[k], o.nkx, o.tnkz), g[k], rtol=0,
                                       atol=1e-13 * np.abs(g[k]).max())
            np.testing.assert_allclose(helpers.dealiased(ic_state[k], o.nkx, o.tnkz), g[f"ic_{k}"], rtol=0,
                                       atol=1e-13 * max(np.abs(g[f"ic_{k}"]).max(), 1e-300))
    else:
        fp = mg.fingerprint(st, o.nkx, o.tnkz, o.n_th)
        for k, v in fp.items():
            np.testing.assert_allclose(v, g[k], rtol=0, atol=1e-12 * np.abs(g[k]).max())


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small_case", "channel_re180"])
def test_gpu_matches_golden(name):
    """Seeded IC built on the device, mg.STEPS[name] RK3 steps, compared
    with the fixture at the 10-step parity bar (1e-10 relative L2)."""
    from diablo_amd import channel, configs
    print(f"[golden] {name}", flush=True)
    g = _load(name)
    p, gy, ic = configs.CONFIGS[name]()
    ch = channel.Channel(p, gy)
    ch.initial_condition(ic)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.time_step(mg.STEPS[name])
    t, dt, hb, ts = ch.clock()
    assert ts == int(g["time_step"])
    assert abs(t - g["time"]) <= 1e-12 * max(1.0, abs(g["time"])) and abs(dt - g["dt"]) <= 1e-12 * g["dt"]
    st = ch.download()
    ch.close()
    nkx, tnkz = p["nx"] // 3, 2 * (p["nz"] // 3)
    errs = {}
    if name == "small_case":
        for k in mg.FIELDS:
            errs[k] = _rel(helpers.dealiased(st[k], nkx, tnkz), g[k])
        errs["cth"] = _rel(helpers.dealiased(st["cth"], nkx, tnkz), g["cth"])
    else:
        fp = mg.fingerprint(st, nkx, tnkz, p["n_th"])
        errs = {k: _rel(v, g[k]) for k, v in fp.items()}
    print("  ", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-10, errs
