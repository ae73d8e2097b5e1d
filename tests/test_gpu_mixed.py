"""The FP32/FP64 mixed form of the substep (dh_set_mixed_precision) against
the FP64 oracle (needs an MI355X).

The mixed form transforms only the O(dt) increments U_new - U_old of the
substep and the fluctuations about the plane means in FP32; the Fourier
state, the plane means ((kx,kz) = (0,0) coefficients) and every term with a
mean factor, CF, the RK combination, the solves and COURANT stay FP64.
After 10 RK3 steps from the same state as the oracle every state field of
each deck must agree to relative L2 <= 1e-5 (the gate of the mixed form),
and DELTA_T to 1e-5 relative: the small all-terms deck (scalar, rotation,
mixed BCs), the KH shear layer with free-slip walls, and the Re_tau=180
channel (128x129x128, variable dt)."""
import numpy as np
import pytest

from diablo_amd import channel
from tests import helpers

pytestmark = pytest.mark.gpu

STATE = ("cu1", "cu2", "cu3", "cp")
BAR = 1e-5


def _say(*a):
    print(*a, flush=True)


def _fmt(d):
    return {k: f"{v:.1e}" for k, v in d.items()}


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


@pytest.mark.parametrize("name", ["small_case", "kh_shear_layer", "channel_re180"])
def test_mixed_ten_steps(dor, name):
    steps = 10
    _say(f"[mixed] {steps} steps {name}")
    o, p, gy, ic = helpers.make_oracle(dor, name)
    o2, *_ = helpers.make_oracle(dor, name)
    helpers.perturb(dor, o2, 1e-16)
    ch = channel.Channel(p, gy)
    ch.upload(helpers.oracle_state(dor, o))
    ch.set_mixed_precision(True)
    assert ch.mixed_precision
    ch.set_clock(0.0, p["delta_t"], 0)
    o.time_step(steps)
    o2.time_step(steps)
    ch.time_step(steps)
    t_o, dt_o, ts_o = o.clock()
    t_g, dt_g, hb, ts_g = ch.clock()
    _say(f"  time oracle {t_o!r} mixed {t_g!r}; dt oracle {dt_o!r} mixed {dt_g!r}")
    assert ts_g == ts_o == steps
    assert abs(dt_g - dt_o) <= BAR * dt_o
    st = ch.download()
    keys = STATE + (("th",) if p["n_th"] > 0 else ())
    errs = helpers.field_errors(dor, o, st, keys=keys)
    floor = helpers.field_errors(dor, o, helpers.oracle_state(dor, o2), keys=keys)
    _say("  err  ", _fmt(errs))
    _say("  floor", _fmt(floor))
    bad = {k: e for k, e in errs.items() if not e <= BAR}
    assert not bad, bad
    ch.close()


def test_switching_back_is_the_plain_path(dor):
    """Mixed on, a step, off again: the next steps are the FP64 path's,
    bit for bit (the captured step graphs of the other form are dropped)."""
    _say("[mixed] switch back small_case")
    o, p, gy, ic = helpers.make_oracle(dor, "small_case")
    state = helpers.oracle_state(dor, o)
    ref = channel.Channel(p, gy)
    ref.upload(state)
    ref.set_clock(0.0, p["delta_t"], 0)
    ref.time_step(3)
    a = ref.download()
    ch = channel.Channel(p, gy)
    ch.upload(state)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.set_mixed_precision(True)
    ch.time_step(1)
    ch.set_mixed_precision(False)
    assert not ch.mixed_precision
    ch.upload(state)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.time_step(3)
    b = ch.download()
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k
    ref.close()
    ch.close()


SPINUP_STEPS = 20000


def test_mixed_long_run_statistics():
    """Long-run statistics of the mixed form against the FP64 path on the
    developed turbulent Re_tau=180 channel (128x129x128, CFL 0.5): a seeded
    transition IC (large-scale modes of finite amplitude over the laminar
    profile) is run SPINUP_STEPS steps in FP64 to developed turbulence
    (checked: u_tau from the wall shear within 10 % of 1, the forcing's
    value, and the rms peak of u beyond 2 u_tau); from that state each form
    runs 2000 steps, with the plane statistics (dh_save_stats) averaged every
    10 steps over the last 1000.  Tolerances of the mixed form: mean U+
    within 2 % and the rms velocities within 5 % (relative L2 over the
    profile)."""
    import math

    from diablo_amd import configs, turbulence
    print(f"[mixed] long-run statistics, channel_re180, {SPINUP_STEPS} steps of spin-up", flush=True)
    p, gy, ic = configs.channel_re180()
    ch = channel.Channel(p, gy)
    turbulence.transition_ic(ch, p, gy, seed=1)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.time_step(SPINUP_STEPS)
    st0 = ch.download()
    clk = ch.clock()
    ch.close()
    prof = {}
    for mixed in (False, True):
        ch = channel.Channel(p, gy)
        ch.upload(st0)
        ch.set_clock(clk[0], clk[1], clk[3])
        ch.set_mixed_precision(mixed)
        ch.time_step(1000)
        stats, _ = turbulence.run_statistics(ch, p, 1000, every=10)
        prof[mixed] = stats
        ch.close()
    ref = prof[False]
    utau = math.sqrt(ref["tau_w"])
    _say(f"  FP64 u_tau {utau:.3f}, U_b {ref['ubulk']:.2f}, max urms {ref['urms'].max():.2f}")
    assert abs(utau - 1.0) <= 0.1 and ref["urms"].max() >= 2.0  # developed turbulence
    rel = {k: float(np.linalg.norm(prof[True][k] - ref[k]) / np.linalg.norm(ref[k]))
           for k in ("ume", "urms", "vrms", "wrms")}
    _say("  rel L2 mixed vs FP64", _fmt(rel))
    assert rel["ume"] <= 0.02, rel
    assert max(rel["urms"], rel["vrms"], rel["wrms"]) <= 0.05, rel
