"""Parity of the device constant-Smagorinsky LES (les_chan and the NU_T
terms of the implicit solves) with the oracle's restatement of les.f and
channel.f (needs an MI355X).

The bar is that of the plain FP64 path: relative L2 <= max(1e-10,
4 x floor) per field after the steps run, where the floor is the oracle's
own spread under a 1e-16 perturbation of its input velocity; NU_T (a
nonlinear function of the strain, with |.| and sqrt) is held to the same
bar against the oracle's NU_T.  Cases: the LES channel of the reference's
LES configuration at 64 x 97 x 64 (Dirichlet walls, variable dt), a small
deck with every term, Neumann walls (APPLY_BC_LES) and a scalar (KAPPA_T),
and the spin-up gate of channel.f:303 crossed between steps 100 and 101."""
import numpy as np
import pytest

from diablo_amd import channel
from tests import devmem, helpers

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _fmt(d):
    return {k: f"{v:.1e}" for k, v in d.items()}


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


def _pair(dor, name, **kw):
    o, p, gy, ic = helpers.make_oracle(dor, name, **kw)
    o2, *_ = helpers.make_oracle(dor, name, **kw)
    helpers.perturb(dor, o2, 1e-16)
    ch = channel.Channel(p, gy)
    assert ch.les
    ch.upload(helpers.oracle_state(dor, o))
    return o, o2, ch, p


def _nut_gpu(ch):
    n = (ch.ny + 2) * ch.nz * ch.nx
    out = np.empty(n)
    ch.synchronize()
    assert devmem.hip().hipMemcpy(out.ctypes.data, ch.field_ptr(300), n * 8, 2) == 0
    return out.reshape(ch.ny + 2, ch.nz, ch.nx)


def _nut_oracle(dor, o):
    # a copy: the view dies with the oracle
    return o.rarr(dor.NUT)[:, : o.nz, : o.nx].copy()


def _run(dor, name, steps, step0=0, **kw):
    o, o2, ch, p = _pair(dor, name, **kw)
    for x in (o, o2):
        x.set_clock(0.0, p["delta_t"], step0)
    ch.set_clock(0.0, p["delta_t"], step0)
    o.time_step(steps)
    o2.time_step(steps)
    ch.time_step(steps)
    st = ch.download()
    keys = ("cu1", "cu2", "cu3", "cp")
    errs = helpers.field_errors(dor, o, st, keys=keys)
    floor = helpers.field_errors(dor, o, helpers.oracle_state(dor, o2), keys=keys)
    no, ng, no2 = _nut_oracle(dor, o), _nut_gpu(ch), _nut_oracle(dor, o2)
    errs["nut"] = helpers.rel_l2(ng, no)
    floor["nut"] = helpers.rel_l2(no2, no)
    _say("  err  ", _fmt(errs))
    _say("  floor", _fmt(floor))
    _say(f"  |NU_T| max oracle {np.abs(no).max():.3e} gpu {np.abs(ng).max():.3e}")
    t_o, dt_o, ts_o = o.clock()
    t_g, dt_g, hb, ts_g = ch.clock()
    assert ts_g == ts_o == step0 + steps
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    ch.close()
    return errs, floor, no


@pytest.mark.parametrize("name,steps", [("les_small", 4), ("les_channel", 3)])
def test_les_steps(dor, name, steps):
    _say(f"[les] {steps} steps {name}")
    errs, floor, nut = _run(dor, name, steps)
    assert np.abs(nut).max() > 0  # the model ran
    bad = helpers.within(errs, floor, 1e-10)
    assert not bad, bad


def test_les_spin_up_gate(dor):
    """A new flow: the model first runs in step 101; NU_T is zero before."""
    _say("[les] gate steps 99..102 les_small (new flow)")
    errs, floor, nut = _run(dor, "les_small", 3, step0=99, create_new_flow=1)
    assert np.abs(nut).max() > 0
    bad = helpers.within(errs, floor, 1e-10)
    assert not bad, bad
    errs, floor, nut = _run(dor, "les_small", 1, step0=0, create_new_flow=1)
    assert np.abs(nut).max() == 0.0
    bad = helpers.within(errs, floor, 1e-10)
    assert not bad, bad


def test_les_refuses_mixed_form(dor):
    p, gy, ic = helpers.configs.CONFIGS["les_small"]()
    ch = channel.Channel(p, gy)
    with pytest.raises(channel.DiabloError):
        ch.set_mixed_precision(True)
    ch.set_les(False)
    ch.set_mixed_precision(True)
    with pytest.raises(channel.DiabloError):
        ch.set_les(True)
    ch.close()


def _les_pair(dor, name, step0, **kw):
    o, p, gy, ic = helpers.make_oracle(dor, name, **kw)
    st = helpers.oracle_state(dor, o)
    chs = []
    for _ in range(2):
        ch = channel.Channel(p, gy)
        ch.upload(st)
        ch.set_clock(0.0, p["delta_t"], step0)
        chs.append(ch)
    return chs


def test_les_autotune_leaves_state(dor):
    """dh_autotune's trial steps cross the spin-up gate (steps 99..102 of a
    new flow) and must leave the state, the clock and NU_T as they were: the
    tuned context then steps bit-identically to an untuned one at the same
    block size."""
    _say("[les] autotune across the gate")
    a, b = _les_pair(dor, "les_small", 98, create_new_flow=1)
    t = a.autotune(4)
    b.set_plane_threads(t)
    assert np.abs(_nut_gpu(a)).max() == 0.0  # the trial steps left NU_T as it was
    for x in (a, b):
        x.time_step(4)
    sa, sb = a.download(), b.download()
    for k in ("cu1", "cu2", "cu3", "cp", "cth"):
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(_nut_gpu(a), _nut_gpu(b))
    assert a.clock() == b.clock()
    for x in (a, b):
        x.close()


def test_les_member_in_ensemble(dor):
    """An ensemble with an LES member steps its members one by one (the LES
    kernels have no batched forms); each equals its solo run bit for bit."""
    _say("[les] ensemble with an LES member")
    les = _les_pair(dor, "les_small", 0)
    plain = []
    o, p, gy, ic = helpers.make_oracle(dor, "small_case")
    st = helpers.oracle_state(dor, o)
    for _ in range(2):
        ch = channel.Channel(p, gy)
        ch.upload(st)
        ch.set_clock(0.0, p["delta_t"], 0)
        plain.append(ch)
    solo = [les[0], plain[0]]
    together = [les[1], plain[1]]
    for ch in solo:
        ch.time_step(3)
    ens = channel.Ensemble(together)
    ens.time_step(3)
    ens.synchronize()
    ens.close()  # before its members
    for a, b in zip(solo, together):
        sa, sb = a.download(), b.download()
        for k in ("cu1", "cu2", "cu3", "cp"):
            assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(_nut_gpu(solo[0]), _nut_gpu(together[0]))
    for ch in solo + together:
        ch.close()
