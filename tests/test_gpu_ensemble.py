"""Graph-captured ensembles of uneven shear-layer instances
(dh_ensemble_*, needs an MI355X).

Six KH instances (NY = NX + 1, NZ = 4) are stepped together: two of NX = 64
and two of NX = 128 with different Ri, Re and seeds, one of NX = 96, and one
more of NX = 64 in the mixed form.  Members of one shape and form run as
one batch per launch (the kernels' batched variants), so the test covers
batches of two, batches of one and a group of its own for the mixed form.
Each member must come out bit-identical to the same instance stepped alone
with dh_time_step: the batch runs the same kernels with the same arguments
in the same order.  The first NX = 64 member is also compared with the
oracle at its roundoff floor (this size runs the LDS row transforms)."""
import numpy as np
import pytest

from diablo_amd import channel, configs
from tests import helpers

pytestmark = [pytest.mark.gpu, pytest.mark.skip(reason=helpers.BATCHED_ENSEMBLE_OFF)]


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


def _oracle_errors(dor, p, gy, ic, ch, steps):
    """The member's state against the oracle after `steps` steps, with the
    oracle's own floor."""
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    o.create_th(ic["th_ic_type"])
    o.create_flow(ic["ic_type"], ic["seed"], ic["amp"], ic["ymask"], ic["ubulk0"])
    o2 = dor.Oracle(p)
    o2.set_grid_gy(gy)
    o2.create_th(ic["th_ic_type"])
    o2.create_flow(ic["ic_type"], ic["seed"], ic["amp"], ic["ymask"], ic["ubulk0"])
    helpers.perturb(dor, o2, 1e-16)
    for x in (o, o2):
        x.set_clock(0.0, p["delta_t"], 0)
        x.time_step(steps)
    return helpers.field_errors(dor, o, ch.download()), helpers.field_errors(dor, o, helpers.oracle_state(dor, o2))


def _decks():
    out = []
    for i, (nx, ri, re, mixed) in enumerate([(64, 0.05, 800.0, False), (96, 0.2, 3000.0, False),
                                             (128, 0.1, 1500.0, False), (64, 0.15, 2000.0, False),
                                             (128, 0.02, 600.0, False), (64, 0.08, 1200.0, True)]):
        p, gy, ic = configs.kh_shear_layer(nx=nx, ny=nx + 1, nz=4, ri=ri, re=re)
        out.append((p, gy, dict(ic, seed=200 + i), mixed))
    return out


def _make(p, gy, ic, mixed):
    ch = channel.Channel(p, gy)
    ch.initial_condition(ic)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.set_mixed_precision(mixed)
    return ch


def test_ensemble_equals_members_alone(dor):
    print("[ensemble] 6 KH instances, NX = 64 x 2, 96, 128 x 2, 64 mixed", flush=True)
    decks = _decks()
    alone = [_make(*d) for d in decks]
    together = [_make(*d) for d in decks]
    for ch in alone:
        ch.time_step(3)
    ens = channel.Ensemble(together)
    ens.time_step(2)
    ens.time_step(1)  # a second call replays the same graph
    for a, b in zip(alone, together):
        sa, sb = a.download(), b.download()
        for k in ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3", "cth"):
            assert np.array_equal(sa[k], sb[k]), (a.nx, k)
        assert a.clock() == b.clock()
    # member 0 (NX = 64) against the oracle: ICs and 3 steps
    p, gy, ic, _ = decks[0]
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    o.create_th(ic["th_ic_type"])
    o.create_flow(ic["ic_type"], ic["seed"], ic["amp"], ic["ymask"], ic["ubulk0"])
    o2 = dor.Oracle(p)
    o2.set_grid_gy(gy)
    o2.create_th(ic["th_ic_type"])
    o2.create_flow(ic["ic_type"], ic["seed"], ic["amp"], ic["ymask"], ic["ubulk0"])
    helpers.perturb(dor, o2, 1e-16)
    for x in (o, o2):
        x.set_clock(0.0, p["delta_t"], 0)
        x.time_step(3)
    errs = helpers.field_errors(dor, o, together[0].download())
    floor = helpers.field_errors(dor, o, helpers.oracle_state(dor, o2))
    print("  member NX=64 vs oracle", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    bad = helpers.within(errs, floor, 1e-10)
    assert not bad, bad
    ens.close()
    for ch in alone + together:
        ch.close()


def test_ensemble_large_members(dor):
    """The two largest sizes of the 64-instance ensemble, NX = 192 and 256
    (two each, so both run as batches), plus a mixed-form NX = 256 member:
    each bit-identical to its solo run, and the FP64 NX = 192 and 256 members
    at the oracle's floor."""
    print("[ensemble] NX = 192 x 2, 256 x 2, 256 mixed", flush=True)
    decks = []
    for i, (nx, ri, re, mixed) in enumerate([(192, 0.05, 900.0, False), (256, 0.12, 2500.0, False),
                                             (192, 0.2, 1800.0, False), (256, 0.03, 700.0, False),
                                             (256, 0.1, 1500.0, True)]):
        p, gy, ic = configs.kh_shear_layer(nx=nx, ny=nx + 1, nz=4, ri=ri, re=re)
        decks.append((p, gy, dict(ic, seed=300 + i), mixed))
    alone = [_make(*d) for d in decks]
    together = [_make(*d) for d in decks]
    for ch in alone:
        ch.time_step(3)
    ens = channel.Ensemble(together)
    ens.time_step(3)
    for a, b in zip(alone, together):
        sa, sb = a.download(), b.download()
        for k in ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3", "cth"):
            assert np.array_equal(sa[k], sb[k]), (a.nx, k)
    for m in (0, 1):
        p, gy, ic, _ = decks[m]
        errs, floor = _oracle_errors(dor, p, gy, ic, together[m], 3)
        print(f"  member NX={p['nx']} vs oracle", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        bad = helpers.within(errs, floor, 1e-10)
        assert not bad, bad
    # the mixed member: within the mixed form's bar of the FP64 oracle
    p, gy, ic, _ = decks[4]
    errs, _ = _oracle_errors(dor, p, gy, ic, together[4], 3)
    print("  mixed NX=256 vs oracle", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-5, errs
    ens.close()
    for ch in alone + together:
        ch.close()
