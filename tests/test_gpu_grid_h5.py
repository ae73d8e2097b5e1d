"""CREATE_GRID_CHAN from the reference's own grid file (dh_set_grid_h5 on
tests/golden/kh_test_grid.h5, the reference's KH_test/grid.h5): the KH
deck at the file's resolution (NY = 65) run one substep of each RK stage on
the device against the oracle given the same GY (needs an MI355X).  The
device forms GYF, DY and DYF from the file itself, so agreement to the
roundoff floor checks that path end to end."""
import os

import pytest

from diablo_amd import channel, configs
from tests import helpers

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "kh_test_grid.h5")
KEYS = ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3")


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


def _oracle(dor, p, gy, ic, eps=0.0):
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    o.create_th(ic["th_ic_type"])
    o.create_flow(ic["ic_type"], ic["seed"], ic["amp"], ic["ymask"], ic["ubulk0"])
    if eps:
        helpers.perturb(dor, o, eps)
    o.set_clock(0.0, p["delta_t"], 1)
    return o


def test_substeps_on_reference_grid_file(dor):
    print("[grid_h5] KH deck, NY=65, grid from the reference's grid.h5", flush=True)
    p, _, ic = configs.kh_shear_layer(ny=65)
    gy = channel.read_h5_f64(FIX, "/grids/y")
    o, o2 = _oracle(dor, p, gy, ic), _oracle(dor, p, gy, ic, 1e-16)
    ch = channel.Channel(p, configs.make_gy(65, 20.0, 0.0, 2))  # replaced by the file's grid below
    ch.set_grid_h5(FIX)
    ch.upload(helpers.oracle_state(dor, o))
    ch.set_clock(0.0, p["delta_t"], 1)
    for rk in (1, 2, 3):
        o.rk_chan_1(rk)
        o2.rk_chan_1(rk)
        ch.rk_substep(rk)
        st = ch.download()
        errs = helpers.field_errors(dor, o, st, keys=KEYS)
        floor = helpers.field_errors(dor, o, helpers.oracle_state(dor, o2), keys=KEYS)
        print(f"  rk={rk} err", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        bad = helpers.within(errs, floor, 1e-11)
        assert not bad, f"rk={rk}: {bad}"
    ch.close()
