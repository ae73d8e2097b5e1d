"""The ensemble's instance shapes stepped one at a time through the plain
context path (dh_time_step; needs an MI355X): shear-layer instances of
NX = 192 and 256 (NY = NX + 1, NZ = 4), the two largest sizes of the
64-instance ensemble, in FP64 at the oracle's floor (the test_gpu_parity
bar: max(1e-10, 4 x floor) per field), and an NX = 256 instance in the
mixed form within 1e-5 of the FP64 oracle after 3 steps (the mixed form's
bar in test_gpu_mixed).  Both sizes run the LDS row transforms (the
register rows serve NX = 64 and 128 only)."""
import pytest

from diablo_amd import channel, configs
from tests import helpers

pytestmark = pytest.mark.gpu

SHAPES = {
    "nx192": (192, 0.05, 900.0, False),
    "nx256": (256, 0.12, 2500.0, False),
    "nx256_mixed": (256, 0.1, 1500.0, True),
}


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
    o.set_clock(0.0, p["delta_t"], 0)
    return o


@pytest.mark.parametrize("name", list(SHAPES))
def test_kh_instance_shapes(dor, name):
    nx, ri, re, mixed = SHAPES[name]
    p, gy, ic = configs.kh_shear_layer(nx=nx, ny=nx + 1, nz=4, ri=ri, re=re)
    ic = dict(ic, seed=300 + nx)
    print(f"[kh shapes] {name}: NX={nx} Ri={ri} Re={re} mixed={mixed}, 3 steps", flush=True)
    ch = channel.Channel(p, gy)
    ch.initial_condition(ic)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.set_mixed_precision(mixed)
    ch.time_step(3)
    o, o2 = _oracle(dor, p, gy, ic), _oracle(dor, p, gy, ic, 1e-16)
    o.time_step(3)
    o2.time_step(3)
    errs = helpers.field_errors(dor, o, ch.download())
    floor = helpers.field_errors(dor, o, helpers.oracle_state(dor, o2))
    ch.close()
    print("  err  ", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    print("  floor", {k: f"{v:.1e}" for k, v in floor.items()}, flush=True)
    if mixed:
        assert max(errs.values()) <= 1e-5, errs
    else:
        bad = helpers.within(errs, floor, 1e-10)
        assert not bad, bad
