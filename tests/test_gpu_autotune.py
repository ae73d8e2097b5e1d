"""Launch-shape tuning (dh_set_plane_threads, dh_autotune,
dh_ensemble_autotune; needs an MI355X).

The block size of the plane transforms and of the nonlinear pass's LDS
variant only decides which thread computes which butterfly or mode, so
results must be bit-identical for every valid size: on the LDS path
(small_case, 16x17x8), on the register path (the 128x129x128 channel) and
on the NX = 64 register rows with the LES model (64x97x64, 256 and 512
threads; such planes offer no larger size).
dh_autotune must leave the state and clock exactly as it found them, and
an autotuned ensemble must still equal its members stepped alone."""
import numpy as np
import pytest

from diablo_amd import channel, configs
from tests import helpers

pytestmark = pytest.mark.gpu

KEYS = ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3")


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


def _pair(dor, name):
    o, p, gy, ic = helpers.make_oracle(dor, name)
    st = helpers.oracle_state(dor, o)
    chs = []
    for _ in range(2):
        ch = channel.Channel(p, gy)
        ch.upload(st)
        ch.set_clock(0.0, p["delta_t"], 0)
        chs.append(ch)
    return chs


@pytest.mark.parametrize("name,other", [("small_case", 256), ("channel_re180", 512), ("les_channel", 512)])
def test_block_size_does_not_change_results(dor, name, other):
    print(f"[autotune] block size {name}", flush=True)
    a, b = _pair(dor, name)
    assert a.plane_threads != other
    b.set_plane_threads(other)
    assert b.plane_threads == other
    a.time_step(2)
    b.time_step(2)
    sa, sb = a.download(), b.download()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    with pytest.raises(channel.DiabloError):
        b.set_plane_threads(96 + 1)
    if name == "les_channel":
        with pytest.raises(channel.DiabloError):
            b.set_plane_threads(1024)
    a.close()
    b.close()


def test_autotune_keeps_state(dor):
    print("[autotune] small_case", flush=True)
    a, b = _pair(dor, "small_case")
    a.time_step(1)
    b.time_step(1)
    before = b.download()
    clock = b.clock()
    best = b.autotune(2)
    assert best == b.plane_threads and best % 64 == 0
    after = b.download()
    for k in KEYS:
        assert np.array_equal(before[k], after[k]), k
    assert b.clock() == clock
    a.time_step(2)
    b.time_step(2)
    sa, sb = a.download(), b.download()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


@pytest.mark.skip(reason=helpers.BATCHED_ENSEMBLE_OFF)
def test_autotuned_ensemble_equals_members_alone():
    print("[autotune] ensemble of 4 KH instances", flush=True)
    decks = []
    for i, (nx, ri) in enumerate([(64, 0.05), (96, 0.2), (64, 0.15), (96, 0.1)]):
        p, gy, ic = configs.kh_shear_layer(nx=nx, ny=nx + 1, nz=4, ri=ri)
        decks.append((p, gy, dict(ic, seed=300 + i)))

    def make(p, gy, ic):
        ch = channel.Channel(p, gy)
        ch.initial_condition(ic)
        ch.set_clock(0.0, p["delta_t"], 0)
        return ch

    alone = [make(*d) for d in decks]
    together = [make(*d) for d in decks]
    ens = channel.Ensemble(together)
    ens.time_step(1)
    sizes = ens.autotune(1)
    print("  block sizes", sizes, flush=True)
    assert sizes[0] == sizes[2] and sizes[1] == sizes[3]  # one size per shape
    ens.time_step(2)
    for ch in alone:
        ch.time_step(3)
    for a, b in zip(alone, together):
        sa, sb = a.download(), b.download()
        for k in KEYS:
            assert np.array_equal(sa[k], sb[k]), (a.nx, k)
    ens.close()
    for ch in alone + together:
        ch.close()
