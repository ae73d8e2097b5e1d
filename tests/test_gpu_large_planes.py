"""Planes too large for LDS (needs an MI355X): x-z planes of 256 x 256 and
512 x 512 points, whose transforms and nonlinear pass run on scratch planes
in device memory, against the oracle (FP64).

The transforms (FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL, mpi.f:1477-1606)
must agree to 1e-13 relative L2 on random planes; whole steps must agree to
max(1e-10, 4 x the oracle's own floor) per field, as in test_gpu_parity: the
channel at 256 x 17 x 256 (variable dt) and 512 x 9 x 512, and the small
deck's every term (scalar, buoyancy, rotation, oscillatory forcing, mixed
BCs) at 256 x 9 x 256."""
import numpy as np
import pytest

from diablo_amd import channel, configs
from tests import devmem, helpers

pytestmark = pytest.mark.gpu

DECKS = {
    "channel_256": ("channel_re180", dict(nx=256, ny=17, nz=256), 5),
    "small_256": ("small_case", dict(nx=256, ny=9, nz=256), 5),
    "channel_512": ("channel_re180", dict(nx=512, ny=9, nz=512), 2),
}


def _say(*a):
    print(*a, flush=True)


def _fmt(d):
    return {k: f"{v:.1e}" for k, v in d.items()}


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


@pytest.mark.parametrize("name", ["channel_256", "channel_512"])
def test_large_plane_transforms(dor, name):
    base, kw, _ = DECKS[name]
    _say(f"[large planes] transforms {name}")
    p, gy, ic = configs.CONFIGS[base](**kw)
    ch = channel.Channel(p, gy)
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    nx, nz, ny = p["nx"], p["nz"], p["ny"]
    nkx, tnkz = nx // 3, 2 * (nz // 3)
    npl = ny + 2
    rng = np.random.default_rng(5)
    phys = rng.standard_normal((npl, nz, nx))
    v = o.rarr(dor.CS1)
    v[:] = 0.0
    v[:, :nz, :nx] = phys
    o.fft_to_fourier(dor.CS1, 0, ny + 1)
    ref = o.carr(dor.CS1)[:, : tnkz + 1, : nkx + 1].copy()
    dp = devmem.DeviceArray(phys.nbytes)
    df = devmem.DeviceArray(ref.nbytes)
    dp.put(phys)
    ch.fft_to_fourier_dev(dp.ptr, df.ptr, npl)
    ch.synchronize()
    e_f = np.linalg.norm(df.get(np.complex128, ref.shape) - ref) / np.linalg.norm(ref)
    spec = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    c = o.carr(dor.CS1)
    c[:] = 0.0
    c[:, : tnkz + 1, : nkx + 1] = spec
    o.fft_to_physical(dor.CS1, 0, ny + 1)
    ref_p = o.rarr(dor.CS1)[:, :nz, :nx].copy()
    df.put(spec)
    ch.fft_to_physical_dev(df.ptr, dp.ptr, npl)
    ch.synchronize()
    e_i = np.linalg.norm(dp.get(np.float64, ref_p.shape) - ref_p) / np.linalg.norm(ref_p)
    _say(f"  forward rel L2 {e_f:.2e}  inverse rel L2 {e_i:.2e}")
    dp.free()
    df.free()
    ch.close()
    assert e_f <= 1e-13 and e_i <= 1e-13


@pytest.mark.parametrize("name", list(DECKS))
def test_large_plane_steps(dor, name):
    base, kw, steps = DECKS[name]
    _say(f"[large planes] {steps} steps {name}")
    o, p, gy, ic = helpers.make_oracle(dor, base, **kw)
    o2, *_ = helpers.make_oracle(dor, base, **kw)
    helpers.perturb(dor, o2, 1e-16)
    ch = channel.Channel(p, gy)
    ch.upload(helpers.oracle_state(dor, o))
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.time_step(steps)
    o.time_step(steps)
    o2.time_step(steps)
    t_o, dt_o, ts_o = o.clock()
    t_g, dt_g, _, ts_g = ch.clock()
    assert ts_g == ts_o == steps
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o and abs(t_g - t_o) <= 1e-12 * max(1.0, abs(t_o))
    st = ch.download()
    keys = ("cu1", "cu2", "cu3", "cp")
    errs = helpers.field_errors(dor, o, st, keys=keys)
    floor = helpers.field_errors(dor, o, helpers.oracle_state(dor, o2), keys=keys)
    _say("  err  ", _fmt(errs))
    _say("  floor", _fmt(floor))
    ch.close()
    bad = helpers.within(errs, floor, 1e-10)
    assert not bad, bad


def test_large_planes_refuse_les_and_mixed():
    p, gy, ic = configs.channel_re180(nx=256, ny=17, nz=256)
    ch = channel.Channel(p, gy)
    with pytest.raises(channel.DiabloError):
        ch.set_les(True)
    with pytest.raises(channel.DiabloError):
        ch.set_mixed_precision(True)
    ch.close()
