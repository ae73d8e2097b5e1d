"""FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL of the HIP library against the
oracle's restatement of mpi.f:1477-1606 on random planes (FP64, relative
L2 <= 1e-13 per batch).  The decks cover the LDS row transforms (16 x 8
and 128 x 4 planes), the register rows for NX = 128 (channel_re180) and
for NX = 64 (les_channel, 64 x 64 planes)."""
import numpy as np
import pytest

from diablo_amd import channel, configs
from tests import devmem

pytestmark = pytest.mark.gpu

GRIDS = {"small_case": {}, "kh_shear_layer": {}, "channel_re180": {}, "les_channel": {}}


@pytest.fixture(scope="module")
def dor():
    from oracle import dor as d
    d.load()
    return d


@pytest.mark.parametrize("name", list(GRIDS))
def test_forward_and_inverse_planes(dor, name):
    print(f"[fft] {name}", flush=True)
    p, gy, ic = configs.CONFIGS[name]()
    ch = channel.Channel(p, gy)
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    nx, nz, ny = p["nx"], p["nz"], p["ny"]
    nkx, tnkz = nx // 3, 2 * (nz // 3)
    npl = ny + 2
    rng = np.random.default_rng(11)
    # forward: physical planes -> de-aliased spectrum
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
    got = df.get(np.complex128, ref.shape)
    e_f = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    # inverse: random de-aliased spectrum -> physical planes
    spec = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    c = o.carr(dor.CS1)
    c[:] = 0.0
    c[:, : tnkz + 1, : nkx + 1] = spec
    o.fft_to_physical(dor.CS1, 0, ny + 1)
    ref_p = o.rarr(dor.CS1)[:, :nz, :nx].copy()
    df.put(spec)
    ch.fft_to_physical_dev(df.ptr, dp.ptr, npl)
    ch.synchronize()
    got_p = dp.get(np.float64, ref_p.shape)
    e_i = np.linalg.norm(got_p - ref_p) / np.linalg.norm(ref_p)
    print(f"  forward rel L2 {e_f:.2e}  inverse rel L2 {e_i:.2e}", flush=True)
    dp.free()
    df.free()
    ch.close()
    assert e_f <= 1e-13 and e_i <= 1e-13
