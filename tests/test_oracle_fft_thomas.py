"""Pins the oracle's transforms and tridiagonal sweeps against library
routines and closed forms (CPU only).

FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL (mpi.f:1477-1606): unnormalised
FFTW DFTs (forward e^{-i}), x r2c scaled by 1/NX and truncated at NKX, z c2c
scaled by 1/NZ and packed as kz = 0..NKZ, -NKZ..-1.  Checked against
numpy.fft (pocketfft), single modes, and the round trip.
THOMAS_*_MPI (mpi.f:630-865): against scipy.linalg.solve_banded."""
import math

import numpy as np
import pytest
import scipy.linalg

from tests import helpers


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 16, 30, 64, 96, 128, 192, 256])
def test_c2c_matches_numpy(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = oracle_mod.fft_c2c(x, -1)
    b = oracle_mod.fft_c2c(x, +1)
    s = np.abs(np.fft.fft(x)).max()
    assert np.abs(f - np.fft.fft(x)).max() <= 1e-14 * s
    assert np.abs(b - np.fft.ifft(x) * n).max() <= 1e-14 * s


def _params(nx, nz, ny=5):
    from diablo_amd import configs
    p = configs.base_params(nx, ny, nz, 0)
    p.update(lx=2 * math.pi, lz=2 * math.pi)
    return p


def _pack_numpy(spec_full, nkx, nkz):
    """numpy spectrum [..., z-freq, kx] -> reference packing [..., 0..TNKZ, 0..NKX]."""
    nz = spec_full.shape[-2]
    out = np.zeros(spec_full.shape[:-2] + (2 * nkz + 1, nkx + 1), np.complex128)
    out[..., : nkz + 1, :] = spec_full[..., : nkz + 1, : nkx + 1]
    for K in range(1, nkz + 1):
        out[..., nkz + K, :] = spec_full[..., nz - 1 + K - nkz, : nkx + 1]
    return out


@pytest.mark.parametrize("nx,nz", [(16, 8), (128, 4), (24, 12), (64, 64)])
def test_fft_xz_to_fourier_matches_numpy(oracle_mod, nx, nz):
    dor = oracle_mod
    o = dor.Oracle(_params(nx, nz))
    ny = o.ny
    rng = np.random.default_rng(nx * nz)
    phys = rng.standard_normal((ny + 2, nz, nx))
    v = o.rarr(dor.CS1)
    v[:] = 0
    v[:, :nz, :nx] = phys
    o.fft_to_fourier(dor.CS1, 0, ny + 1)
    got = o.carr(dor.CS1)
    nkx, nkz = nx // 3, nz // 3
    ref = _pack_numpy(np.fft.fft(np.fft.rfft(phys, axis=-1) / nx, axis=-2) / nz, nkx, nkz)
    np.testing.assert_allclose(got[:, : 2 * nkz + 1, : nkx + 1], ref, rtol=0, atol=1e-15 * np.abs(ref).max() * 10)
    # truncated modes are exactly zero (mpi.f:1501-1506)
    assert np.all(got[:, : 2 * nkz + 1, nkx + 1:] == 0)


@pytest.mark.parametrize("nx,nz", [(16, 8), (128, 4), (24, 12), (64, 64)])
def test_fft_xz_to_physical_matches_numpy(oracle_mod, nx, nz):
    dor = oracle_mod
    o = dor.Oracle(_params(nx, nz))
    ny = o.ny
    nkx, nkz = nx // 3, nz // 3
    rng = np.random.default_rng(7 + nx)
    spec = rng.standard_normal((ny + 2, 2 * nkz + 1, nkx + 1)) + 1j * rng.standard_normal((ny + 2, 2 * nkz + 1, nkx + 1))
    c = o.carr(dor.CS1)
    c[:] = 0
    c[:, : 2 * nkz + 1, : nkx + 1] = spec
    o.fft_to_physical(dor.CS1, 0, ny + 1)
    got = o.rarr(dor.CS1)[:, :nz, :nx]
    full = np.zeros((ny + 2, nz, nx // 2 + 1), np.complex128)
    full[:, : nkz + 1, : nkx + 1] = spec[:, : nkz + 1]
    for K in range(1, nkz + 1):
        full[:, nz - 1 + K - nkz, : nkx + 1] = spec[:, nkz + K]
    zt = np.fft.ifft(full, axis=-2) * nz
    ref = np.fft.irfft(zt, n=nx, axis=-1) * nx  # imaginary part of the DC term ignored, as c2r
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_single_mode_amplitudes(oracle_mod):
    dor = oracle_mod
    nx, nz = 32, 16
    o = dor.Oracle(_params(nx, nz))
    x = np.arange(nx) * 2 * math.pi / nx
    z = np.arange(nz) * 2 * math.pi / nz
    X, Z = np.meshgrid(x, z)
    v = o.rarr(dor.CS1)
    v[:] = 0
    v[:, :nz, :nx] = np.cos(3 * X) + 2 * np.sin(2 * Z)
    o.fft_to_fourier(dor.CS1, 0, o.ny + 1)
    c = o.carr(dor.CS1)[2]
    nkz = nz // 3
    assert abs(c[0, 3] - 0.5) < 1e-15           # cos(3x) -> 1/2 at kx=3
    assert abs(c[2, 0] - (-1j)) < 1e-15          # 2 sin(2z) -> -i at kz=+2
    assert abs(c[nkz + (nkz + 1 - 2), 0] - 1j) < 1e-15  # and +i at kz=-2 (packed)
    mask = np.ones_like(c, bool)
    mask[0, 3] = mask[2, 0] = mask[nkz + nkz - 1, 0] = False
    assert np.abs(c[: 2 * nkz + 1, : nx // 3 + 1][mask[: 2 * nkz + 1, : nx // 3 + 1]]).max() < 1e-15


def test_round_trip_of_dealiased_field(oracle_mod):
    dor = oracle_mod
    nx, nz = 48, 24
    o = dor.Oracle(_params(nx, nz))
    nkx, nkz = nx // 3, nz // 3
    rng = np.random.default_rng(5)
    phys = rng.standard_normal((o.ny + 2, nz, nx))
    v = o.rarr(dor.CS1)
    v[:] = 0
    v[:, :nz, :nx] = phys
    o.fft_to_fourier(dor.CS1, 0, o.ny + 1)
    spec = o.carr(dor.CS1)[:, : 2 * nkz + 1, : nkx + 1].copy()
    o.fft_to_physical(dor.CS1, 0, o.ny + 1)
    o.fft_to_fourier(dor.CS1, 0, o.ny + 1)
    back = o.carr(dor.CS1)[:, : 2 * nkz + 1, : nkx + 1]
    assert np.abs(back - spec).max() <= 1e-15 * 10 * np.abs(spec).max()


def _banded(a, b, c, d):
    n = len(b)
    ab = np.zeros((3, n), dtype=np.result_type(a, d))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    return scipy.linalg.solve_banded((1, 1), ab, d)


def test_thomas_real_matches_solve_banded(oracle_mod):
    rng = np.random.default_rng(1)
    ny, nx = 33, 7
    a = rng.uniform(-1, 0, (ny + 2, nx))
    c = rng.uniform(-1, 0, (ny + 2, nx))
    b = 2.5 + rng.uniform(0, 1, (ny + 2, nx))
    a[0] = 0
    c[-1] = 0
    d = rng.standard_normal((ny + 2, nx))
    ref = np.stack([_banded(a[:, i], b[:, i], c[:, i], d[:, i]) for i in range(nx)], axis=1)
    A, B, C, G = (np.ascontiguousarray(t) for t in (a.copy(), b.copy(), c.copy(), d.copy()))
    oracle_mod.thomas_real(A, B, C, G)
    np.testing.assert_allclose(G, ref, rtol=1e-13, atol=1e-14)


def test_thomas_complex_matches_solve_banded_and_zero_guard(oracle_mod):
    """Rows 1..NY are solved (mpi.f:708-865); a column whose pivots are zero
    (the uninitialised column I=NXP of REM_DIV) returns zeros."""
    rng = np.random.default_rng(2)
    ny, nxp = 21, 5
    a = rng.uniform(0.1, 1, (ny + 2, nxp + 1))
    c = rng.uniform(0.1, 1, (ny + 2, nxp + 1))
    b = -(a + c) - rng.uniform(0.5, 2, (ny + 2, nxp + 1))
    d = rng.standard_normal((ny + 2, nxp + 1)) + 1j * rng.standard_normal((ny + 2, nxp + 1))
    a[:, nxp] = b[:, nxp] = c[:, nxp] = 0.0  # guarded column
    A, B, C = a.copy(), b.copy(), c.copy()
    G = d.copy()
    oracle_mod.thomas_complex(A, B, C, G)
    for i in range(nxp):
        aa, bb, cc, dd = a[1:ny + 1, i].copy(), b[1:ny + 1, i], c[1:ny + 1, i].copy(), d[1:ny + 1, i]
        aa[0] = 0
        cc[-1] = 0
        ref = _banded(aa, bb, cc, dd)
        np.testing.assert_allclose(G[1:ny + 1, i], ref, rtol=1e-12, atol=1e-13)
    assert np.all(G[1:ny + 1, nxp] == 0)
