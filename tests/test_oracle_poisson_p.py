"""Pins the oracle's POISSON_P_CHAN restatement (channel.f:1217-1453 with
APPLY_BC_POISSON_MPI, mpi.f:1280-1310) against closed forms (CPU only).

The right-hand side is sum_ij (dui/dxj)(duj/dxi) plus RI d(theta)/dy, and
the pressure solves (d2/dx2 + d2/dz2 + delta_yy) p = -rhs with Neumann rows
at both walls (a Dirichlet gauge at the lower wall for the (0,0) mode).
Every case below isolates one term of the right-hand side with fields whose
discrete derivatives and products are exact, so a dropped term, a wrong
sign, factor or index fails it:

- x-z terms give y-independent pressure modes, p = rhs_k / |k|^2;
- the (0,0) mode with a constant right-hand side c is the quadratic
  p = -c (y^2 - y1^2)/2 + b (y - y1), b = c (y_NY + y_NY-1)/2, which the
  3-point operator reproduces exactly on a uniform grid;
- terms with a y-derivative use fields linear in y, whose centred and
  one-sided differences are exact.
"""
import math

import numpy as np
import pytest

from diablo_amd import configs


def _setup(dor):
    nx, ny, nz = 16, 9, 8
    p = configs.base_params(nx, ny, nz, 0)
    p.update(lx=2 * math.pi, lz=2 * math.pi)
    o = dor.Oracle(p)
    o.set_grid_gy(configs.make_gy(ny, 2.0, 0.0, 2))  # uniform
    x = np.arange(nx) * 2 * math.pi / nx
    z = np.arange(nz) * 2 * math.pi / nz
    return o, p, x, z


def _set_phys(o, which, values_jzx):
    v = o.rarr(which)
    v[:] = 0.0
    v[:, : o.nz, : o.nx] = values_jzx
    o.fft_to_fourier(which, 0, o.ny + 1)


def _pressure_phys(dor, o):
    """CP (planes 1..NY) in physical space, through a scratch array."""
    cs = o.carr(dor.CR1)
    cs[:] = o.carr(dor.CP)
    o.fft_to_physical(dor.CR1, 0, o.ny + 1)
    return o.rarr(dor.CR1)[1: o.ny + 1, : o.nz, : o.nx].copy()


def _field(o, f):
    return np.broadcast_to(f, (o.ny + 2, o.nz, o.nx))


def test_xz_cross_term(oracle_mod):
    """u1 = cos z, u3 = cos x: rhs = 2 (du1/dz)(du3/dx) = 2 sin x sin z,
    p = sin x sin z on every plane."""
    dor = oracle_mod
    o, p, x, z = _setup(dor)
    _set_phys(o, dor.CU1, _field(o, np.cos(z)[:, None] + 0 * x[None, :]))
    _set_phys(o, dor.CU3, _field(o, np.cos(x)[None, :] + 0 * z[:, None]))
    o.poisson_p()
    want = np.sin(x)[None, None, :] * np.sin(z)[None, :, None]
    np.testing.assert_allclose(_pressure_phys(dor, o), np.broadcast_to(want, (o.ny, o.nz, o.nx)), atol=1e-13)


def test_diagonal_x_term_and_mean_mode(oracle_mod):
    """u1 = sin x: rhs = cos^2 x = 1/2 + cos(2x)/2; the cos 2x mode gives
    p = cos(2x)/8, the mean mode the quadratic with c = 1/2."""
    dor = oracle_mod
    o, p, x, z = _setup(dor)
    _set_phys(o, dor.CU1, _field(o, np.sin(x)[None, :] + 0 * z[:, None]))
    o.poisson_p()
    y = o.vec(dor.GYF)
    ny = o.ny
    c = 0.5
    b = c * (y[ny] + y[ny - 1]) / 2
    mean = -c * (y[1: ny + 1] ** 2 - y[1] ** 2) / 2 + b * (y[1: ny + 1] - y[1])
    want = mean[:, None, None] + (np.cos(2 * x) / 8)[None, None, :]
    np.testing.assert_allclose(_pressure_phys(dor, o), np.broadcast_to(want, (ny, o.nz, o.nx)), atol=1e-13)


def test_diagonal_z_and_y_terms(oracle_mod):
    """u3 = sin z: (du3/dz)^2 = 1/2 + cos(2z)/2; u2 = GY (linear, on the
    staggered grid): (du2/dy)^2 = 1 on every plane.  Mean mode c = 3/2."""
    dor = oracle_mod
    o, p, x, z = _setup(dor)
    gy = o.vec(dor.GY)
    _set_phys(o, dor.CU3, _field(o, np.sin(z)[:, None] + 0 * x[None, :]))
    _set_phys(o, dor.CU2, np.broadcast_to(gy[:, None, None], (o.ny + 2, o.nz, o.nx)))
    o.poisson_p()
    y = o.vec(dor.GYF)
    ny = o.ny
    c = 1.5
    b = c * (y[ny] + y[ny - 1]) / 2
    mean = -c * (y[1: ny + 1] ** 2 - y[1] ** 2) / 2 + b * (y[1: ny + 1] - y[1])
    want = mean[:, None, None] + (np.cos(2 * z) / 8)[None, :, None]
    np.testing.assert_allclose(_pressure_phys(dor, o), np.broadcast_to(want, (ny, o.nz, o.nx)), atol=1e-13)


@pytest.mark.parametrize("comp", ["u1", "u3"])
def test_y_cross_terms(oracle_mod, comp):
    """u1 = GYF, u2 = sin x: rhs = 2 (du1/dy)(du2/dx) = 2 cos x, p = 2 cos x.
    u3 = GYF, u2 = sin z: rhs = 2 (du3/dy)(du2/dz) = 2 cos z, p = 2 cos z."""
    dor = oracle_mod
    o, p, x, z = _setup(dor)
    gyf = o.vec(dor.GYF)
    lin = np.broadcast_to(gyf[:, None, None], (o.ny + 2, o.nz, o.nx))
    if comp == "u1":
        _set_phys(o, dor.CU1, lin)
        _set_phys(o, dor.CU2, _field(o, np.sin(x)[None, :] + 0 * z[:, None]))
        want = 2 * np.cos(x)[None, None, :] + 0 * z[None, :, None]
    else:
        _set_phys(o, dor.CU3, lin)
        _set_phys(o, dor.CU2, _field(o, np.sin(z)[:, None] + 0 * x[None, :]))
        want = 2 * np.cos(z)[None, :, None] + 0 * x[None, None, :]
    o.poisson_p()
    np.testing.assert_allclose(_pressure_phys(dor, o), np.broadcast_to(want, (o.ny, o.nz, o.nx)), atol=1e-12)


def test_buoyancy_term(oracle_mod):
    """theta = GYF cos x, RI = 0.3: rhs = RI cos x, p = 0.3 cos x."""
    dor = oracle_mod
    nx, ny, nz = 16, 9, 8
    p = configs.base_params(nx, ny, nz, 1)
    p.update(lx=2 * math.pi, lz=2 * math.pi)
    p["ri"][0] = 0.3
    o = dor.Oracle(p)
    o.set_grid_gy(configs.make_gy(ny, 2.0, 0.0, 2))
    x = np.arange(nx) * 2 * math.pi / nx
    gyf = o.vec(dor.GYF)
    th = o.rarr(dor.CTH)
    th = th.reshape(-1, *th.shape[-3:])[0] if th.ndim == 4 else th
    th[:] = 0.0
    th[:, :nz, :nx] = gyf[:, None, None] * np.cos(x)[None, None, :]
    o.fft_to_fourier(dor.CTH, 0, ny + 1)
    o.poisson_p()
    want = 0.3 * np.cos(x)[None, None, :] + np.zeros((ny, nz, 1))
    np.testing.assert_allclose(_pressure_phys(dor, o), want, atol=1e-13)


def test_rhs_on_stretched_grid(oracle_mod):
    """On the tanh grid (CS = 1.75), u2 = GY has (U2(J+1)-U2(J))/DYF(J) = 1
    exactly, so the right-hand side left in CS1 is 1 in the mean mode of
    planes 2..NY (a one-sided difference over DY instead of DYF, or a
    centred one, is not)."""
    dor = oracle_mod
    nx, ny, nz = 16, 17, 8
    p = configs.base_params(nx, ny, nz, 0)
    o = dor.Oracle(p)
    o.set_grid_gy(configs.make_gy(ny, 2.0, 1.75, 1))
    gy = o.vec(dor.GY)
    assert np.ptp(np.diff(gy[1:])) > 1e-3  # really stretched
    v = o.rarr(dor.CU2)
    v[:] = 0.0
    v[:, :nz, :nx] = gy[:, None, None]
    o.fft_to_fourier(dor.CU2, 0, ny + 1)
    o.poisson_p()
    np.testing.assert_allclose(o.carr(dor.CS1)[2: ny + 1, 0, 0], 1.0, rtol=0, atol=1e-13)
