"""Pins the oracle's wall-normal flux and interpolation forms of RK_CHAN_1
(channel.f:467-571 velocity, 699-720 scalar, 209-243 buoyancy) with closed
forms on a uniform grid (CPU only).

After one substep at RK 1, CF holds the explicit terms of the start state:
  CF1 = -nu k^2 CU1 - ikz(U3U1) - ikx(U1U1) - d/dy(U1 U2)
  CF2 = -nu k^2 CU2 - ikx(interp(U1) U2) - ikz(interp(U3) U2) - d/dy(U2 U2)
        + RI grav_y interp(TH)                      ((0,0) mode skipped)
  CF3 = -nu k^2 CU3 - ikx(U3U1) - ikz(U3U3) - d/dy(U3 U2)
  CFTH = -(nu/Pr) k^2 CTH - ikx(U1 TH) - ikz(U3 TH) - d/dy(TH U2).
Fields linear in y times one x or z mode make every staggered difference
and every midpoint interpolation exact on a uniform grid, so each flux
below has a closed form.  A wrong plane (J-1 for J+1), factor (1/2), sign
or derivative direction fails it; every other mode must stay zero.  U2 is
taken both constant and linear in y so that the U2 plane each flux reads is
pinned too.  The order of the DYF weights in the two interpolations
(channel.f:472-473 for velocity, 214 for buoyancy) is pinned on a stretched
grid: for a field f linear in y, the buoyancy form is the exact linear
interpolation f(GY(J)), and the velocity form (which weights the other way
round) gives f(GY(J)) + f'(y) (DYF(J) - DYF(J-1))/2."""
import math

import numpy as np
import pytest

from diablo_amd import configs

NX, NY, NZ = 16, 9, 8
NU, C = 0.01, 0.5


def _oracle(dor, nth=0, ri=0.0, pr=1.0, cs=0.0):
    p = configs.base_params(NX, NY, NZ, nth)
    p.update(nu=NU, lx=2 * math.pi, lz=2 * math.pi, delta_t=1e-3, f_type=3, i_ro=0.0, beta=1.0,
             u_bc_ymin=1, u_bc_ymax=1, w_bc_ymin=1, w_bc_ymax=1, v_bc_ymin=0, v_bc_ymax=0)
    if nth:
        p["ri"][0] = ri
        p["pr"][0] = pr
        p["grav_y"] = 1.0
        p["th_bc_ymin"] = [1]
        p["th_bc_ymax"] = [1]
    o = dor.Oracle(p)
    o.set_grid_gy(configs.make_gy(NY, 2.0, cs, 2))
    o.set_clock(0.0, p["delta_t"], 1)
    return o


def _set(o, which, f_jzx):
    v = o.rarr(which)
    if v.ndim == 4:
        v = v[0]
    v[:] = 0.0
    v[:, :NZ, :NX] = f_jzx
    o.fft_to_fourier(which, 0, NY + 1)


def _grid():
    x = np.arange(NX) * 2 * math.pi / NX
    z = np.arange(NZ) * 2 * math.pi / NZ
    return x, z


def _expect_only(cf, rows, modes):
    """cf[J, K, I] equals modes[(K, I)][J] on rows, and 0 elsewhere there."""
    for J in rows:
        got = cf[J, : 2 * (NZ // 3) + 1, : NX // 3 + 1].copy()
        for (K, I), val in modes.items():
            np.testing.assert_allclose(got[K, I], val[J], rtol=0, atol=1e-13, err_msg=f"J={J} mode {(K, I)}")
            got[K, I] = 0.0
        assert np.abs(got).max() < 1e-13, (J, np.abs(got).max())


@pytest.mark.parametrize("s", [0, 1])
@pytest.mark.parametrize("comp", [1, 3])
def test_flux_and_interp_of_u_times_u2(oracle_mod, comp, s):
    """u = GYF cos(x) (comp 1) or GYF cos(z) (comp 3), u2 = C + s GY:
    d/dy(u U2) = (C + 2 s GYF) cos, interp(u) U2 = GY (C + s GY) cos,
    u u = GYF^2 cos^2 and d/dy(U2 U2) = 2 s (C + s GY)."""
    dor = oracle_mod
    o = _oracle(dor)
    x, z = _grid()
    gyf, gy = o.vec(dor.GYF).copy(), o.vec(dor.GY).copy()
    wave = (np.cos(x)[None, :] + 0 * z[:, None]) if comp == 1 else (np.cos(z)[:, None] + 0 * x[None, :])
    _set(o, dor.CU1 if comp == 1 else dor.CU3, gyf[:, None, None] * wave[None])
    _set(o, dor.CU2, (C + s * gy)[:, None, None] + np.zeros((NY + 2, NZ, NX)))
    o.rk_chan_1(1)
    cfu = o.carr(dor.CF1 if comp == 1 else dor.CF3)
    cf2 = o.carr(dor.CF2)
    rows = range(2, NY + 1)
    # CF_u: -nu k^2 u - d/dy(u U2) at |k| = 1; -ik(u u) at |k| = 2, where
    # u u = GYF^2 (1 + cos 2x)/2.  CF2: -ik interp(u) U2 at |k| = 1 and
    # -d/dy(U2 U2) in the mean mode.
    lin = -NU * gyf / 2 - (C + 2 * s * gyf) / 2
    sq = -1j * 2 * gyf ** 2 / 4
    itp = -1j * gy * (C + s * gy) / 2
    mean2 = {(0, 0): -2 * s * (C + s * gy)} if s else {}
    if comp == 1:
        want = {(0, 1): lin, (0, 2): sq}
        want2 = {(0, 1): itp, **mean2}
    else:
        # complex z transform: kz = +1, +2 at K = 1, 2 and kz = -1, -2 at
        # K = TNKZ, TNKZ - 1 (TNKZ = 2 (NZ/3)); -ikz flips sign with kz
        t = 2 * (NZ // 3)
        want = {(1, 0): lin, (2, 0): sq, (t, 0): lin, (t - 1, 0): -sq}
        want2 = {(1, 0): itp, (t, 0): -itp, **mean2}
    _expect_only(cfu, rows, want)
    _expect_only(cf2, rows, want2)


def test_flux_of_u2_squared(oracle_mod):
    """u2 = GY cos x: d/dy(U2U2) = cos^2 x (GYF(J)^2 - GYF(J-1)^2)/DY(J)
    = GY(J) (1 + cos 2x), so CF2 = -GY at (0,0), -nu GY/2 at kx=1 and
    -GY/2 at kx=2."""
    dor = oracle_mod
    o = _oracle(dor)
    x, z = _grid()
    gy = o.vec(dor.GY).copy()
    _set(o, dor.CU2, gy[:, None, None] * (np.cos(x)[None, None, :] + 0 * z[None, :, None]))
    o.rk_chan_1(1)
    _expect_only(o.carr(dor.CF2), range(2, NY + 1), {(0, 0): -gy, (0, 1): -NU * gy / 2, (0, 2): -gy / 2})
    _expect_only(o.carr(dor.CF1), range(2, NY + 1), {})


@pytest.mark.parametrize("s", [0, 1])
def test_scalar_flux_and_buoyancy_interpolation(oracle_mod, s):
    """theta = GYF (1 + cos x), u2 = C + s GY, RI = 0.3, Pr = 0.7:
    d/dy(theta U2) = (C + 2 s GYF)(1 + cos x), so CFTH = -(C + 2 s GYF) in
    the mean mode and -(nu/Pr)(GYF/2) - (C + 2 s GYF)/2 at kx=1; CF2 gains
    RI interp(theta) = RI GY/2 at kx=1 but nothing in the mean mode, which
    buoyancy skips, so there CF2 = -d/dy(U2 U2) = -2 s (C + s GY)."""
    dor = oracle_mod
    ri, pr = 0.3, 0.7
    o = _oracle(dor, nth=1, ri=ri, pr=pr)
    x, z = _grid()
    gyf, gy = o.vec(dor.GYF).copy(), o.vec(dor.GY).copy()
    _set(o, dor.CTH, gyf[:, None, None] * (1 + np.cos(x)[None, None, :] + 0 * z[None, :, None]))
    _set(o, dor.CU2, (C + s * gy)[:, None, None] + np.zeros((NY + 2, NZ, NX)))
    o.rk_chan_1(1)
    cfth = o.carr(dor.CFTH)
    cfth = cfth[0] if cfth.ndim == 4 else cfth
    flux = C + 2 * s * gyf
    _expect_only(cfth, range(2, NY + 1), {(0, 0): -flux, (0, 1): -(NU / pr) * gyf / 2 - flux / 2})
    want2 = {(0, 1): ri * gy / 2}
    if s:
        want2[(0, 0)] = -2 * s * (C + s * gy)
    _expect_only(o.carr(dor.CF2), range(2, NY + 1), want2)


@pytest.mark.parametrize("axis", ["x", "z"])
def test_interpolation_weights_on_stretched_grid(oracle_mod, axis):
    """Stretched grid (CS = 1.5), theta = u = GYF cos (u = u1 along x, u3
    along z), u2 = C, RI = 0.3.  CF2 at |k| = 1 holds RI theta(GY)/2 (real:
    the buoyancy form interpolates exactly) and -ik C (GY + (DYF(J) -
    DYF(J-1))/2)/2 (imaginary: the velocity form's weights); no other CF2
    mode is set."""
    dor = oracle_mod
    ri = 0.3
    o = _oracle(dor, nth=1, ri=ri, cs=1.5)
    x, z = _grid()
    gyf, gy, dyf = (o.vec(w).copy() for w in (dor.GYF, dor.GY, dor.DYF))
    rows = range(2, NY + 1)
    assert max(abs(dyf[j] - dyf[j - 1]) for j in rows) > 1e-2  # the grid is stretched
    if axis == "x":
        wave = np.cos(x)[None, None, :] + 0 * z[None, :, None]
    else:
        wave = np.cos(z)[None, :, None] + 0 * x[None, None, :]
    _set(o, dor.CTH, gyf[:, None, None] * wave)
    _set(o, dor.CU1 if axis == "x" else dor.CU3, gyf[:, None, None] * wave)
    _set(o, dor.CU2, np.full((NY + 2, NZ, NX), C))
    o.rk_chan_1(1)
    vel = gy.copy()
    vel[1:] += (dyf[1:] - dyf[:-1]) / 2
    if axis == "x":
        want = {(0, 1): ri * gy / 2 - 1j * C * vel / 2}
    else:
        t = 2 * (NZ // 3)
        want = {(1, 0): ri * gy / 2 - 1j * C * vel / 2, (t, 0): ri * gy / 2 + 1j * C * vel / 2}
    _expect_only(o.carr(dor.CF2), rows, want)
