"""Pins the oracle outputs that the whole-step invariants leave open (CPU only):

- the pressure RK_CHAN_1 accumulates, CP += phi/T4 on J = JSTART..JEND
  (channel.f:1059-1065): for the Taylor-Green vortex, whose advection is the
  gradient of p_TG = (cos 2x + cos 2z)/4, CP after substep 1 is
  BETA_BAR(1) A^2 p_TG and after substep 2 (ZETA_BAR(2) A^2 + BETA_BAR(2)
  a1^2) p_TG, with a1 the amplitude after substep 1; CP(0), CP(1) and
  CP(NY+1) stay zero;
- every branch of COURANT (courant.f:1-65) made the binding minimum in turn,
  its two clamps and H_BAR;
- CREATE_FLOW_CHAN's noise-free profiles (set_ics.f:26-36, 51-73) and the
  boundary rows of APPLY_BC_VEL_LOWER/UPPER (channel.f:2285-2492) as the
  discrete boundary conditions they impose;
- the Gaussian generator that stands in for RANDOM_NUMBER.
"""
import math

import numpy as np
import pytest

from diablo_amd import configs

BETA = (1.0, 25.0 / 8.0, 9.0 / 4.0)
ZETA = (0.0, -17.0 / 8.0, -5.0 / 4.0)
HFRAC = (8.0 / 15.0, 2.0 / 15.0, 5.0 / 15.0)


def _set_phys(dor, o, which, values_jzx):
    v = o.rarr(which)
    v[:] = 0.0
    v[:, : o.nz, : o.nx] = values_jzx
    o.fft_to_fourier(which, 0, o.ny + 1)


def _metrics(gy):
    """GYF, DY, DYF on 0..NY+1 from GY(1..NY+1) (channel.f:1640-1652)."""
    ny = len(gy) - 1
    GY = np.zeros(ny + 2)
    GY[1:] = gy
    GYF = np.zeros(ny + 2)
    GYF[1: ny + 1] = 0.5 * (GY[1: ny + 1] + GY[2: ny + 2])
    GYF[0] = 2 * GYF[1] - GYF[2]
    GYF[ny + 1] = 2 * GYF[ny] - GYF[ny - 1]
    DY = np.zeros(ny + 2)
    DY[1:] = GYF[1:] - GYF[:-1]
    DYF = np.zeros(ny + 2)
    DYF[1: ny + 1] = GY[2: ny + 2] - GY[1: ny + 1]
    DYF[ny + 1] = DYF[ny]
    return GYF, DY, DYF


# ---------------------------------------------------------------- pressure
@pytest.mark.parametrize("amp", [1.0, 0.6])
def test_taylor_green_pressure_accumulation(oracle_mod, amp):
    dor = oracle_mod
    nx, ny, nz, nu, dt = 16, 9, 16, 0.2, 0.05
    p = configs.base_params(nx, ny, nz, 0)
    p.update(nu=nu, lx=2 * math.pi, lz=2 * math.pi, delta_t=dt, f_type=3,
             u_bc_ymin=1, u_bc_ymax=1, w_bc_ymin=1, w_bc_ymax=1, v_bc_ymin=0, v_bc_ymax=0)
    o = dor.Oracle(p)
    o.set_grid_gy(configs.make_gy(ny, 2.0, 0.0, 2))
    x = np.arange(nx) * 2 * math.pi / nx
    z = np.arange(nz) * 2 * math.pi / nz
    X, Z = np.meshgrid(x, z)
    _set_phys(dor, o, dor.CU1, np.broadcast_to(amp * np.sin(X) * np.cos(Z), (ny + 2, nz, nx)))
    _set_phys(dor, o, dor.CU3, np.broadcast_to(-amp * np.cos(X) * np.sin(Z), (ny + 2, nz, nx)))
    o.set_clock(0.0, dt, 1)
    tnkz = o.tnkz

    # p_TG = (cos 2x + cos 2z)/4: coefficient 1/8 at kx = 2, kz = +-2
    phat = np.zeros((nz + 2, nx // 2 + 1), np.complex128)
    phat[0, 2] = 0.125
    phat[2, 0] = 0.125
    phat[tnkz + 1 - 2, 0] = 0.125

    def check_cp(scale):
        cp = o.carr(dor.CP)
        for J in range(2, ny + 1):  # JSTART..JEND
            np.testing.assert_allclose(cp[J], scale * phat, rtol=0, atol=1e-14 * max(1.0, abs(scale)))
        for J in (0, 1, ny + 1):  # never updated by the loop
            assert np.abs(cp[J]).max() == 0.0, J

    o.rk_chan_1(1)
    # single Fourier mode, k^2 = 2: the explicit horizontal viscosity alone
    # changes the amplitude, a1 = A (1 - 2 nu T5), T5 = BETA_BAR(1) H_BAR(1)
    a1 = amp * (1.0 - 2.0 * nu * BETA[0] * HFRAC[0] * dt)
    np.testing.assert_allclose(o.carr(dor.CU1)[2: ny, 1, 1].imag * -4.0, a1, rtol=1e-13)
    check_cp(BETA[0] * amp * amp)
    o.rk_chan_1(2)
    check_cp(ZETA[1] * amp * amp + BETA[1] * a1 * a1)
    # T4 = T5 at substep 1, so only substep 2 tells the divisor apart: phi/T5
    # would give A^2 + (ZETA A^2 - A^2 + BETA a1^2)/BETA, off by ~1e-2 here,
    # and leaving out the T3 carry of CF would give BETA a1^2
    exact = ZETA[1] * amp * amp + BETA[1] * a1 * a1
    by_t5 = amp * amp + (ZETA[1] * amp * amp - amp * amp + BETA[1] * a1 * a1) / BETA[1]
    no_t3 = BETA[1] * a1 * a1
    assert min(abs(exact - by_t5), abs(exact - no_t3)) > 1e-3 * amp * amp


# ----------------------------------------------------------------- COURANT
def _courant_case(dor, *, n_th=0, gy=None, u=(0.0, 0.0, 0.0), planes=None, **kw):
    nx, ny, nz = 8, 9, 4
    p = configs.base_params(nx, ny, nz, n_th)
    p.update(nu=0.01, cfl=0.5, variable_dt=1, lx=1.0, lz=2.0)
    p.update(kw)
    o = dor.Oracle(p)
    o.set_grid_gy(configs.make_gy(ny, 2.0, 1.2, 1) if gy is None else gy)
    for w, val in zip((dor.CU1, dor.CU2, dor.CU3), u):
        v = o.rarr(w)
        v[:] = 0.0
        if planes is None:
            v[1: ny + 1, :nz, :nx] = val
        else:
            for J in planes:
                v[J, :nz, :nx] = val
    o.courant()
    _, dt, _ = o.clock()
    return o, p, dt


def test_courant_viscous_limit_config1(oracle_mod):
    """Config 1 (Re_tau = 180, tanh grid CS = 1.75): with the flow at rest the
    viscous limit 0.5 DY(1)^2 / NU binds; DY(1) < DX."""
    dor = oracle_mod
    p, gy, _ = configs.channel_re180(nx=16, nz=16)  # DX = 4 pi / 16, same y grid
    p.update(variable_dt=1)
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    for w in (dor.CU1, dor.CU2, dor.CU3):
        o.rarr(w)[:] = 0.0
    o.courant()
    _, dt, _ = o.clock()
    _, DY, _ = _metrics(gy)
    assert DY[1] < p["lx"] / p["nx"]
    expect = 0.5 * DY[1] ** 2 / p["nu"]
    assert abs(dt - expect) <= 2e-15 * expect
    # the figure the full 128x129x128 deck runs at (DX = 4 pi/128 > DY(1) too)
    assert abs(dt - 1.0768573899829742e-3) <= 2e-15 * dt
    hb = o.vec(dor.HBAR)
    for i in range(3):
        assert hb[i] == dt * HFRAC[i]


def test_courant_viscous_limit_takes_smaller_of_dx_dy(oracle_mod):
    dor = oracle_mod
    # uniform grid of height 200: DY = 25 > DX = 1/8
    gy = configs.make_gy(9, 200.0, 0.0, 2)
    o, p, dt = _courant_case(dor, gy=gy)
    dx = p["lx"] / p["nx"]
    assert abs(dt - 0.5 * dx * dx / p["nu"]) <= 2e-15 * dt


def test_courant_scalar_diffusivity(oracle_mod):
    dor = oracle_mod
    o0, _, dt0 = _courant_case(dor)
    o, p, dt = _courant_case(dor, n_th=1, pr=[0.25], ri=[0.0])
    assert abs(dt - 0.25 * dt0) <= 2e-15 * dt
    # PR >= 1 does not bind
    o, p, dt = _courant_case(dor, n_th=1, pr=[3.0], ri=[0.0])
    assert dt == dt0


def test_courant_rotation(oracle_mod):
    dor = oracle_mod
    o, p, dt = _courant_case(dor, i_ro=-1000.0)
    expect = 2.0 * math.pi / 1000.0 / 20.0
    assert abs(dt - expect) <= 2e-15 * expect


def test_courant_buoyancy_default_real_literal(oracle_mod):
    """0.1 in courant.f:30 is a default-real literal: 0.1f, not 0.1d0."""
    dor = oracle_mod
    o, p, dt = _courant_case(dor, n_th=1, ri=[50.0], th_bc_ymin_c1=[-2.0], pr=[1.0])
    expect32 = float(np.float32(0.1)) * 2.0 * math.pi / 10.0
    expect64 = 0.1 * 2.0 * math.pi / 10.0
    assert abs(dt - expect32) <= 2e-15 * dt
    assert abs(dt - expect64) > 1e-9 * dt


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_courant_cfl_each_direction(oracle_mod, comp):
    dor = oracle_mod
    vel = [0.0, 0.0, 0.0]
    vel[comp] = -40.0  # |U| enters
    if comp == 1:
        # U2 on one interior plane only: CFL DY(J)/|U2| at that J (J = 3,
        # off the centre of the symmetric grid, where DY(J) != DY(J+1))
        o, p, dt = _courant_case(dor, u=vel, planes=[3])
        _, DY, _ = _metrics(configs.make_gy(9, 2.0, 1.2, 1))
        assert abs(DY[3] - DY[4]) > 1e-3 * DY[3]
        expect = p["cfl"] * DY[3] / 40.0
    else:
        o, p, dt = _courant_case(dor, u=vel)
        d = p["lx"] / p["nx"] if comp == 0 else p["lz"] / p["nz"]
        expect = p["cfl"] * d / 40.0
    assert abs(dt - expect) <= 2e-15 * expect


def test_courant_cfl_range_j_1_to_ny(oracle_mod):
    """The CFL loop runs over J = 1..NY: velocity on the ghost planes 0 and
    NY+1 does not bind, on plane 1 and plane NY it does."""
    dor = oracle_mod
    _, _, dt0 = _courant_case(dor)
    for J, binds in ((0, False), (10, False), (1, True), (9, True)):
        o, p, dt = _courant_case(dor, u=(1e4, 0.0, 0.0), planes=[J])
        if binds:
            assert dt == p["cfl"] * (p["lx"] / p["nx"]) / 1e4, J
        else:
            assert dt == dt0, J


def test_courant_clamps(oracle_mod):
    dor = oracle_mod
    # dt <= 0 (a negative CFL number): DELTA_T = 1e-4
    o, p, dt = _courant_case(dor, cfl=-0.5, u=(1.0, 0.0, 0.0))
    assert dt == 0.0001
    # dt >= 999 (nothing binds below it): DELTA_T = 999
    o, p, dt = _courant_case(dor, nu=1e-12)
    assert dt == 999.0
    hb = o.vec(dor.HBAR)
    assert hb[0] == 999.0 * (8.0 / 15.0) and hb[1] == 999.0 * (2.0 / 15.0) and hb[2] == 999.0 * (5.0 / 15.0)


# ------------------------------------------------------- initial conditions
@pytest.mark.parametrize("ic_type", [0, 2, 3])
def test_create_flow_profiles(oracle_mod, ic_type):
    """Noise-free CREATE_FLOW_CHAN: the (0,0) mode of U1 is the profile at
    GYF(J) on the interior rows, every other mode and U2, U3 are zero (the
    profile is divergence-free, so REM_DIV leaves it), and the Dirichlet
    rows hold the wall values."""
    dor = oracle_mod
    nx, ny, nz = 8, 17, 4
    p = configs.base_params(nx, ny, nz, 0)
    p.update(u_bc_ymin_c1=0.3, u_bc_ymax_c1=-0.4, w_bc_ymin_c1=0.25, w_bc_ymax_c1=0.7)
    gy = configs.make_gy(ny, 2.0, 1.5, 1)
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    ub = 2.5
    o.create_flow(ic_type, 1, 0.0, 0.0, ub)
    GYF, DY, DYF = _metrics(gy)
    prof = {0: 1.5 * ub * (1.0 - GYF ** 2), 2: GYF, 3: np.tanh(GYF)}[ic_type]
    cu1, cu2, cu3 = (o.carr(w) for w in (dor.CU1, dor.CU2, dor.CU3))
    np.testing.assert_allclose(cu1[2:ny, 0, 0].real, prof[2:ny], rtol=1e-14, atol=1e-15)
    assert np.abs(cu1[2:ny, 0, 0].imag).max() == 0.0
    rest = cu1[:, : o.tnkz + 1, : o.nkx + 1].copy()
    rest[:, 0, 0] = 0
    assert np.abs(rest).max() < 1e-15
    assert np.abs(cu2).max() < 1e-15  # REM_DIV's correction of roundoff only
    # Dirichlet rows: the mean equals the wall value on rows 0, 1 and NY;
    # the upper W ghost row is zeroed ("ghost cell not used")
    assert cu1[0, 0, 0] == 0.3 and cu1[1, 0, 0] == 0.3
    assert cu1[ny, 0, 0] == -0.4 and cu1[ny + 1, 0, 0] == -0.4
    assert cu3[0, 0, 0] == 0.25 and cu3[1, 0, 0] == 0.25
    assert cu3[ny, 0, 0] == 0.7 and cu3[ny + 1, 0, 0] == 0.0
    assert np.abs(cu3[2:ny]).max() < 1e-15
    assert np.abs(o.carr(dor.CP)).max() == 0.0


def _random_velocity(dor, o, seed):
    rng = np.random.default_rng(seed)
    for w in (dor.CU1, dor.CU2, dor.CU3):
        a = o.carr(w)
        a[:] = 0
        blk = a[:, : o.tnkz + 1, : o.nkx + 1]
        blk[:] = rng.standard_normal(blk.shape) + 1j * rng.standard_normal(blk.shape)
        a[:, 0, 0] = a[:, 0, 0].real  # the mean of a real field is real


def test_apply_bc_vel_dirichlet_rows(oracle_mod):
    dor = oracle_mod
    nx, ny, nz = 8, 13, 8
    p = configs.base_params(nx, ny, nz, 0)
    p.update(u_bc_ymin_c1=0.3, u_bc_ymax_c1=-0.4, w_bc_ymin_c1=0.25, w_bc_ymax_c1=0.7,
             v_bc_ymin_c1=0.15, v_bc_ymax_c1=-0.05)
    gy = configs.make_gy(ny, 2.0, 1.5, 1)
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    _random_velocity(dor, o, 1)
    before = [o.carr(w).copy() for w in (dor.CU1, dor.CU2, dor.CU3)]
    o.apply_bc_vel()
    cu1, cu2, cu3 = (o.carr(w) for w in (dor.CU1, dor.CU2, dor.CU3))
    m = (slice(None), slice(0, o.tnkz + 1), slice(0, o.nkx + 1))
    for c, lo, hi in ((cu1, 0.3, -0.4), (cu3, 0.25, 0.7)):
        for J, val in ((0, lo), (1, lo), (ny, hi)):
            blk = c[J][m[1:]].copy()
            assert blk[0, 0] == val
            blk[0, 0] = 0
            assert np.abs(blk).max() == 0.0
    assert cu1[ny + 1, 0, 0] == -0.4 and cu3[ny + 1, 0, 0] == 0.0
    # interior rows untouched
    for a, b in zip((cu1, cu2, cu3), before):
        assert np.array_equal(a[2:ny], b[2:ny])
    # V: the wall value is the average of rows 1 and 2 for every mode at
    # the lower wall, of rows NY and NY+1 for the mean only at the upper
    np.testing.assert_allclose((cu2[1] + cu2[2])[m[1:]], np.full((o.tnkz + 1, o.nkx + 1), 2 * 0.15), atol=1e-15)
    assert np.array_equal(cu2[0], cu2[1])
    assert abs((cu2[ny, 0, 0] + cu2[ny + 1, 0, 0]) - 2 * (-0.05)) < 1e-15
    rest = cu2[ny + 1].copy()
    rest[0, 0] = before[1][ny + 1, 0, 0]
    assert np.array_equal(rest, before[1][ny + 1])
    assert np.array_equal(cu2[ny], before[1][ny])


def test_apply_bc_vel_neumann_rows(oracle_mod):
    dor = oracle_mod
    nx, ny, nz = 8, 13, 8
    p = configs.base_params(nx, ny, nz, 0)
    p.update(u_bc_ymin=1, u_bc_ymax=1, w_bc_ymin=1, w_bc_ymax=1, v_bc_ymin=1, v_bc_ymax=1,
             u_bc_ymin_c1=0.3, u_bc_ymax_c1=-0.4, w_bc_ymin_c1=0.25, w_bc_ymax_c1=0.7,
             v_bc_ymin_c1=0.15, v_bc_ymax_c1=-0.05)
    gy = configs.make_gy(ny, 2.0, 1.5, 1)
    o = dor.Oracle(p)
    o.set_grid_gy(gy)
    GYF, DY, DYF = _metrics(gy)
    _random_velocity(dor, o, 2)
    before = [o.carr(w).copy() for w in (dor.CU1, dor.CU2, dor.CU3)]
    o.apply_bc_vel()
    cu1, cu2, cu3 = (o.carr(w) for w in (dor.CU1, dor.CU2, dor.CU3))
    m = (slice(0, o.tnkz + 1), slice(0, o.nkx + 1))

    def fluct_equal(a, b):
        x, y = a[m].copy(), b[m].copy()
        x[0, 0] = y[0, 0] = 0
        return np.array_equal(x, y)

    for c, lo, hi in ((cu1, 0.3, -0.4), (cu3, 0.25, 0.7)):
        # zero gradient for the fluctuations, the prescribed one for the mean:
        # one-sided differences over DY(2), DY(1) below and DY(NY) (twice) above
        assert fluct_equal(c[1], c[2]) and fluct_equal(c[0], c[1])
        assert fluct_equal(c[ny], c[ny - 1]) and fluct_equal(c[ny + 1], c[ny])
        np.testing.assert_allclose((c[2, 0, 0] - c[1, 0, 0]) / DY[2], lo, rtol=1e-12)
        np.testing.assert_allclose((c[1, 0, 0] - c[0, 0, 0]) / DY[1], lo, rtol=1e-12)
        np.testing.assert_allclose((c[ny, 0, 0] - c[ny - 1, 0, 0]) / DY[ny], hi, rtol=1e-12)
        np.testing.assert_allclose((c[ny + 1, 0, 0] - c[ny, 0, 0]) / DY[ny], hi, rtol=1e-12)
    # V: the lower rows sit on GY(1), GY(2) and take DYF(1); the upper mean
    # takes its gradient only into the ghost row
    assert fluct_equal(cu2[1], cu2[2]) and fluct_equal(cu2[0], cu2[1])
    np.testing.assert_allclose((cu2[2, 0, 0] - cu2[1, 0, 0]) / DYF[1], 0.15, rtol=1e-12)
    np.testing.assert_allclose((cu2[1, 0, 0] - cu2[0, 0, 0]) / DYF[1], 0.15, rtol=1e-12)
    assert np.array_equal(cu2[ny], cu2[ny - 1])
    np.testing.assert_allclose((cu2[ny + 1, 0, 0] - cu2[ny, 0, 0]) / DY[ny], -0.05, rtol=1e-12)
    for a, b in zip((cu1, cu2, cu3), before):
        assert np.array_equal(a[2:ny], b[2:ny])


def test_gaussian_generator_statistics(oracle_mod):
    """The counter-based generator behind the seeded ICs: standard normal
    moments and tail mass, and distinct streams for distinct seeds."""
    dor = oracle_mod
    n = 40000
    x = np.array([dor.gauss(7, c) for c in range(n)])
    assert abs(x.mean()) < 4.5 / math.sqrt(n)
    assert abs(x.var() - 1.0) < 4.5 * math.sqrt(2.0 / n)
    assert abs(np.mean(np.abs(x) > 1.96) - 0.05) < 4.5 * math.sqrt(0.05 * 0.95 / n)
    assert abs(np.mean(x ** 3)) < 4.5 * math.sqrt(15.0 / n)
    y = np.array([dor.gauss(8, c) for c in range(n)])
    assert abs(np.corrcoef(x, y)[0, 1]) < 4.5 / math.sqrt(n)
