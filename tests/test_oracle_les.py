"""Pins the oracle's constant-Smagorinsky LES (les.f:1-336, compute_strain_chan
les.f:526-569, the NU_T terms of channel.f:631-659, 748-760, 793-807,
862-875, 927-941, 991-1005 and the 100-step gate of channel.f:303) with
closed forms and exact equivalences (CPU only).

* A single x (or z) mode u = A cos x, uniform in y, has only S11 = -A sin x
  (S33 = -A sin z): |S| = sqrt(2) A |sin x| at both point sets, so
  NU_T = C_s^2 (DX DY 2 DZ)^(2/3) sqrt(2) A |sin x| on J = 3..NY-1 (rows 1,
  2, NY, NY+1 are zeroed by the one-rank ghost rule) and the SGS divergence
  is -ik DYFW F[|S| S11] on J = 2..NY-1 only, with F the truncated x
  transform taken here with numpy.
* A mean shear u1 = c y is removed before |S| is formed: NU_T = 0.
* A constant NU_T = c is, in the implicit rows and the explicit half, the
  molecular viscosity raised by c (NU_V_SCALE + c/NU), and for a scalar
  with PR = 1 the diffusivity likewise.
* A mean profile with constant discrete flux NU_T(J) (u(J) - u(J-1))/DY(J)
  is a fixed point of the eddy-viscosity terms alone (NU = 0): every wrong
  pairing of NU_T(J), NU_T(J+1) with the differences moves it.
* The scalars' horizontal eddy diffusivity (KAPPA_T = NU_T) has a closed
  form for a scalar varying in z only under the single-x-mode NU_T above."""
import math

import numpy as np
import pytest

from diablo_amd import configs

C_SMAG = 0.13


def _oracle(dor, nx=16, ny=12, nz=8, nth=0, les=1, cnf=0, nu=1e-3, cs=0.0, **over):
    p = configs.base_params(nx, ny, nz, nth)
    p.update(nu=nu, lx=2 * math.pi, lz=math.pi, delta_t=1e-3, f_type=3, les=les, create_new_flow=cnf)
    p.update(over)
    o = dor.Oracle(p)
    o.set_grid_gy(configs.make_gy(ny, 2.0, cs, 2))
    o.set_clock(0.0, p["delta_t"], 1)
    return o, p


def _set_phys(o, which, f_jzx):
    v = o.rarr(which)
    if v.ndim == 4:
        v = v[0]
    v[:] = 0.0
    v[:, : o.nz, : o.nx] = f_jzx
    o.fft_to_fourier(which, 0, o.ny + 1)


def _widths(o, p, dor):
    dx, dz = p["lx"] / o.nx, p["lz"] / o.nz
    dy, dyf = o.vec(dor.DY), o.vec(dor.DYF)
    w = lambda d: -2.0 * C_SMAG ** 2 * (dx * d * 2.0 * dz) ** (2.0 / 3.0)
    return np.array([w(d) for d in dy]), np.array([w(d) for d in dyf])


@pytest.mark.parametrize("axis", ["x", "z"])
def test_single_mode_nut_and_sgs_divergence(oracle_mod, axis):
    dor = oracle_mod
    A = 0.7
    o, p = _oracle(dor, u_bc_ymin=1, u_bc_ymax=1, w_bc_ymin=1, w_bc_ymax=1)
    nx, ny, nz = o.nx, o.ny, o.nz
    x = np.arange(nx) * 2 * math.pi / nx
    z = np.arange(nz) * p["lz"] / nz
    kz1 = 2 * math.pi / p["lz"]
    if axis == "x":
        wave = np.broadcast_to(np.cos(x)[None, :], (nz, nx))
        s = -A * np.broadcast_to(np.sin(x)[None, :], (nz, nx))
    else:
        wave = np.broadcast_to(np.cos(kz1 * z)[:, None], (nz, nx))
        s = -A * kz1 * np.broadcast_to(np.sin(kz1 * z)[:, None], (nz, nx))
    _set_phys(o, dor.CU1 if axis == "x" else dor.CU3, A * np.broadcast_to(wave, (ny + 2, nz, nx)))
    for f in (dor.CF1, dor.CF2, dor.CF3):
        o.carr(f)[:] = 0.0
    o.les_chan()
    dyw, dyfw = _widths(o, p, dor)
    absS = math.sqrt(2.0) * np.abs(s)
    nut = o.rarr(dor.NUT)[:, :nz, :nx]
    for J in range(ny + 2):
        want = -0.5 * dyw[J] * absS if 3 <= J <= ny - 1 else np.zeros_like(absS)
        np.testing.assert_allclose(nut[J], want, rtol=1e-12, atol=1e-15, err_msg=f"NU_T J={J}")
    # truncated transform of |S| S of the plane, normalised as FFT_XZ_TO_FOURIER
    g = absS * s
    G = np.fft.fft(np.fft.rfft(g, axis=1), axis=0) / (nx * nz)
    nkx, nkz = nx // 3, nz // 3
    cf_own = o.carr(dor.CF1 if axis == "x" else dor.CF3)
    cf_other = o.carr(dor.CF3 if axis == "x" else dor.CF1)
    for J in range(ny + 2):
        got = cf_own[J]
        want = np.zeros_like(got)
        if 2 <= J <= ny - 1:
            if axis == "x":
                for i in range(nkx + 1):
                    want[0, i] = -1j * i * dyfw[J] * G[0, i]
            else:
                for K in range(2 * nkz + 1):
                    kk = K if K <= nkz else K - 2 * nkz - 1
                    kz = kk * kz1
                    want[K, 0] = -1j * kz * dyfw[J] * G[kk % nz, 0]
        sl = (slice(0, 2 * nkz + 1), slice(0, nkx + 1))
        np.testing.assert_allclose(got[sl], want[sl], rtol=0, atol=1e-13, err_msg=f"CF J={J}")
        assert np.abs(cf_other[J][sl]).max() < 1e-14
        assert np.abs(o.carr(dor.CF2)[J][sl]).max() < 1e-14


def test_mean_shear_is_removed(oracle_mod):
    """u1 = c GYF (mean only) has S4 = c/2, all of it mean shear: NU_T = 0 and
    the SGS terms vanish."""
    dor = oracle_mod
    o, p = _oracle(dor, u_bc_ymin=1, u_bc_ymax=1, u_bc_ymin_c1=0.3, u_bc_ymax_c1=0.3)
    gyf = o.vec(dor.GYF).copy()
    _set_phys(o, dor.CU1, 0.3 * gyf[:, None, None] + np.zeros((o.ny + 2, o.nz, o.nx)))
    for f in (dor.CF1, dor.CF2, dor.CF3):
        o.carr(f)[:] = 0.0
    o.les_chan()
    assert np.abs(o.rarr(dor.NUT)).max() < 1e-14
    for f in (dor.CF1, dor.CF2, dor.CF3):
        assert np.abs(o.carr(f)).max() < 1e-14


def _seeded_state(o, dor, seed, nth):
    rng = np.random.default_rng(seed)
    o.create_flow(1, seed, 0.05)
    if nth:
        o.create_th(1)
        th = o.carr(dor.CTH)
        th = th[0] if th.ndim == 4 else th
        th[1:-1, : 2 * (o.nz // 3) + 1, : o.nx // 3 + 1] += 1e-2 * (
            rng.standard_normal(th[1:-1, : 2 * (o.nz // 3) + 1, : o.nx // 3 + 1].shape)
        )
        th[1:-1, 0, 0] = th[1:-1, 0, 0].real


def test_constant_nut_is_raised_viscosity(oracle_mod):
    """Gate closed (new flow, step 1): les_chan does not run and NU_T = c
    only enters the implicit rows and their explicit half; a substep then
    equals one without LES at NU_V_SCALE + c/NU (PR = 1 for the scalar)."""
    dor = oracle_mod
    c, nu = 2e-3, 1e-3
    outs = []
    for les in (1, 0):
        o, p = _oracle(dor, nth=1, les=les, cnf=1, nu=nu, nu_v_scale=1.0 if les else 1.0 + c / nu, pr=[1.0],
                       th_bc_ymin=[0], th_bc_ymax=[0], th_bc_ymin_c1=[0.0], th_bc_ymax_c1=[0.0], ri=[0.0])
        _seeded_state(o, dor, 5, 1)
        if les:
            o.rarr(dor.NUT)[:] = c
        o.rk_chan_1(1)
        o.rk_chan_1(2)
        outs.append([o.carr(f).copy() for f in (dor.CU1, dor.CU2, dor.CU3, dor.CP, dor.CTH)])
    for a, b, name in zip(outs[0], outs[1], ("CU1", "CU2", "CU3", "CP", "CTH")):
        scale = max(np.abs(b).max(), 1e-30)
        assert np.abs(a - b).max() / scale < 1e-11, name


def _flux_profile(o, dor, nut_y, F=3e-3):
    """u(J) with NU_T(J) (u(J) - u(J-1))/DY(J) = F on J = 2..NY."""
    ny, dy = o.ny, o.vec(dor.DY)
    prof = np.zeros(ny + 2)
    prof[1] = 0.2
    for J in range(2, ny + 1):
        prof[J] = prof[J - 1] + dy[J] * F / nut_y[J]
    prof[0], prof[ny + 1] = prof[1], prof[ny]
    return prof


def _nut_profile(o, dor):
    gy = o.vec(dor.GY)
    nut_y = 1e-2 * (1.5 + np.sin(3.0 * gy))  # positive, varying
    nut_y[0] = nut_y[1]
    return nut_y


@pytest.mark.parametrize("comp", ["u1", "th"])
def test_constant_flux_profile_is_fixed_point(oracle_mod, comp):
    """NU = 0, only a mean profile of U1 (or TH) with NU_T (u(J)-u(J-1))/DY(J)
    = F on J = 2..NY and NU_T varying in y: the substep returns the profile
    unchanged."""
    dor = oracle_mod
    probe, _ = _oracle(dor, nu=0.0, cs=1.2)
    nut_y = _nut_profile(probe, dor)
    prof = _flux_profile(probe, dor, nut_y)
    ny = probe.ny
    if comp == "u1":
        o, _ = _oracle(dor, cnf=1, nu=0.0, cs=1.2, u_bc_ymin_c1=prof[1], u_bc_ymax_c1=prof[ny])
        field = lambda: o.carr(dor.CU1)
    else:
        o, _ = _oracle(dor, nth=1, cnf=1, nu=0.0, cs=1.2, pr=[1.0], ri=[0.0], th_bc_ymin=[0], th_bc_ymax=[0],
                       th_bc_ymin_c1=[prof[1]], th_bc_ymax_c1=[prof[ny]])
        field = lambda: o.carr(dor.CTH)[0] if o.carr(dor.CTH).ndim == 4 else o.carr(dor.CTH)
    o.rarr(dor.NUT)[:] = nut_y[:, None, None]
    f = field()
    f[:] = 0.0
    f[:, 0, 0] = prof
    o.rk_chan_1(1)
    got = field()[:, 0, 0].real
    for J in range(1, ny + 1):
        assert abs(got[J] - prof[J]) < 1e-13 * max(1.0, abs(prof[J])), (J, got[J], prof[J])


def test_constant_flux_detects_shifted_pairing(oracle_mod):
    """The fixed point above is not trivial: the same profile moves under a
    NU_T shifted by one row (what a J/J+1 mix-up amounts to)."""
    dor = oracle_mod
    probe, _ = _oracle(dor, nu=0.0, cs=1.2)
    nut_y = _nut_profile(probe, dor)
    prof = _flux_profile(probe, dor, nut_y)
    ny = probe.ny
    o, _ = _oracle(dor, cnf=1, nu=0.0, cs=1.2, u_bc_ymin_c1=prof[1], u_bc_ymax_c1=prof[ny])
    o.rarr(dor.NUT)[:] = np.roll(nut_y, 1)[:, None, None]
    cu = o.carr(dor.CU1)
    cu[:] = 0.0
    cu[:, 0, 0] = prof
    o.rk_chan_1(1)
    got = o.carr(dor.CU1)[:, 0, 0].real
    assert max(abs(got[J] - prof[J]) for J in range(2, ny)) > 1e-6


@pytest.mark.parametrize("cnf,step,runs", [(1, 100, False), (1, 101, True), (0, 1, True)])
def test_spin_up_gate(oracle_mod, cnf, step, runs):
    """channel.f:303: the model runs when the flow is not new or after step 100."""
    dor = oracle_mod
    o, p = _oracle(dor, cnf=cnf)
    _seeded_state(o, dor, 9, 0)
    o.set_clock(0.0, p["delta_t"], step)
    o.rk_chan_1(1)
    assert (np.abs(o.rarr(dor.NUT)).max() > 0) == runs


def test_scalar_eddy_diffusivity(oracle_mod):
    """channel.f:327-375: with u1 = A cos x (Neumann walls, so NU_T = c |sin x|
    on J = 3..NY-1) and a scalar theta = B cos(kz z), ikx theta = 0 and the
    eddy-diffusivity part of CFTH is ikz F[NU_T (-B kz sin(kz z))] on
    JSTART_TH..JEND_TH.  It is the difference of CFTH after a substep with
    the model on (gate open) and off, from the same state: every other
    explicit scalar term is the same in both."""
    dor = oracle_mod
    A, B = 0.7, 0.4
    outs = {}
    for les in (1, 0):
        o, p = _oracle(dor, nth=1, les=les, cnf=0, u_bc_ymin=1, u_bc_ymax=1, w_bc_ymin=1, w_bc_ymax=1,
                       th_bc_ymin=[1], th_bc_ymax=[1], th_bc_ymin_c1=[0.0], th_bc_ymax_c1=[0.0], ri=[0.0], pr=[1.0])
        nx, ny, nz = o.nx, o.ny, o.nz
        x = np.arange(nx) * 2 * math.pi / nx
        z = np.arange(nz) * p["lz"] / nz
        kz1 = 2 * math.pi / p["lz"]
        _set_phys(o, dor.CU1, A * np.broadcast_to(np.cos(x)[None, None, :], (ny + 2, nz, nx)))
        _set_phys(o, dor.CTH, B * np.broadcast_to(np.cos(kz1 * z)[None, :, None], (ny + 2, nz, nx)))
        o.rk_chan_1(1)
        cf = o.carr(dor.CFTH)
        outs[les] = (cf[0] if cf.ndim == 4 else cf).copy()
        if les:
            nut = o.rarr(dor.NUT)[:, :nz, :nx].copy()
    dyw, _ = _widths(o, p, dor)
    for J in range(3, ny):  # the closed form of NU_T the model produced
        np.testing.assert_allclose(nut[J], -0.5 * dyw[J] * math.sqrt(2.0) * A * np.abs(np.sin(x))[None, :]
                                   * np.ones((nz, 1)), rtol=1e-12, atol=1e-15)
    nkx, nkz = nx // 3, nz // 3
    diff = outs[1] - outs[0]
    for J in range(1, ny + 1):  # JSTART_TH = 1 (Neumann), JEND_TH = NY
        flux = nut[J] * (-B * kz1 * np.sin(kz1 * z))[:, None]
        G = np.fft.fft(np.fft.rfft(flux, axis=1), axis=0) / (nx * nz)
        want = np.zeros((2 * nkz + 1, nkx + 1), complex)
        for K in range(2 * nkz + 1):
            kk = K if K <= nkz else K - 2 * nkz - 1
            for i in range(nkx + 1):
                want[K, i] = 1j * kk * kz1 * G[kk % nz, i]
        got = diff[J, : 2 * nkz + 1, : nkx + 1]
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13, err_msg=f"J={J}")
