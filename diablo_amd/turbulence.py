"""Developed channel turbulence on the device, and its statistics.

* :func:`transition_ic` -- the laminar profile of ``CREATE_FLOW_CHAN``
  (``IC_TYPE`` 0, ``set_ics.f:26-36``) plus a seeded perturbation of the low
  x-z modes with a wall-vanishing y profile, made divergence-free by
  ``APPLY_BC_VEL`` and ``REM_DIV_CHAN`` as ``set_ics.f:171-186`` do.  Large
  scales of finite amplitude make the flow transition within a few
  eddy-turnover times; white noise at the grid scale alone decays.
* :func:`run_statistics` -- steps the flow and averages the plane profiles of
  ``SAVE_STATS_CHAN`` (``save_stats.f:136-262``, computed on the device by
  ``dh_save_stats``) over samples taken every few steps.
* :func:`wall_units` -- the friction velocity from the mean wall shear and
  the profiles in wall units, folded over the two halves of the channel.
"""
import math

import numpy as np


def transition_ic(ch, p, gy, seed=1, ubulk=15.7, amp=0.3, kx_max=4, kz_max=6):
    """Laminar Poiseuille of bulk velocity ubulk plus random Fourier modes
    0 < |k| of U1..U3, kx <= kx_max and |kz| <= kz_max, of amplitude amp *
    ubulk * (1 - y^2)^2 (seeded), then the BCs and the projection."""
    ch.create_flow(0, seed, 0.0, 0.0, ubulk)
    st = ch.download()
    gyf = ch_gyf(gy)
    prof = (1.0 - np.clip(gyf, -1.0, 1.0) ** 2) ** 2
    rng = np.random.default_rng(seed)
    nz, tnkz = p["nz"], 2 * (p["nz"] // 3)
    kzs = [k for k in range(-kz_max, kz_max + 1)]
    for key in ("cu1", "cu2", "cu3"):
        a = st[key]
        for kx in range(0, kx_max + 1):
            for kz in kzs:
                if kx == 0 and kz <= 0:
                    continue  # the kx = 0 column is set with its Hermitian partner
                c = (rng.standard_normal() + 1j * rng.standard_normal()) * amp * ubulk / math.sqrt(2.0)
                kzi = kz if kz >= 0 else tnkz + 1 + kz
                a[:, kzi, kx] += c * prof
                if kx == 0:
                    a[:, tnkz + 1 - kz, 0] += np.conj(c) * prof
    ch.upload(st)
    ch.apply_bc_vel()
    ch.rem_div()
    return st


def ch_gyf(gy):
    """GYF(0..NY+1) of a grid GY(1..NY+1) (channel.f:1640-1645)."""
    ny = len(gy) - 1
    gyf = np.zeros(ny + 2)
    gyf[1: ny + 1] = 0.5 * (gy[:-1] + gy[1:])
    gyf[0] = 2 * gyf[1] - gyf[2]
    gyf[ny + 1] = 2 * gyf[ny] - gyf[ny - 1]
    return gyf


def wall_shear(stats, nu, ny):
    """Mean wall shear from DUDY at the two walls (rows 2 and NY: the first
    differences next to GYF(1) and GYF(NY))."""
    return 0.5 * nu * (stats["dudy"][2] - stats["dudy"][ny])


def run_statistics(ch, p, nsteps, every=20, ly=2.0, on_sample=None):
    """Advance nsteps steps, sampling dh_save_stats every `every` steps;
    returns the sample-mean profiles (dict of (NY+2,) arrays, plus 'ubulk',
    'tau_w', 'nsamples') and the per-sample history [(time, tau_w, ubulk)]."""
    acc, hist, n = None, [], 0
    done = 0
    while done < nsteps:
        k = min(every, nsteps - done)
        ch.time_step(k)
        done += k
        s = ch.save_stats(ly)
        t, dt, _, ts = ch.clock()
        tw = wall_shear(s, p["nu"], p["ny"])
        hist.append((t, tw, s["ubulk"]))
        if on_sample is not None:
            on_sample(ts, t, tw, s["ubulk"])
        if acc is None:
            acc = {k2: np.array(v, dtype=np.float64) if np.ndim(v) else float(v) for k2, v in s.items()}
        else:
            for k2, v in s.items():
                acc[k2] = acc[k2] + v
        n += 1
    out = {k2: v / n for k2, v in acc.items()}
    out["tau_w"] = float(np.mean([h[1] for h in hist]))
    out["nsamples"] = n
    return out, hist


def wall_units(stats, nu, gy, utau=None):
    """u_tau (from the mean wall shear unless given), Re_tau and the folded
    profiles U+ (at GYF), urms+, vrms+, wrms+ and -uv+ (as seen from the
    nearer wall) against y+ (distance from the nearer wall) on rows 2..NY-1."""
    ny = len(gy) - 1
    gyf = ch_gyf(gy)
    if utau is None:
        utau = math.sqrt(max(stats["tau_w"], 0.0))
    rows = np.arange(2, ny)
    yw = 1.0 - np.abs(gyf[rows])
    order = np.argsort(yw, kind="stable")
    out = {"u_tau": utau, "re_tau": utau / nu, "y_plus": yw[order] * utau / nu}
    for key, src in (("U_plus", "ume"), ("urms_plus", "urms"), ("vrms_plus", "vrms"), ("wrms_plus", "wrms")):
        out[key] = np.asarray(stats[src])[rows][order] / max(utau, 1e-300)
    # -uv seen from the nearer wall: uv changes sign with y
    out["uv_plus"] = np.sign(gyf[rows][order]) * np.asarray(stats["uv"])[rows][order] / max(utau * utau, 1e-300)
    return out
