"""The learned eddy-viscosity closure: its network on the device, training
samples from filtered DNS of this solver, training, and the LES runs that
compare it with the constant-Smagorinsky model (no counterpart in the
reference; the network stands in for NU_T = C_s^2 Delta^2 |S| of
``les.f:196-227``, and the strain it reads is ``compute_strain_chan``'s,
``les.f:526-569``).

Pipeline (``python -m diablo_amd.closure``):

1. DNS of the Re_tau = 180 channel (128 x 129 x 128) from a transition IC
   to developed turbulence, then plane statistics and snapshots.
2. Each snapshot, box-filtered in x and z (width 2 Delta of the LES grid)
   and interpolated in y, gives the resolved LES velocity (64 x 65 x 64, same
   domain) and the subgrid stress tau_ij = filter(u_i u_j) - u~_i u~_j; the
   LES context forms the closure's inputs and the resolved strain S from the
   filtered velocity (dh_les_features), and the targets are eddy viscosities
   that match the subgrid dissipation -tau^d_ij S_ij with 2 NU_T S_kl S_kl
   (pointwise or per plane; see VARIANTS).
3. Adam on the device (dh_mlp_grad_dev, dh_mlp_adam_dev: BF16 MFMA forward
   and backward passes), one network per target form.
4. A-priori errors on the held-out last snapshot, and LES runs from that
   filtered snapshot with each network, with the Smagorinsky model and with
   no model, whose mean-velocity and rms profiles are compared with the
   DNS's; the network with the smallest profile error is kept.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import channel, configs, turbulence

NIN = 6


def _torch():
    import torch
    return torch


_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        channel.load_library()
        lib = ctypes.CDLL("libamdhip64.so.7")
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _hip = lib
    return _hip


def d2d(dst, src, nbytes):
    """Device-to-device copy between library buffers and torch tensors."""
    _torch().cuda.synchronize()
    h = _hiprt()
    if h.hipDeviceSynchronize() != 0 or h.hipMemcpy(dst, src, nbytes, 3) != 0:
        raise RuntimeError("hipMemcpy failed")


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"libdiablo_hip: {channel.load_library().dh_mlp_last_error().decode()} ({rc})")


# ---------------------------------------------------------------- network
def init_params(seed=0):
    """He-style initial weights (normal, std sqrt(2 / fan_in)), zero biases,
    output bias 0.5; the parameter layout of dh_mlp_*."""
    H, INPAD = 64, 32
    rng = np.random.default_rng(seed)
    P = np.zeros(channel.load_library().dh_mlp_nparams(), np.float32)
    o = 0
    W1 = np.zeros((INPAD, H))
    W1[:NIN] = rng.standard_normal((NIN, H)) * math.sqrt(2.0 / NIN)
    P[o:o + INPAD * H] = W1.ravel(); o += INPAD * H + H
    P[o:o + H * H] = (rng.standard_normal((H, H)) * math.sqrt(2.0 / H)).ravel(); o += H * H + H
    P[o:o + H * H] = (rng.standard_normal((H, H)) * math.sqrt(2.0 / H)).ravel(); o += H * H + H
    P[o:o + H] = rng.standard_normal(H) * math.sqrt(1.0 / H); o += H
    P[o] = 0.5
    return P


class Network:
    """Parameters, Adam moments and the packed BF16 operands on the device."""

    def __init__(self, params=None, seed=0, device="cuda:0"):
        torch = _torch()
        self.lib = channel.load_library()
        self.dev = device
        P = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.P = torch.from_numpy(P.copy()).to(device)
        self.m = torch.zeros_like(self.P)
        self.v = torch.zeros_like(self.P)
        self.g = torch.zeros_like(self.P)
        self.loss = torch.zeros(1, dtype=torch.float32, device=device)
        self.packed = torch.empty(self.lib.dh_mlp_packed_bytes(), dtype=torch.uint8, device=device)
        self.work = None
        self.step_no = 0
        self.pack()

    def pack(self):
        _check(self.lib.dh_mlp_pack_dev(self.P.data_ptr(), self.packed.data_ptr(), None))

    def _work(self, n):
        torch = _torch()
        need = self.lib.dh_mlp_work_bytes(n)
        if self.work is None or self.work.numel() < need:
            self.work = torch.empty(need, dtype=torch.uint8, device=self.dev)
        return self.work

    def grad(self, X, t):
        """Loss and gradient for X [6][n] float32, t [n] (n a multiple of 32)."""
        n = t.numel()
        w = self._work(n)
        _check(self.lib.dh_mlp_grad_dev(self.P.data_ptr(), self.packed.data_ptr(), X.data_ptr(), t.data_ptr(), n,
                                        w.data_ptr(), w.numel(), self.g.data_ptr(), self.loss.data_ptr(), None))
        return self.loss, self.g

    def step(self, X, t, lr):
        self.grad(X, t)
        self.step_no += 1
        _check(self.lib.dh_mlp_adam_dev(self.P.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.g.data_ptr(),
                                        float(lr), self.step_no, self.packed.data_ptr(), None))
        return self.loss

    def predict(self, X, scale=1.0):
        torch = _torch()
        n = X.shape[1]
        out = torch.empty(n, dtype=torch.float32, device=self.dev)
        _check(self.lib.dh_mlp_infer_dev(self.packed.data_ptr(), self.P.data_ptr(), X.data_ptr(), X.stride(0), n,
                                         float(scale), out.data_ptr(), None))
        return out

    def params(self):
        return self.P.cpu().numpy()


# ----------------------------------------------------------- the samples
def les_re180(nx=64, ny=65, nz=64, dt=2e-3):
    """LES of the config-1 channel (Re_tau = 180, 4 pi x 2 x 4 pi/3, tanh
    grid CS = 1.75, no-slip walls, constant mean pressure gradient) on the
    coarse grid, the model on from the first step."""
    p, gy, ic = configs.channel_re180(nx=nx, ny=ny, nz=nz, dt=dt)
    p.update(les=1, create_new_flow=0)
    return p, configs.make_gy(ny, 2.0, 1.75, 1), ic


def dns_re395(nx=256, ny=193, nz=256):
    """The DNS behind the Re_tau = 395 LES (configs.les_channel's domain,
    2 pi x 2 x pi, tanh grid CS = 1.75, constant mean pressure gradient) on
    the 256 x 193 x 256 grid, no model."""
    p, _, ic = configs.les_channel(nx=nx, ny=ny, nz=nz)
    p.update(les=0)
    return p, configs.make_gy(ny, 2.0, 1.75, 1), ic


def les_re395():
    """The Re_tau = 395 LES: configs.les_channel (64 x 97 x 64), the model
    on from the first step."""
    return configs.les_channel()


def _y_positions(gy):
    """GYF (0:NY+1) and GY with GY(0) taken one spacing below GY(1)."""
    return turbulence.ch_gyf(gy), np.concatenate([[2 * gy[0] - gy[1]], gy])


def refine_state(src, ps, gys, dst, pdst, gyd):
    """The velocity of context src, on a coarser grid of the same domain,
    into context dst: the Fourier modes both grids keep carried over, linear
    interpolation in y (U1 and U3 at GYF, U2 at GY), then APPLY_BC_VEL and
    REM_DIV on dst."""
    torch = _torch()
    dev = "cuda:0"
    nkxs, nkzs = ps["nx"] // 3, ps["nz"] // 3
    nkxd, tnkzd = pdst["nx"] // 3, 2 * (pdst["nz"] // 3)
    shs = (ps["ny"] + 2, 2 * nkzs + 1, nkxs + 1)
    shd = (pdst["ny"] + 2, tnkzd + 1, nkxd + 1)
    kz_s = np.array([k if k <= nkzs else k - 2 * nkzs - 1 for k in range(2 * nkzs + 1)])
    kzmap = torch.from_numpy(np.where(kz_s >= 0, kz_s, tnkzd + 1 + kz_s)).to(dev)
    (fs, ys), (fd, yd) = _y_positions(gys), _y_positions(gyd)
    mat = lambda a, b: torch.from_numpy(_interp_matrix(a, b)).to(dev).to(torch.complex128)
    Mf, My = mat(fs, fd), mat(ys, yd)
    for c in range(3):
        f = torch.empty(shs, dtype=torch.complex128, device=dev)
        d2d(f.data_ptr(), src.field_ptr(c), f.numel() * 16)
        t = torch.einsum("ij,jkl->ikl", My if c == 1 else Mf, f)
        out = torch.zeros(shd, dtype=torch.complex128, device=dev)
        out[:, kzmap, : nkxs + 1] = t
        d2d(dst.field_ptr(c), out.data_ptr(), out.numel() * 16)
    dst.apply_bc_vel()
    dst.rem_div()


def _interp_matrix(src, dst):
    """Linear interpolation in y: rows of dst positions from src positions
    (both increasing), extrapolating at the ends."""
    M = np.zeros((len(dst), len(src)))
    for i, y in enumerate(dst):
        j = int(np.clip(np.searchsorted(src, y) - 1, 0, len(src) - 2))
        w = (y - src[j]) / (src[j + 1] - src[j])
        M[i, j], M[i, j + 1] = 1.0 - w, w
    return M


class Sampler:
    """Training samples of the closure from snapshots of a DNS context,
    through an LES context on the coarse grid (same domain)."""

    def __init__(self, dns, pd, gyd, les, pl, gyl, width=2.0):
        torch = _torch()
        self.t = torch
        self.dns, self.les, self.pd, self.pl = dns, les, pd, pl
        dev = "cuda:0"
        self.dev = dev
        nkxd, tnkzd = pd["nx"] // 3, 2 * (pd["nz"] // 3)
        nkxl, nkzl, tnkzl = pl["nx"] // 3, pl["nz"] // 3, 2 * (pl["nz"] // 3)
        self.shd = (pd["ny"] + 2, tnkzd + 1, nkxd + 1)
        self.shl = (pl["ny"] + 2, tnkzl + 1, nkxl + 1)
        # LES modes among the DNS's (the same wavenumbers: same domain)
        kz_l = np.array([k if k <= nkzl else k - tnkzl - 1 for k in range(tnkzl + 1)])
        self.kzmap = torch.from_numpy(np.where(kz_l >= 0, kz_l, tnkzd + 1 + kz_l)).to(dev)
        kx = np.arange(nkxl + 1) * 2 * math.pi / pd["lx"]
        kz = kz_l * 2 * math.pi / pd["lz"]
        dx, dz = width * pd["lx"] / pl["nx"], width * pd["lz"] / pl["nz"]
        G = np.sinc(kz[:, None] * dz / (2 * math.pi)) * np.sinc(kx[None, :] * dx / (2 * math.pi))  # box filter
        # x-z filters: the box filter of width `width` LES spacings, and the
        # sharp cutoff at the LES grid's retained modes (the implicit filter of
        # a pseudo-spectral LES on that grid)
        self.G = {"box": torch.from_numpy(G).to(dev).to(torch.complex128),
                  "sharp": torch.ones(G.shape, dtype=torch.complex128, device=dev)}
        # y positions: GYF and GY of both grids (rows 0..NY+1)
        fd, yd = _y_positions(gyd)
        fl, yl = _y_positions(gyl)
        m = lambda s, d: torch.from_numpy(_interp_matrix(s, d)).to(dev).to(torch.complex128)
        self.f2f, self.y2y, self.f2y = m(fd, fl), m(yd, yl), m(fd, yl)
        self.nu = pd["nu"]

    def _field(self, ctx, which, shape):
        t = self.t.empty(shape, dtype=self.t.complex128, device=self.dev)
        d2d(t.data_ptr(), ctx.field_ptr(which), t.numel() * 16)
        return t

    def _to_phys(self, ctx, four, p):
        out = self.t.empty((four.shape[0], p["nz"], p["nx"]), dtype=self.t.float64, device=self.dev)
        src = four.clone().contiguous()
        self.t.cuda.synchronize()
        ctx.fft_to_physical_dev(src.data_ptr(), out.data_ptr(), four.shape[0])
        ctx.synchronize()
        return out

    def _to_four(self, ctx, phys, p, shape):
        out = self.t.zeros(shape, dtype=self.t.complex128, device=self.dev)
        src = phys.contiguous()
        self.t.cuda.synchronize()
        ctx.fft_to_fourier_dev(src.data_ptr(), out.data_ptr(), phys.shape[0])
        ctx.synchronize()
        return out

    def _filter(self, four, M, kind):
        """DNS Fourier planes -> filtered LES modes (filter `kind`),
        y-interpolated by M."""
        f = four[:, self.kzmap][:, :, : self.shl[2]] * self.G[kind]
        return self.t.einsum("ij,jkl->ikl", M, f).contiguous()

    def _load(self, cu, kind):
        ub = [self._filter(cu[0], self.f2f, kind), self._filter(cu[1], self.y2y, kind),
              self._filter(cu[2], self.f2f, kind)]
        for c in range(3):
            d2d(self.les.field_ptr(c), ub[c].data_ptr(), ub[c].numel() * 16)

    def load_resolved(self, kind="sharp"):
        """The DNS context's current velocity, filtered by `kind`, into the
        LES context's CU1..3 (U1, U3 at GYF, U2 at GY of the LES grid)."""
        self._load([self._field(self.dns, c, self.shd) for c in range(3)], kind)

    def snapshot(self, kinds=("box", "sharp")):
        """{kind: samples} for each filter kind ("box", "sharp"; the LES
        context is left holding the last kind's resolved velocity): samples of
        the DNS context's current state on rows J = 3..NY-1 of the LES grid,
        flattened row by row (float32 [n], n = rows NZ NX):
        X [6][n] the closure's inputs; the eddy viscosities (units of NU)
        that match the subgrid dissipation D = -tau^d_ij S_ij with
        2 NU_T S_kl S_kl, pointwise (t_point, clipped to [0, 50]; t_signed,
        clipped to [-50, 50]) and per plane (t_plane, the ratio of the plane
        means, broadcast); SS = S_kl S_kl and D themselves; and |S| / NU for
        the Smagorinsky baseline."""
        torch = self.t
        pd, pl = self.pd, self.pl
        cu = [self._field(self.dns, c, self.shd) for c in range(3)]
        u = [self._to_phys(self.dns, cu[c], pd) for c in range(3)]
        u2f = u[1].clone()
        u2f[1:-1] = 0.5 * (u[1][1:-1] + u[1][2:])  # V at GYF
        prods = [u[0] * u[0], u2f * u2f, u[2] * u[2], u[0] * u2f, u[0] * u[2], u2f * u[2]]
        fqs = [self._to_four(self.dns, q, pd, self.shd) for q in prods]
        out = {}
        for kind in kinds:
            out[kind] = self._samples(cu, fqs, kind)
        return out

    def _samples(self, cu, fqs, kind):
        torch = self.t
        pd, pl = self.pd, self.pl
        pbar = [self._to_phys(self.les, self._filter(fq, self.f2y, kind), pl) for fq in fqs]
        # the resolved velocity: into the LES context (U1, U3 at GYF, U2 at GY) ...
        self._load(cu, kind)
        # ... and at GY for the stress
        ug = [self._to_phys(self.les, self._filter(cu[0], self.f2y, kind), pl),
              self._to_phys(self.les, self._filter(cu[1], self.y2y, kind), pl),
              self._to_phys(self.les, self._filter(cu[2], self.f2y, kind), pl)]
        tau = [pbar[0] - ug[0] * ug[0], pbar[1] - ug[1] * ug[1], pbar[2] - ug[2] * ug[2],
               pbar[3] - ug[0] * ug[1], pbar[4] - ug[0] * ug[2], pbar[5] - ug[1] * ug[2]]
        self.les.les_features()
        nyl = pl["ny"]
        feat = torch.empty((13, nyl + 2, pl["nz"], pl["nx"]), dtype=torch.float64, device=self.dev)
        d2d(feat.data_ptr(), self.les.field_ptr(301), feat.numel() * 8)
        S = feat[6:12]  # S11, S22, S33, S12, S13, S23
        tr = (tau[0] + tau[1] + tau[2]) / 3.0
        tdS = (tau[0] - tr) * S[0] + (tau[1] - tr) * S[1] + (tau[2] - tr) * S[2] + 2.0 * (
            tau[3] * S[3] + tau[4] * S[4] + tau[5] * S[5])
        SS = S[0] ** 2 + S[1] ** 2 + S[2] ** 2 + 2.0 * (S[3] ** 2 + S[4] ** 2 + S[5] ** 2)
        D = -tdS
        rows = slice(3, nyl)  # J = 3..NY-1
        nrow = nyl - 3
        f32 = lambda a: a.reshape(-1).to(torch.float32).contiguous()
        ratio = D / (2.0 * SS + 1e-300) / self.nu
        Dm = D[rows].reshape(nrow, -1).mean(dim=1)
        SSm = SS[rows].reshape(nrow, -1).mean(dim=1)
        plane = torch.clamp(Dm / (2.0 * SSm) / self.nu, min=0.0)
        return {"X": feat[:NIN, rows].reshape(NIN, -1).to(torch.float32).contiguous(),
                "t_point": f32(torch.clamp(ratio[rows], min=0.0, max=50.0)),
                "t_signed": f32(torch.clamp(ratio[rows], min=-50.0, max=50.0)),
                "t_plane": f32(plane[:, None].expand(nrow, pl["nz"] * pl["nx"])),
                "SS": f32(SS[rows]), "D": f32(D[rows]),
                "smag": f32(feat[12, rows] / self.nu)}


def smagorinsky_scale(pl, cs=0.13):
    """C_s^2 Delta^2 per row of the LES grid (les.f:196-227 with DY at GY),
    for the a-priori baseline NU_T = C_s^2 Delta^2 |S|."""
    gy = configs.make_gy(pl["ny"], 2.0, 1.75, 1)
    _, _, dy, _ = channel.grid_metrics(gy)
    dx, dz = pl["lx"] / pl["nx"], pl["lz"] / pl["nz"]
    return cs * cs * (dx * dy * 2.0 * dz) ** (2.0 / 3.0)


# ------------------------------------------------------- a-posteriori runs
def les_profiles(state, steps, model="smagorinsky", params=None, scales=None, every=20, utau=1.0, deck=None):
    """LES of the Re_tau = 180 channel on the coarse grid (les_re180) from
    the resolved velocity in `state` (CU1..3 in the host layout): `model`
    "none" (no subgrid model), "smagorinsky" (les.f's constant C_s) or
    "learned" (params and scales = (ts, ls, nus)).  The BCs and the
    projection are applied first; then 10 steps, half the steps timed (wall
    seconds per unit of simulated time), and the other half averaged
    (plane statistics every `every` steps).  The plane kernels' block size
    is tuned first (dh_autotune: the state is restored and the results do
    not depend on it).  Returns the profiles in wall units with the
    forcing's u_tau, and {u_tau from the wall shear, wall seconds per unit
    time, simulated time}.  deck: the LES deck (les_re180 unless given)."""
    pl, gyl, _ = (deck or les_re180)()
    ch = channel.Channel(dict(pl, les=0 if model == "none" else 1), gyl)
    if model == "learned":
        ts, ls, nus = scales
        ch.set_les_closure(params, ts=ts, ls=ls, nus=nus)
    up = ch.download()
    for k in ("cu1", "cu2", "cu3"):
        up[k] = np.asarray(state[k], dtype=np.complex128)
    ch.upload(up)
    ch.apply_bc_vel()
    ch.rem_div()
    ch.set_clock(0.0, pl["delta_t"], 0)
    ch.autotune(2)
    half = steps // 2
    ch.time_step(10)
    ch.synchronize()
    tl0, ts0 = time.time(), ch.clock()[0]
    ch.time_step(half)
    ch.synchronize()
    rate = (time.time() - tl0) / (ch.clock()[0] - ts0)
    s, _ = turbulence.run_statistics(ch, pl, steps - half, every=every)
    w = turbulence.wall_units(s, pl["nu"], gyl, utau=utau)
    out = {"u_tau_wall_shear": math.sqrt(max(s["tau_w"], 0.0)), "wall_s_per_unit_time": rate, "t_sim": ch.clock()[0]}
    ch.close()
    return w, out


def profile_errors(w, dns, filtered):
    """Relative RMS errors of LES profiles `w` against the DNS's (`dns`,
    interpolated to the LES's y+) and against the DNS as the LES grid
    resolves it (`filtered`, on the same rows); score = the mean's error
    against the DNS plus the mean of the three rms errors against the
    filtered DNS."""
    e = {}
    rel = lambda a, ref: float(np.sqrt(np.mean((a - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    for key in PROFILE_KEYS:
        e[key] = rel(w[key], np.interp(w["y_plus"], dns["y_plus"], dns[key]))
        e[key + "_vs_filtered"] = rel(w[key], filtered[key])
    e["score"] = e["U_plus"] + sum(e[k + "_vs_filtered"] for k in PROFILE_KEYS[1:]) / 3.0
    return e


# --------------------------------------------------------------- pipeline
# Target forms, one network each: (name, x-z filter, target, sampling of the
# samples).  The filter is the box filter of two LES spacings or the sharp
# cutoff at the LES grid's modes (Sampler.snapshot).  pointwise: the dissipation-matching NU_T at each point, negative values
# (backscatter) set to 0; signed: the same with its sign kept (a softplus
# output cannot go below 0, so the fit settles on the conditional mean where
# that is positive); dissipation: the signed target with samples drawn with
# probability (S_kl S_kl)^2, which makes the mean-square loss the least-squares
# fit of the dissipation D by 2 NU_T S_kl S_kl; plane: the ratio of the plane
# means of D and 2 S_kl S_kl, the eddy viscosity that gives each plane its
# mean subgrid dissipation.
VARIANTS = (("box_pointwise", "box", "t_point", None), ("box_signed", "box", "t_signed", None),
            ("box_plane", "box", "t_plane", None), ("sharp_signed", "sharp", "t_signed", None),
            ("sharp_dissipation", "sharp", "t_signed", "SS2"), ("sharp_plane", "sharp", "t_plane", None))
FILTERS = ("box", "sharp")
PROFILE_KEYS = ("U_plus", "urms_plus", "vrms_plus", "wrms_plus")

# Cases of the pipeline: the Re_tau = 180 channel (DNS 128 x 129 x 128,
# LES 64 x 65 x 64), and config 3's Re_tau = 395 channel (DNS 256 x 193 x
# 256, spun up on 128 x 129 x 128 and refined; LES 64 x 97 x 64; 200
# snapshots).
CASES = {
    "re180": dict(dns=configs.channel_re180, coarse=None, les=les_re180, ubulk=15.7, spin=30000, adjust=0,
                  stat=6000, nsnap=12, iters=3000),
    "re395": dict(dns=dns_re395, coarse=lambda: dns_re395(128, 129, 128), les=les_re395, ubulk=17.5, spin=20000,
                  adjust=4000, stat=5000, nsnap=200, iters=6000),
}


def run_pipeline(out_dir, case="re180", spin=None, stat=None, every=20, nsnap=None, iters=None, batch=16384,
                 les_steps=12000, adjust=None, log=print):
    torch = _torch()
    C = CASES[case]
    spin = C["spin"] if spin is None else spin
    stat = C["stat"] if stat is None else stat
    nsnap = C["nsnap"] if nsnap is None else nsnap
    iters = C["iters"] if iters is None else iters
    adjust = C["adjust"] if adjust is None else adjust
    t_start = time.time()
    rep = {"case": case}
    pd, gyd, icd = C["dns"]()
    fmt = lambda tag: (lambda ts, t, tw, ub: log(f"[{tag}] step {ts} t {t:.2f} u_tau {math.sqrt(max(tw, 0)):.3f} "
                                                 f"U_b {ub:.2f}"))
    hist_row = lambda h: (round(h[0], 3), round(math.sqrt(max(h[1], 0)), 4), round(h[2], 3))
    t0 = time.time()
    if C["coarse"] is None:
        dns = channel.Channel(pd, gyd)
        turbulence.transition_ic(dns, pd, gyd, seed=1, ubulk=C["ubulk"])
        dns.set_clock(0.0, pd["delta_t"], 0)
        dns.time_step(10)
        dns.synchronize()
        t_sim0, t_rate0 = dns.clock()[0], time.time()
        st, hist = turbulence.run_statistics(dns, pd, spin, every=500, on_sample=fmt("dns"))
        dns_rate = (time.time() - t_rate0) / (dns.clock()[0] - t_sim0)  # wall s per unit time, stepping
        rep["dns_spinup"] = {"steps": spin, "wall_s": time.time() - t0, "u_tau_last": math.sqrt(max(hist[-1][1], 0)),
                             "history": [hist_row(h) for h in hist]}
    else:
        # spin-up on the coarse grid, then the state refined onto the DNS grid
        # and settled there (the DNS's own rate is measured while it settles)
        pc, gyc, _ = C["coarse"]()
        src = channel.Channel(pc, gyc)
        turbulence.transition_ic(src, pc, gyc, seed=1, ubulk=C["ubulk"])
        src.set_clock(0.0, pc["delta_t"], 0)
        _, hist = turbulence.run_statistics(src, pc, spin, every=500, on_sample=fmt("coarse"))
        dns = channel.Channel(pd, gyd)
        refine_state(src, pc, gyc, dns, pd, gyd)
        t_src, dt_src, _, _ = src.clock()
        src.close()
        dns.set_clock(t_src, dt_src / 3.0, 0)  # COURANT sets the fine grid's step from the first step on
        dns.time_step(10)
        dns.synchronize()
        t_sim0, t_rate0 = dns.clock()[0], time.time()
        _, hist2 = turbulence.run_statistics(dns, pd, adjust, every=250, on_sample=fmt("dns"))
        dns_rate = (time.time() - t_rate0) / (dns.clock()[0] - t_sim0)
        rep["dns_spinup"] = {"coarse_grid": [pc["nx"], pc["ny"], pc["nz"]], "coarse_steps": spin, "fine_steps": adjust,
                             "wall_s": time.time() - t0, "u_tau_last": math.sqrt(max(hist2[-1][1], 0)),
                             "history": [hist_row(h) for h in hist], "history_fine": [hist_row(h) for h in hist2]}
    pl, gyl, _ = C["les"]()
    les = channel.Channel(pl, gyl)
    les.set_les(True)
    les.set_les_closure(init_params(0), ts=pd["nu"], ls=pd["nu"], nus=pd["nu"])  # its inputs need it on
    sampler = Sampler(dns, pd, gyd, les, pl, gyl)
    snaps = {kind: [] for kind in FILTERS}
    acc, nacc = None, 0
    fsamples = []

    def filtered_sample(ts, t, tw, ub):
        # plane statistics of the DNS velocity as the LES grid resolves it
        # (sharp cutoff at its modes, y-interpolated to its rows)
        sampler.load_resolved("sharp")
        fsamples.append(les.save_stats(2.0))

    t0 = time.time()
    t_stat0 = dns.clock()[0]
    per = max(1, stat // nsnap)
    done = 0
    while done < stat:
        k = min(per, stat - done)
        s, _ = turbulence.run_statistics(dns, pd, k, every=every, on_sample=filtered_sample)
        done += k
        acc = s if acc is None else {q: acc[q] + s[q] for q in s if q != "nsamples"}
        nacc += 1
        sn_all = sampler.snapshot(FILTERS)  # the sharp cutoff last: its resolved field stays in the LES context
        for kind in FILTERS:
            snaps[kind].append(sn_all[kind])
            sn = snaps[kind][-1]
            log(f"[samples] {kind} snapshot {len(snaps[kind])}: {sn['t_point'].numel()} samples, pointwise target "
                f"mean {float(sn['t_point'].mean()):.3f} NU, plane target mean {float(sn['t_plane'].mean()):.3f} NU")
    dns_wall = time.time() - t0
    dstats = {q: acc[q] / nacc for q in acc}
    # wall units of the forcing: PX0 = -1 on a half height of 1 gives u_tau = 1
    ut = math.sqrt(-pd["px0"])
    wu_d = turbulence.wall_units(dstats, pd["nu"], gyd, utau=ut)
    fstats = {q: np.mean([f[q] for f in fsamples], axis=0) for q in fsamples[0]}
    fstats["tau_w"] = dstats["tau_w"]
    wu_f = turbulence.wall_units(fstats, pl["nu"], gyl, utau=ut)  # the LES grid's rows
    rep["dns"] = {"u_tau_wall_shear": math.sqrt(max(dstats["tau_w"], 0.0)), "stat_steps": stat, "stat_wall_s": dns_wall,
                  "dt_developed": (dns.clock()[0] - t_stat0) / stat}
    # training on all but the last snapshot, one network per target form
    cat = lambda kind, key: torch.cat([s[key] for s in snaps[kind][:-1]], dim=-1).contiguous()
    vals = {kind: snaps[kind][-1] for kind in FILTERS}
    n = vals["box"]["X"].shape[1] * (len(snaps["box"]) - 1)
    nrow = pl["ny"] - 3
    gen = torch.Generator(device="cuda:0").manual_seed(0)
    nets = {}
    rep["train"] = {}
    chunk = 250
    for name, kind, tkey, weight in VARIANTS:
        Xtr, ttr = cat(kind, "X"), cat(kind, tkey)
        w = None
        if weight == "SS2":
            w = cat(kind, "SS").double() ** 2
            w = torch.clamp(w, max=float(torch.quantile(w[::31].float(), 0.999))).float()
        net = Network(seed=0)
        t0 = time.time()
        losses = []
        if w is not None:  # inverse-CDF draws (more samples than torch.multinomial takes categories)
            cdf = torch.cumsum(w.double(), 0)
        for it in range(iters):
            if it % chunk == 0 and w is not None:
                u = torch.rand(chunk * batch, dtype=torch.float64, device="cuda:0", generator=gen) * cdf[-1]
                draws = torch.clamp(torch.searchsorted(cdf, u), max=n - 1)
            lr = 1e-3 * (0.1 ** (it / iters))
            if w is None:
                idx = torch.randint(0, n, (batch,), device="cuda:0", generator=gen)
            else:
                idx = draws[(it % chunk) * batch:(it % chunk + 1) * batch]
            L = net.step(Xtr[:, idx].contiguous(), ttr[idx].contiguous(), lr)
            if it % chunk == 0 or it == iters - 1:
                losses.append((it, float(L)))
        torch.cuda.synchronize()
        nets[name] = net
        rep["train"][name] = {"filter": kind, "target": tkey, "sampling": weight or "uniform", "iters": iters, "batch": batch,
                              "samples": n, "wall_s": time.time() - t0, "loss": losses}
        log(f"[train] {name}: loss {losses[0][1]:.4f} -> {losses[-1][1]:.4f}")

    # a priori on the held-out snapshot: the eddy viscosity against the
    # pointwise target, and the subgrid dissipation 2 NU_T S_kl S_kl against
    # D pointwise and as plane means
    def corr(a, b):
        a0, b0 = a - a.mean(), b - b.mean()
        return float((a0 * b0).sum() / torch.sqrt((a0 * a0).sum() * (b0 * b0).sum()))

    def apriori(pred, val):
        Dp = 2.0 * pred * pd["nu"] * val["SS"]
        pm = Dp.view(nrow, -1).mean(dim=1)
        tm = val["D"].view(nrow, -1).mean(dim=1)
        return {"nut_corr": corr(pred, val["t_point"]), "diss_corr": corr(Dp, val["D"]),
                "plane_diss_rel_l2": float(torch.linalg.vector_norm(pm - tm) / torch.linalg.vector_norm(tm))}

    cs2d2 = smagorinsky_scale(pl)
    rows = np.repeat(cs2d2[3: pl["ny"]], pl["nx"] * pl["nz"])
    rep["a_priori"] = {}
    for kind in FILTERS:
        val = vals[kind]
        smag_pred = val["smag"] * torch.from_numpy(rows).to("cuda:0").to(torch.float32)
        rep["a_priori"][f"smagorinsky_{kind}"] = apriori(smag_pred, val)
        rep["a_priori"][f"dns_{kind}_mean_dissipation"] = float(val["D"].mean())
    for name, kind, _, _ in VARIANTS:
        rep["a_priori"][name] = apriori(nets[name].predict(vals[kind]["X"]), vals[kind])
    log(f"[a priori] {json.dumps(rep['a_priori'])}")

    # a posteriori: LES from the last (sharp-cutoff) filtered snapshot with
    # each model, and without one, against the DNS profiles (interpolated in y+)
    ub_state = les.download()

    def les_run(model):
        params = nets[model].params() if model in nets else None
        w, e = les_profiles(ub_state, les_steps, model="learned" if params is not None else model, params=params,
                            scales=(pd["nu"], pd["nu"], pd["nu"]), every=every, utau=ut, deck=C["les"])
        e.update(profile_errors(w, wu_d, wu_f))
        log(f"[a posteriori] {model}: {json.dumps(e)}")
        return e, w

    comp, walls = {}, {}
    for model in ("none", "smagorinsky") + tuple(nets):
        comp[model], walls[model] = les_run(model)
    rep["a_posteriori"] = comp
    chosen = min(nets, key=lambda m: comp[m]["score"])
    rep["chosen"] = chosen
    # time to statistics: wall seconds per unit of simulated time while stepping
    rep["time_per_unit_time"] = {"dns_s": dns_rate, "les_learned_s": comp[chosen]["wall_s_per_unit_time"],
                                 "les_smagorinsky_s": comp["smagorinsky"]["wall_s_per_unit_time"],
                                 "speedup_learned_vs_dns": dns_rate / comp[chosen]["wall_s_per_unit_time"]}
    rep["profiles"] = {"dns": {k: np.asarray(v).tolist() for k, v in wu_d.items() if isinstance(v, np.ndarray)},
                       "dns_filtered": {k: np.asarray(v).tolist() for k, v in wu_f.items()
                                        if isinstance(v, np.ndarray)},
                       **{m: {k: np.asarray(v).tolist() for k, v in w.items() if isinstance(v, np.ndarray)}
                          for m, w in walls.items()}}
    np.savez(os.path.join(out_dir, "closure_trained.npz"), params=nets[chosen].params(), variant=chosen,
             ts=pd["nu"], ls=pd["nu"], nus=pd["nu"],
             **{f"params_{m}": net.params() for m, net in nets.items()})
    # the a-posteriori case as a fixture: the filtered initial state, and the
    # DNS profiles in wall units, whole and as the LES grid resolves them
    np.savez_compressed(os.path.join(out_dir, "les_case.npz"),
                        **{k: ub_state[k].astype(np.complex64) for k in ("cu1", "cu2", "cu3")},
                        **{f"dns_{k}": np.asarray(wu_d[k]) for k in ("y_plus",) + PROFILE_KEYS},
                        **{f"filtered_{k}": np.asarray(wu_f[k]) for k in ("y_plus",) + PROFILE_KEYS})
    rep["total_wall_s"] = time.time() - t_start
    les.close()
    dns.close()
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="closure_out")
    ap.add_argument("--case", default="re180", choices=sorted(CASES))
    ap.add_argument("--spin", type=int, default=None)
    ap.add_argument("--adjust", type=int, default=None)
    ap.add_argument("--stat", type=int, default=None)
    ap.add_argument("--nsnap", type=int, default=None)
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--les-steps", type=int, default=12000)
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    rep = run_pipeline(a.out, case=a.case, spin=a.spin, stat=a.stat, nsnap=a.nsnap, iters=a.iters,
                       les_steps=a.les_steps, adjust=a.adjust, log=lambda s: print(s, flush=True))
    with open(os.path.join(a.out, "closure_report.json"), "w") as f:
        json.dump(rep, f)
    print(json.dumps({k: v for k, v in rep.items() if k != "profiles"}), flush=True)


if __name__ == "__main__":
    main()
