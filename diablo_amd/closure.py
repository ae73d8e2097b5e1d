"""The learned eddy-viscosity closure: its network on the device, training
samples from filtered DNS of this solver, training, and the LES runs that
judge it against the constant-Smagorinsky model (no counterpart in the
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
   filtered velocity (dh_les_features), and the target is the eddy viscosity
   that best fits tau's deviatoric part, max(0, -tau^d_ij S_ij / (2 S_kl S_kl)).
3. Adam on the device (dh_mlp_grad_dev, dh_mlp_adam_dev: BF16 MFMA forward
   and backward passes).
4. A-priori error on held-out samples, and LES runs with the trained closure
   and with the Smagorinsky model from the same filtered snapshot, whose
   mean-velocity and rms profiles are compared with the DNS's.
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
        self.G = torch.from_numpy(G).to(dev).to(torch.complex128)
        # y positions: GYF and GY of both grids (rows 0..NY+1; GY(0) taken one spacing below)
        def pos(gy):
            gyf = turbulence.ch_gyf(gy)
            g = np.concatenate([[2 * gy[0] - gy[1]], gy])
            return gyf, g
        fd, yd = pos(gyd)
        fl, yl = pos(gyl)
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

    def _filter(self, four, M):
        """DNS Fourier planes -> box-filtered LES modes, y-interpolated by M."""
        f = four[:, self.kzmap][:, :, : self.shl[2]] * self.G
        return self.t.einsum("ij,jkl->ikl", M, f).contiguous()

    def snapshot(self):
        """(X [6][n] float32, target [n] float32 in units of NU, diagnostics)
        from the DNS context's current state; rows 3..NY-1 of the LES grid."""
        torch = self.t
        pd, pl = self.pd, self.pl
        cu = [self._field(self.dns, c, self.shd) for c in range(3)]
        u = [self._to_phys(self.dns, cu[c], pd) for c in range(3)]
        u2f = u[1].clone()
        u2f[1:-1] = 0.5 * (u[1][1:-1] + u[1][2:])  # V at GYF
        prods = [u[0] * u[0], u2f * u2f, u[2] * u[2], u[0] * u2f, u[0] * u[2], u2f * u[2]]
        pbar = []
        for q in prods:
            fq = self._to_four(self.dns, q, pd, self.shd)
            pbar.append(self._to_phys(self.les, self._filter(fq, self.f2y), pl))
        # the resolved velocity: into the LES context (U1, U3 at GYF, U2 at GY) ...
        ub = [self._filter(cu[0], self.f2f), self._filter(cu[1], self.y2y), self._filter(cu[2], self.f2f)]
        for c in range(3):
            d2d(self.les.field_ptr(c), ub[c].data_ptr(), ub[c].numel() * 16)
        # ... and at GY for the stress
        ug = [self._to_phys(self.les, self._filter(cu[0], self.f2y), pl),
              self._to_phys(self.les, self._filter(cu[1], self.y2y), pl),
              self._to_phys(self.les, self._filter(cu[2], self.f2y), pl)]
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
        nut = torch.clamp(-tdS / (2.0 * SS + 1e-300), min=0.0) / self.nu
        rows = slice(3, nyl)  # J = 3..NY-1
        X = feat[:NIN, rows].reshape(NIN, -1).to(torch.float32).contiguous()
        t = torch.clamp(nut[rows].reshape(-1), max=50.0).to(torch.float32).contiguous()
        smag = (feat[12, rows].reshape(-1) / self.nu).to(torch.float32)  # |S| / NU, for the baseline
        return X, t, smag


def smagorinsky_scale(pl, cs=0.13):
    """C_s^2 Delta^2 per row of the LES grid (les.f:196-227 with DY at GY),
    for the a-priori baseline NU_T = C_s^2 Delta^2 |S|."""
    gy = configs.make_gy(pl["ny"], 2.0, 1.75, 1)
    _, _, dy, _ = channel.grid_metrics(gy)
    dx, dz = pl["lx"] / pl["nx"], pl["lz"] / pl["nz"]
    return cs * cs * (dx * dy * 2.0 * dz) ** (2.0 / 3.0)


# --------------------------------------------------------------- pipeline
def run_pipeline(out_dir, spin=30000, stat=6000, every=20, nsnap=12, iters=3000, batch=16384, les_steps=12000,
                 log=print):
    torch = _torch()
    t_start = time.time()
    rep = {}
    pd, gyd, icd = configs.channel_re180()
    dns = channel.Channel(pd, gyd)
    turbulence.transition_ic(dns, pd, gyd, seed=1)
    dns.set_clock(0.0, pd["delta_t"], 0)
    t0 = time.time()
    dns.time_step(10)
    dns.synchronize()
    t_sim0, t_rate0 = dns.clock()[0], time.time()
    st, hist = turbulence.run_statistics(dns, pd, spin, every=500,
                                         on_sample=lambda ts, t, tw, ub: log(f"[dns] step {ts} t {t:.2f} "
                                                                             f"u_tau {math.sqrt(max(tw, 0)):.3f} "
                                                                             f"U_b {ub:.2f}"))
    dns_rate = (time.time() - t_rate0) / (dns.clock()[0] - t_sim0)  # wall s per unit time, stepping
    rep["dns_spinup"] = {"steps": spin, "wall_s": time.time() - t0, "u_tau_last": math.sqrt(max(hist[-1][1], 0)),
                         "history": [(round(h[0], 3), round(math.sqrt(max(h[1], 0)), 4), round(h[2], 3)) for h in hist]}
    pl, gyl, _ = les_re180()
    les = channel.Channel(pl, gyl)
    les.set_les(True)
    net = Network(seed=0)
    les.set_les_closure(net.params(), ts=pd["nu"], ls=pd["nu"], nus=pd["nu"])
    sampler = Sampler(dns, pd, gyd, les, pl, gyl)
    Xs, ts_, sm = [], [], []
    acc, nacc = None, 0
    t0 = time.time()
    per = max(1, stat // nsnap)
    done = 0
    while done < stat:
        k = min(per, stat - done)
        s, _ = turbulence.run_statistics(dns, pd, k, every=every)
        done += k
        acc = s if acc is None else {q: acc[q] + s[q] for q in s if q != "nsamples"}
        nacc += 1
        X, t, smag = sampler.snapshot()
        Xs.append(X); ts_.append(t); sm.append(smag)
        log(f"[samples] snapshot {len(Xs)}: {t.numel()} samples, target mean {float(t.mean()):.3f} NU")
    dns_wall = time.time() - t0
    dstats = {q: acc[q] / nacc for q in acc}
    # wall units of the forcing: PX0 = -1 on a half height of 1 gives u_tau = 1
    ut = math.sqrt(-pd["px0"])
    wu_d = turbulence.wall_units(dstats, pd["nu"], gyd, utau=ut)
    rep["dns"] = {"u_tau_wall_shear": math.sqrt(max(dstats["tau_w"], 0.0)), "stat_steps": stat, "stat_wall_s": dns_wall}
    # training on all but the last snapshot
    Xtr = torch.cat(Xs[:-1], dim=1).contiguous()
    ttr = torch.cat(ts_[:-1]).contiguous()
    Xva, tva, smva = Xs[-1], ts_[-1], sm[-1]
    n = ttr.numel()
    gen = torch.Generator(device="cuda:0").manual_seed(0)
    t0 = time.time()
    losses = []
    for it in range(iters):
        lr = 1e-3 * (0.1 ** (it / iters))
        idx = torch.randint(0, n, (batch,), device="cuda:0", generator=gen)
        L = net.step(Xtr[:, idx].contiguous(), ttr[idx].contiguous(), lr)
        if it % 250 == 0 or it == iters - 1:
            losses.append((it, float(L)))
            log(f"[train] iter {it} loss {float(L):.4f}")
    torch.cuda.synchronize()
    rep["train"] = {"iters": iters, "batch": batch, "samples": n, "wall_s": time.time() - t0, "loss": losses}
    # a priori on the held-out snapshot
    pred = net.predict(Xva)
    cs2d2 = smagorinsky_scale(pl)
    rows = np.repeat(cs2d2[3: pl["ny"]], pl["nx"] * pl["nz"])
    smag_pred = smva * torch.from_numpy(rows).to("cuda:0").to(torch.float32)
    def err(a):
        return float(torch.linalg.vector_norm(a - tva) / torch.linalg.vector_norm(tva))
    def corr(a):
        a0, b0 = a - a.mean(), tva - tva.mean()
        return float((a0 * b0).sum() / torch.sqrt((a0 * a0).sum() * (b0 * b0).sum()))
    rep["a_priori"] = {"learned": {"rel_l2": err(pred), "corr": corr(pred)},
                       "smagorinsky": {"rel_l2": err(smag_pred), "corr": corr(smag_pred)}}
    log(f"[a priori] {json.dumps(rep['a_priori'])}")
    params = net.params()
    np.savez(os.path.join(out_dir, "closure_trained.npz"), params=params, ts=pd["nu"], ls=pd["nu"], nus=pd["nu"],
             X_val=Xva[:, :8192].cpu().numpy(), t_val=tva[:8192].cpu().numpy(), smag_val=smag_pred[:8192].cpu().numpy())
    # a posteriori: LES from the last filtered snapshot, both models
    res = {}
    ub_state = les.download()
    for model in ("smagorinsky", "learned"):
        ch = channel.Channel(pl, gyl)
        ch.set_les(True)
        if model == "learned":
            ch.set_les_closure(params, ts=pd["nu"], ls=pd["nu"], nus=pd["nu"])
        st0 = dict(ub_state)
        ch.upload(st0)
        ch.apply_bc_vel()
        ch.rem_div()
        ch.set_clock(0.0, pl["delta_t"], 0)
        half = les_steps // 2
        ch.time_step(10)
        ch.synchronize()
        tl0, ts0 = time.time(), ch.clock()[0]
        ch.time_step(half)
        ch.synchronize()
        rate = (time.time() - tl0) / (ch.clock()[0] - ts0)
        s, h = turbulence.run_statistics(ch, pl, les_steps - half, every=every)
        w = turbulence.wall_units(s, pl["nu"], gyl, utau=ut)
        res[model] = {"stats": s, "wall": w, "rate": rate, "t_sim": ch.clock()[0],
                      "u_tau_wall_shear": math.sqrt(max(s["tau_w"], 0.0))}
        ch.close()
    # compare with the DNS at the LES rows (interpolation in y+ of the folded profiles)
    comp = {}
    for model, r in res.items():
        w = r["wall"]
        yp = w["y_plus"]
        e = {}
        for key in ("U_plus", "urms_plus", "vrms_plus", "wrms_plus"):
            ref = np.interp(yp, wu_d["y_plus"], wu_d[key])
            e[key] = float(np.sqrt(np.mean((w[key] - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
        e["u_tau_wall_shear"] = r["u_tau_wall_shear"]
        e["wall_s_per_unit_time"] = r["rate"]
        e["t_sim"] = r["t_sim"]
        comp[model] = e
    rep["a_posteriori"] = comp
    # time to statistics: wall seconds per unit of simulated time while stepping
    rep["time_per_unit_time"] = {"dns_s": dns_rate, "les_learned_s": comp["learned"]["wall_s_per_unit_time"],
                                 "les_smagorinsky_s": comp["smagorinsky"]["wall_s_per_unit_time"],
                                 "speedup_learned_vs_dns": dns_rate / comp["learned"]["wall_s_per_unit_time"]}
    rep["profiles"] = {"dns": {k: np.asarray(v).tolist() for k, v in wu_d.items() if isinstance(v, np.ndarray)},
                       **{m: {k: np.asarray(v).tolist() for k, v in r["wall"].items() if isinstance(v, np.ndarray)}
                          for m, r in res.items()}}
    rep["total_wall_s"] = time.time() - t_start
    les.close()
    dns.close()
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="gpurun_out")
    ap.add_argument("--spin", type=int, default=30000)
    ap.add_argument("--stat", type=int, default=6000)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--les-steps", type=int, default=12000)
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    rep = run_pipeline(a.out, spin=a.spin, stat=a.stat, iters=a.iters, les_steps=a.les_steps,
                       log=lambda s: print(s, flush=True))
    with open(os.path.join(a.out, "closure_report.json"), "w") as f:
        json.dump(rep, f)
    print(json.dumps({k: v for k, v in rep.items() if k != "profiles"}), flush=True)


if __name__ == "__main__":
    main()
