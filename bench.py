"""Throughput of the channel timestep on one MI355X.

Workload: turbulent channel at Re_tau=180, 128 x 129 x 128, domain
4*pi x 2 x 4*pi/3, tanh y-grid (CS=1.75), no-slip walls, constant mean
pressure gradient, FP64, CFL=0.5 variable dt; seeded laminar-plus-noise
initial condition built on the device.  One step = the three RK_CHAN_1
substeps of diablo.f:70-84 plus TIME+=DELTA_T.

Prints one JSON line: Mpts*RK3-steps/s = NX*NY*NZ*steps/wall/1e6, the HBM
roofline of the dominant kernel (algorithmic bytes / mean HIP-event
duration), the roofline of the whole step against the algorithmic bytes
per step, and the CPU baseline (the oracle, OpenMP on the host cores);
beside it the mixed form, the LES deck, the learned-LES time to
statistics, the batched solver and 512^2 transforms, and the 64-instance
shear-layer ensemble (each can be left out with its --no-* flag).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts·RK3-steps/s, Re_tau=180 channel; learned-LES speedup at matched U+ stats"
UNIT = "Mpts·RK3-steps/s"
HBM_PEAK_GBS = 8000.0


def plane_bytes(p):
    nkx, nkz = p["nx"] // 3, p["nz"] // 3
    F = 16 * (nkx + 1) * (2 * nkz + 1)  # one de-aliased complex plane
    P = 8 * p["nx"] * p["nz"]            # one physical plane
    return F, P


def step_bytes(p):
    """Algorithmic bytes per RK3 step (FP64, LES off): (66+11 N_TH) F +
    (36+12 N_TH) P with F, P whole fields over NY+2 planes."""
    Fp, Pp = plane_bytes(p)
    nyy, nth = p["ny"] + 2, p["n_th"]
    return (66 + 11 * nth) * Fp * nyy + (36 + 12 * nth) * Pp * nyy


def phase_bytes(p, rk):
    """Algorithmic (unique) bytes each kernel of substep rk must move, for
    the velocity fields (N_TH=0 workload): planes actually read/written."""
    Fp, Pp = plane_bytes(p)
    NY = p["ny"]
    nyy = NY + 2
    nf = 3 + p["n_th"]
    b = {}
    b["to_physical"] = nf * nyy * (Fp + Pp)
    cf_old = 3 * (NY - 1) * Fp if rk > 1 else 0
    b["nonlinear"] = 3 * (NY + 1) * Pp + 3 * (NY - 1) * Fp + NY * Fp + cf_old + 3 * (NY - 1) * Fp + 3 * (NY - 1) * Pp
    b["cn_solve"] = 3 * (NY - 1) * Pp + 3 * (NY + 1) * Pp  # R in, the solution written over it
    b["courant"] = 0
    b["to_fourier"] = nf * nyy * (Fp + Pp)
    b["rem_div"] = 3 * NY * Fp + 3 * NY * Fp + 2 * (NY - 1) * Fp  # velocity in and out, CP updated
    return b


def measured_traffic():
    """Device-memory bytes per launch by phase, from the FETCH_SIZE and
    WRITE_SIZE counter runs summarised in profiles/ (scaled by the reading
    of a plain copy of known size); {} when the file is absent."""
    path = os.path.join(ROOT, "profiles", "traffic_per_launch.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) == 2 and not line.startswith("#"):
                out[parts[0]] = int(parts[1])
    return out


def cpu_baseline(p, gy, ic, budget_s=12.0, max_steps=60):
    from oracle import dor
    from tests import helpers
    o, *_ = helpers.make_oracle(dor, "channel_re180")
    o.time_step(1)  # first step outside the sample (page-in, thread start)
    n, t0 = 0, time.perf_counter()
    while n < max_steps:
        o.time_step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    pts = p["nx"] * p["ny"] * p["nz"]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": pts * n / dt / 1e6, "unit": UNIT, "cores": cores, "kind": "port",
            "sample": f"{n} RK3 steps of the same 128x129x128 channel from the same seeded IC "
                      f"(after 1 untimed step), FP64 C restatement with OpenMP over planes/columns, {dt:.1f} s"}


def copy_calibration(mib, reps=5):
    """Plain device-to-device copies of `mib` MiB (HIP's blit kernel): the
    reference reading for the FETCH_SIZE / WRITE_SIZE counters, which count
    L2-to-fabric traffic, and the achievable copy bandwidth of the device."""
    import ctypes
    from diablo_amd import channel
    channel.load_library()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    n = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for b in (src, dst):
        if hip.hipMalloc(ctypes.byref(b), n) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, n)
    hip.hipDeviceSynchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        if hip.hipMemcpy(dst, src, n, 3) != 0:
            raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        times.append(time.perf_counter() - t0)
    for b in (src, dst):
        hip.hipFree(b)
    best = min(times)
    print(json.dumps({"copy_bytes": n, "read_plus_write_GBs": round(2 * n / best / 1e9, 1),
                      "best_ms": round(best * 1e3, 4)}), flush=True)


def ensemble_run(n, steps, warmup, mixed=False, one_at_a_time=True):
    """Uneven shear-layer instances (NX in {64..256}, NY = NX + 1, NZ = 4)
    stepped as one graph (FP64, or the FP32/FP64 mixed form); aggregate
    sum(NX NY NZ) steps / wall, next to the same instances run one after
    another (each with its own step graph)."""
    from diablo_amd import channel, configs
    decks = configs.kh_ensemble(n)
    chs = []
    for p, gy, ic in decks:
        ch = channel.Channel(p, gy)
        ch.initial_condition(ic)
        ch.set_clock(0.0, p["delta_t"], 0)
        ch.set_mixed_precision(mixed)
        chs.append(ch)
    pts = sum(p["nx"] * p["ny"] * p["nz"] for p, _, _ in decks)
    ens = channel.Ensemble(chs)
    # block size per member shape, tuned on the members' own state
    tuned = ens.autotune(2)
    block = {}
    for (p, _, _), t in zip(decks, tuned):
        block[p["nx"]] = t
    ens.time_step(warmup)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.time_step(steps)
    ens.synchronize()
    wall = time.perf_counter() - t0
    sizes = {}
    for p, _, _ in decks:
        sizes[p["nx"]] = sizes.get(p["nx"], 0) + 1
    out = {"value": round(pts * steps / wall / 1e6, 3), "unit": UNIT, "steps": steps, "warmup": warmup,
           "ms_per_step": round(wall / steps * 1e3, 4),
           "dtype": "f32/f64 mixed" if mixed else "f64",
           "config": {"workload": "kh_ensemble", "instances": n, "nx_counts": sizes, "nz": 4,
                      "points_per_step": pts, "block_size_by_nx": block}}
    if one_at_a_time:
        for ch in chs:
            ch.time_step(1)
            ch.synchronize()
        t1 = time.perf_counter()
        for ch in chs:
            ch.time_step(steps)
            ch.synchronize()
        wall_seq = time.perf_counter() - t1
        out["one_at_a_time"] = {"value": round(pts * steps / wall_seq / 1e6, 3),
                                "ms_per_step": round(wall_seq / steps * 1e3, 4)}
    ens.close()
    for ch in chs:
        ch.close()
    return out


def ensemble_bench(n, steps, warmup):
    out = ensemble_run(n, steps, warmup)
    out.update({"metric": f"Mpts·RK3-steps/s, KH shear-layer ensemble of {n} uneven instances", "n_gpus": 1,
                "higher_is_better": True, "data": "synthetic (seeded KH tanh shear layer + noise per instance)"})
    print(json.dumps(out), flush=True)


def mixed_form(p, gy, ic, steps, warmup):
    """The FP32/FP64 mixed form on the same workload: its throughput, and the
    relative L2 difference of each state field from the FP64 path after 10
    steps from the same state (the FP64 path is itself at the roundoff floor
    of the oracle)."""
    import numpy as np
    from diablo_amd import channel

    ch = channel.Channel(p, gy)
    ch.initial_condition(ic)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.set_mixed_precision(True)
    ch.time_step(warmup)
    ch.synchronize()
    t0 = time.perf_counter()
    ch.time_step(steps)
    ch.synchronize()
    wall = time.perf_counter() - t0
    ch.close()
    pts = p["nx"] * p["ny"] * p["nz"]
    runs = []
    for mixed in (False, True):
        c = channel.Channel(p, gy)
        c.initial_condition(ic)
        c.set_clock(0.0, p["delta_t"], 0)
        c.set_mixed_precision(mixed)
        c.time_step(10)
        runs.append(c.download())
        c.close()
    err = {}
    for k in ("cu1", "cu2", "cu3", "cp"):
        a, b = runs[1][k], runs[0][k]
        err[k] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"value": round(pts * steps / wall / 1e6, 3), "unit": UNIT, "ms_per_step": round(wall / steps * 1e3, 4),
            "dtype": "f32 transforms of increments and fluctuations, f64 state, means and solves",
            "rel_l2_vs_f64_after_10_steps": {k: float(f"{v:.3g}") for k, v in err.items()}}


def les_cpu_baseline(budget_s=4.0, max_steps=20):
    """The oracle's LES channel step (the same deck, model on) timed on the
    host: a bounded sample of a few steps."""
    from oracle import dor
    from tests import helpers
    o, p, *_ = helpers.make_oracle(dor, "les_channel")
    o.time_step(1)  # first step outside the sample
    n, t0 = 0, time.perf_counter()
    while n < max_steps:
        o.time_step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    pts = p["nx"] * p["ny"] * p["nz"]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": pts * n / dt / 1e6, "unit": UNIT, "cores": cores, "kind": "port",
            "sample": f"{n} RK3 steps of the 64x97x64 LES channel (after 1 untimed step), FP64 C restatement "
                      f"of les_chan and the substep with OpenMP over planes/columns, {dt:.1f} s"}


def les_form(steps, warmup, reps=10):
    """The constant-Smagorinsky LES channel (Re_tau = 395, 64 x 97 x 64, the
    model on from the first step): its throughput, the same deck stepped
    without the model and with the learned closure in place of the model's
    NU_T, and the LES substep's phases (the model runs in to_physical, its
    implicit terms in cn_solve).  Each context takes the block size
    dh_autotune finds for it first."""
    from diablo_amd import channel, configs

    import numpy as np

    p, gy, ic = configs.les_channel()
    pts = p["nx"] * p["ny"] * p["nz"]
    # the learned closure: the network trained on filtered DNS when its
    # parameters ship with the package, else one of random weights (the cost
    # does not depend on the values)
    from diablo_amd import closure
    trained = os.path.join(ROOT, "diablo_amd", "data", "closure_re180.npz")
    if os.path.exists(trained):
        z = np.load(trained)
        w, wsrc = z["params"], "trained on filtered Re_tau=180 DNS"
        scales = dict(ts=float(z["ts"]), ls=float(z["ls"]), nus=float(z["nus"]))
    else:
        w, wsrc = closure.init_params(0), "random weights"
        scales = dict(ts=p["nu"], ls=p["nu"], nus=p["nu"])
    res = {}
    for les in (True, False, "closure"):
        ch = channel.Channel(p, gy)
        ch.set_les(bool(les))
        if les == "closure":
            ch.set_les_closure(w, **scales)
        ch.initial_condition(ic)
        ch.set_clock(0.0, p["delta_t"], 0)
        threads = ch.autotune(2)  # block size of the 64x64-plane kernels, state restored
        ch.time_step(warmup)
        ch.synchronize()
        t0 = time.perf_counter()
        ch.time_step(steps)
        ch.synchronize()
        wall = time.perf_counter() - t0
        res[les] = (pts * steps / wall / 1e6, wall / steps * 1e3, threads)
        if les is True:
            phases = {f"rk{rk}:{k}": round(v, 5) for rk in (1, 2, 3) for k, v in ch.profile_substep(rk, reps).items()}
        ch.close()
    return {"workload": "les_channel: constant-Smagorinsky LES, Re_tau=395, 64x97x64, 2pi x 2 x pi, "
                        "tanh grid CS=1.75, FP64, CFL=0.5 variable dt",
            "value": round(res[True][0], 3), "unit": UNIT, "ms_per_step": round(res[True][1], 4),
            "block_size": res[True][2],
            "without_model": {"value": round(res[False][0], 3), "ms_per_step": round(res[False][1], 4)},
            "learned_closure": {"value": round(res["closure"][0], 3), "ms_per_step": round(res["closure"][1], 4),
                                "network": f"6 -> 64 -> 64 -> 64 -> 1 GELU/softplus, BF16 MFMA 16x16x32, {wsrc}"},
            "phases_ms": phases}


def large_plane_dns_ms(steps, warmup=3):
    """Milliseconds per step of config 3's DNS (Re_tau = 395, 256 x 193 x 256,
    FP64), whose 528 KB x-z planes go through the large-plane kernels: from
    the seeded transitional IC, `warmup` steps untimed, then `steps` timed."""
    from diablo_amd import channel, closure, turbulence
    pd, gyd, _ = closure.dns_re395()
    dns = channel.Channel(pd, gyd)
    turbulence.transition_ic(dns, pd, gyd, seed=1, ubulk=17.5)
    dns.set_clock(0.0, 1e-5, 0)
    dns.time_step(warmup)
    dns.synchronize()
    t0 = time.perf_counter()
    dns.time_step(steps)
    dns.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    dns.close()
    return ms


def learned_les(dns_ms_per_step, steps=1200):
    """The second half of the headline metric: how much sooner an LES with
    the trained closure reaches statistics than the DNS, at the statistics
    the closure pipeline matched (python -m diablo_amd.closure; profile
    errors quoted from profiles/closure_re*.json).  Live: each LES's wall
    seconds per unit of simulated time, from the filtered DNS state of its
    case (tests/golden/closure_les_case*.npz), with its own CFL time steps in
    developed turbulence (steps / 2 timed), for the trained network, the
    Smagorinsky model and no model; and the DNS's, as its ms per step
    (timed here) over the DNS time step of developed turbulence the pipeline
    recorded.  Cases: config 3 (Re_tau = 395, DNS 256 x 193 x 256 timed live
    for 10 steps, LES 64 x 97 x 64) and the Re_tau = 180 channel (DNS = this
    bench's config-1 step, LES 64 x 65 x 64)."""
    import numpy as np
    from diablo_amd import closure
    out = {}
    for case, deck, case_file, net_file in (("re395", closure.les_re395, "closure_les_case_re395.npz", "closure_re395.npz"),
                                            ("re180", closure.les_re180, "closure_les_case.npz", "closure_re180.npz")):
        rep = json.load(open(os.path.join(ROOT, "profiles", f"closure_{case}.json")))
        z = np.load(os.path.join(ROOT, "tests", "golden", case_file))
        net = np.load(os.path.join(ROOT, "diablo_amd", "data", net_file))
        live = {}
        for model in ("learned", "smagorinsky", "none"):
            _, info = closure.les_profiles(z, steps, model=model, params=net["params"], deck=deck,
                                           scales=(float(net["ts"]), float(net["ls"]), float(net["nus"])))
            live[model] = round(info["wall_s_per_unit_time"], 5)
        if case == "re395":
            ms = large_plane_dns_ms(10)
        else:
            ms = dns_ms_per_step
        dns_s = ms * 1e-3 / rep["dns"]["dt_developed"]
        ap = rep["a_posteriori"]
        keep = ("U_plus", "urms_plus_vs_filtered", "vrms_plus_vs_filtered", "wrms_plus_vs_filtered",
                "urms_plus", "vrms_plus", "wrms_plus", "score")
        out[case] = {"workload": rep["case"],
                     "speedup_learned_vs_dns": round(dns_s / live["learned"], 3),
                     "dns_ms_per_step": round(ms, 3), "dns_dt_developed": rep["dns"]["dt_developed"],
                     "wall_s_per_unit_time": {"dns": round(dns_s, 5), **{f"les_{m}": v for m, v in live.items()}},
                     "profile_errors": {m: {k: round(ap[m][k], 4) for k in keep} for m in (rep["chosen"], "smagorinsky", "none")},
                     "network": f"6->64->64->64->1 GELU/softplus on BF16 MFMA, target form {rep['chosen']}",
                     "source_of_errors": f"profiles/closure_{case}.json"}
    return out


def _event_ms(fn, reps):
    """Mean milliseconds of fn() over reps launches on the current (null)
    stream, between HIP events recorded on that stream."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def batched_microbench(reps=20):
    """Batched solver and transform microbenchmark: 257 x 512 complex
    Crank-Nicolson systems of NY = 257 rows on the tanh grid (CS = 1.75),
    alpha log-uniform in [1e-3, 10] (seed 2), complex Gaussian right-hand
    sides (seeded on the device), FP64 and FP32 with the residual in FP64;
    and 257 planes of 512 x 512 through the x/z transforms both ways, with
    the error of 2 planes against numpy.fft.  Algorithmic bytes: 32 B per
    row (FP64 solve: right-hand side in, solution out; 16 B in FP32); per
    transform direction 8 B per real point + 16 B per complex mode."""
    import numpy as np
    import torch
    from diablo_amd import batched

    dev = "cuda:0"
    ny, nkx, nkz = 257, 257, 512
    nsys = nkx * nkz
    sl, su = batched.cn_rows(ny)
    alpha = np.exp(np.random.default_rng(2).uniform(np.log(1e-3), np.log(10.0), nsys))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sl_d, su_d, al_d = d(sl), d(su), d(alpha)
    g = torch.Generator(device=dev).manual_seed(3)
    b = torch.complex(torch.randn((ny, nsys), generator=g, device=dev, dtype=torch.float64),
                      torch.randn((ny, nsys), generator=g, device=dev, dtype=torch.float64))
    x = torch.empty_like(b)
    out = {"workload": "batched: 257x512 complex CN systems of NY=257 (tanh grid CS=1.75, alpha log-uniform "
                       "[1e-3,10]) + 257 planes of 512x512 x/z transforms", "rows": ny * nsys}
    res = {}
    for fp32 in (False, True):
        bb = b.to(torch.complex64) if fp32 else b
        xx = torch.empty_like(bb)
        ms = _event_ms(lambda: batched.cn_solve(sl_d, su_d, al_d, bb, xx, fp32=fp32), reps)
        xr = xx.to(torch.complex128)
        r = batched.cn_matrix_apply(sl_d, su_d, al_d, xr) - b
        resid = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b))
        byts = (16 if fp32 else 32) * ny * nsys
        gbs = byts / (ms * 1e-3) / 1e9
        res["fp32" if fp32 else "fp64"] = {"ms": round(ms, 4), "GBs": round(gbs, 1),
                                           "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": byts,
                                           "residual_fp64": float(f"{resid:.3g}")}
        del xx, xr, r
    out["cn_solve"] = res
    del b, x
    torch.cuda.empty_cache()
    n = 257
    rng = np.random.default_rng(4)
    sample = rng.standard_normal((2, 512, 512))
    gp = torch.Generator(device=dev).manual_seed(4)
    fft = {}
    for fp32 in (False, True):
        rdt, cdt = (torch.float32, torch.complex64) if fp32 else (torch.float64, torch.complex128)
        phys = torch.randn((n, 512, 512), generator=gp, device=dev, dtype=torch.float64).to(rdt)
        phys[:2] = torch.from_numpy(sample).to(dev).to(rdt)
        four = torch.empty((n, 512, 257), dtype=cdt, device=dev)
        back = torch.empty_like(phys)
        ms_f = _event_ms(lambda: batched.fft_to_fourier(phys, four, fp32=fp32), reps)
        ms_i = _event_ms(lambda: batched.fft_to_physical(four, back, fp32=fp32), reps)
        ref = np.fft.fft(np.fft.rfft(sample, axis=-1) / 512, axis=-2) / 512
        f2 = four[:2].cpu().numpy().astype(np.complex128)
        err = float(np.linalg.norm(f2 - ref) / np.linalg.norm(ref))
        rt = float(torch.linalg.vector_norm((back - phys).double()) / torch.linalg.vector_norm(phys.double()))
        esz = 4 if fp32 else 8
        byts = n * (esz * 512 * 512 + 2 * esz * 512 * 257)
        fft["fp32" if fp32 else "fp64"] = {
            "forward_ms": round(ms_f, 4), "inverse_ms": round(ms_i, 4), "bytes_per_direction": byts,
            "forward_GBs": round(byts / (ms_f * 1e-3) / 1e9, 1), "inverse_GBs": round(byts / (ms_i * 1e-3) / 1e9, 1),
            "forward_frac": round(byts / (ms_f * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "inverse_frac": round(byts / (ms_i * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "rel_l2_vs_numpy_2_planes": float(f"{err:.3g}"), "round_trip_rel_l2": float(f"{rt:.3g}")}
        del phys, four, back
        torch.cuda.empty_cache()
    out["fft512"] = fft
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="leave out the FP32/FP64 mixed form's line")
    ap.add_argument("--no-les", action="store_true", help="leave out the LES channel's line")
    ap.add_argument("--no-batched", action="store_true", help="leave out the batched solver/transform line")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="leave out the 64-instance KH ensemble (FP64 and mixed); ensembles step member by "
                         "member, each member's step graph on its own stream, while batched ensemble stepping "
                         "is switched off in the library")
    ap.add_argument("--profile-reps", type=int, default=10)
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="only time N uneven shear-layer instances stepped as one graph")
    ap.add_argument("--large-plane", type=int, default=0, metavar="N",
                    help="only time N steps of the 256 x 193 x 256 DNS (large-plane kernels)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU y-slab design is documented, not built", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.ensemble:
        ensemble_bench(args.ensemble, args.steps, args.warmup)
        return
    if args.large_plane:
        ms = large_plane_dns_ms(args.large_plane)
        print(json.dumps({"workload": "DNS Re_tau=395, 256x193x256, FP64 (large-plane kernels)",
                          "steps": args.large_plane, "ms_per_step": round(ms, 3),
                          "mpts_steps_per_s": round(256 * 193 * 256 / (ms * 1e-3) / 1e6, 1)}), flush=True)
        return

    from diablo_amd import channel, configs

    p, gy, ic = configs.channel_re180()
    ch = channel.Channel(p, gy)
    ch.initial_condition(ic)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.time_step(args.warmup)
    ch.synchronize()
    t0 = time.perf_counter()
    ch.time_step(args.steps)
    ch.synchronize()
    wall = time.perf_counter() - t0
    t_sim, dt_now, _, ts = ch.clock()
    pts = p["nx"] * p["ny"] * p["nz"]
    value = pts * args.steps / wall / 1e6
    ms_step = wall / args.steps * 1e3

    # per-kernel HIP-event timings on the context stream (state keeps evolving)
    phases = {}
    for rk in (1, 2, 3):
        for k, v in ch.profile_substep(rk, args.profile_reps).items():
            phases[f"rk{rk}:{k}"] = v
    per_kernel = {}
    for k, v in phases.items():
        name = k.split(":", 1)[1]
        per_kernel.setdefault(name, []).append((int(k[2]), v))
    dominant = max(per_kernel, key=lambda n: sum(v for _, v in per_kernel[n]))
    launches = per_kernel[dominant]
    byts = sum(phase_bytes(p, rk)[dominant] for rk, _ in launches) / len(launches)
    mean_ms = sum(v for _, v in launches) / len(launches)
    achieved = byts / (mean_ms * 1e-3) / 1e9
    B = step_bytes(p)
    step_gbs = B / (ms_step * 1e-3) / 1e9

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": UNIT, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded laminar Poiseuille + Gaussian noise, projected divergence-free on the device)",
        "config": {"workload": "channel_re180: turbulent channel Re_tau=180, 128x129x128, 4pi x 2 x 4pi/3, "
                               "tanh grid CS=1.75, FP64, CFL=0.5 variable dt",
                   "nx": p["nx"], "ny": p["ny"], "nz": p["nz"], "n_th": p["n_th"]},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic().get(dominant),
                     "bytes_per_launch": int(byts), "mean_ms": round(mean_ms, 5)},
        "step_roofline": {"bytes_per_step": B, "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "phases_ms": {k: round(v, 5) for k, v in phases.items()},
        "sim": {"time": t_sim, "dt": dt_now, "time_step": ts},
    }
    if not args.no_mixed:
        out["mixed_form"] = mixed_form(p, gy, ic, args.steps, args.warmup)
    if not args.no_les:
        out["les_form"] = les_form(args.steps, args.warmup)
        if not args.no_cpu_baseline:
            out["les_form"]["cpu_baseline"] = les_cpu_baseline()
    if not args.no_les:
        out["learned_les"] = learned_les(ms_step)
    if not args.no_batched:
        out["batched_microbench"] = batched_microbench()
    if not args.no_ensemble:
        ens = ensemble_run(64, 100, 10)
        ens["mixed_form"] = ensemble_run(64, 100, 10, mixed=True, one_at_a_time=False)
        out["kh_ensemble"] = ens
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, gy, ic)
    ch.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
