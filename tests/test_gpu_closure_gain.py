"""What the trained closure gains over the constant-Smagorinsky model of
les.f:196-227, a posteriori (needs an MI355X).

The case (tests/golden/closure_les_case.npz) and the network
(diablo_amd/data/closure_re180.npz) come from one run of
``python -m diablo_amd.closure``: the Re_tau = 180 channel (128 x 129 x 128)
run to developed turbulence, its velocity cut to the modes of the 64 x 65 x 64
LES grid (the sharp cutoff a pseudo-spectral LES resolves) for the initial
state, and the DNS's mean and rms profiles in wall units, whole and as that
grid resolves them.  From that state an LES with each model runs 8000 steps,
the last 4000 averaged.  The network must beat the Smagorinsky model on the
mean velocity against the DNS and on every rms profile against the resolved
DNS, and its summed error must be under half of Smagorinsky's (the pipeline's
own 12000-step runs: 0.05 against 0.37)."""
import os

import numpy as np
import pytest

from diablo_amd import closure

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASE = os.path.join(ROOT, "tests", "golden", "closure_les_case.npz")
NET = os.path.join(ROOT, "diablo_amd", "data", "closure_re180.npz")


def _profiles(z, prefix):
    return {k: z[f"{prefix}_{k}"] for k in ("y_plus",) + closure.PROFILE_KEYS}


def test_trained_closure_beats_smagorinsky():
    z = np.load(CASE)
    net = np.load(NET)
    dns, filtered = _profiles(z, "dns"), _profiles(z, "filtered")
    errs = {}
    for model in ("smagorinsky", "learned"):
        print(f"[closure gain] LES 64x65x64, {model}, 8000 steps", flush=True)
        w, info = closure.les_profiles(z, 8000, model=model, params=net["params"],
                                       scales=(float(net["ts"]), float(net["ls"]), float(net["nus"])))
        errs[model] = closure.profile_errors(w, dns, filtered)
        print("  " + ", ".join(f"{k} {v:.3f}" for k, v in errs[model].items()) +
              f"; u_tau {info['u_tau_wall_shear']:.3f}, {info['wall_s_per_unit_time']:.3f} s per unit time",
              flush=True)
    le, se = errs["learned"], errs["smagorinsky"]
    assert le["U_plus"] < se["U_plus"]
    for k in closure.PROFILE_KEYS[1:]:
        assert le[k + "_vs_filtered"] < se[k + "_vs_filtered"], k
    assert le["score"] < 0.5 * se["score"]
