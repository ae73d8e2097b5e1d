"""What the trained closure gains over the constant-Smagorinsky model of
les.f:196-227, a posteriori (needs an MI355X).

Each case (tests/golden/closure_les_case*.npz) and its network
(diablo_amd/data/closure_re*.npz) come from one run of
``python -m diablo_amd.closure --case ...``:
* re180: the Re_tau = 180 channel (DNS 128 x 129 x 128, LES 64 x 65 x 64);
* re395 (config 3): the Re_tau = 395 channel, DNS 256 x 193 x 256 (spun up on
  128 x 129 x 128 and refined), 200 snapshots, LES 64 x 97 x 64.
The case holds the DNS velocity cut to the LES grid's modes (the sharp
cutoff a pseudo-spectral LES resolves) as the initial state, and the DNS's
mean and rms profiles in wall units, whole and as the LES grid resolves
them.  From that state an LES with each model runs 8000 steps, the last 4000
averaged.  The network must beat the Smagorinsky model on the mean velocity
against the DNS and on every rms profile against the resolved DNS, and its
summed error must be under half of Smagorinsky's (the pipelines' own
12000-step runs: re180 0.05 against 0.37, re395 0.09 against 0.20)."""
import os

import numpy as np
import pytest

from diablo_amd import closure

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = {
    "re180": ("closure_les_case.npz", "closure_re180.npz", closure.les_re180),
    "re395": ("closure_les_case_re395.npz", "closure_re395.npz", closure.les_re395),
}


def _profiles(z, prefix):
    return {k: z[f"{prefix}_{k}"] for k in ("y_plus",) + closure.PROFILE_KEYS}


@pytest.mark.parametrize("case", sorted(CASES))
def test_trained_closure_beats_smagorinsky(case):
    case_file, net_file, deck = CASES[case]
    z = np.load(os.path.join(ROOT, "tests", "golden", case_file))
    net = np.load(os.path.join(ROOT, "diablo_amd", "data", net_file))
    dns, filtered = _profiles(z, "dns"), _profiles(z, "filtered")
    errs = {}
    for model in ("smagorinsky", "learned"):
        print(f"[closure gain] {case} LES, {model}, 8000 steps", flush=True)
        w, info = closure.les_profiles(z, 8000, model=model, params=net["params"],
                                       scales=(float(net["ts"]), float(net["ls"]), float(net["nus"])), deck=deck)
        errs[model] = closure.profile_errors(w, dns, filtered)
        print("  " + ", ".join(f"{k} {v:.3f}" for k, v in errs[model].items()) +
              f"; u_tau {info['u_tau_wall_shear']:.3f}, {info['wall_s_per_unit_time']:.3f} s per unit time",
              flush=True)
    le, se = errs["learned"], errs["smagorinsky"]
    assert le["U_plus"] < se["U_plus"]
    for k in closure.PROFILE_KEYS[1:]:
        assert le[k + "_vs_filtered"] < se[k + "_vs_filtered"], k
    assert le["score"] < 0.5 * se["score"]
