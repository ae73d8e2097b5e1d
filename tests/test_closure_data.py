"""The closure's shipped network and its a-posteriori case are well formed
(CPU): the parameter count of dh_mlp_*, finite values, positive scales, and
a case whose profiles are ordered in y+ and whose velocity has the LES
grid's shape."""
import os

import numpy as np
import pytest

from oracle import mlp_ref as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", ["closure_re180.npz", "closure_re395.npz"])
def test_shipped_network(name):
    z = np.load(os.path.join(ROOT, "diablo_amd", "data", name))
    assert z["params"].shape == (M.NP,) and z["params"].dtype == np.float32
    assert np.all(np.isfinite(z["params"]))
    assert all(float(z[k]) > 0.0 for k in ("ts", "ls", "nus"))
    # the padded inputs 6..31 of layer 1 stay zero
    assert np.all(z["params"][M.OW1 + M.IN * M.H: M.OB1] == 0.0)


@pytest.mark.parametrize("name, ny, re_tau", [("closure_les_case.npz", 65, 180.0), ("closure_les_case_re395.npz", 97, 395.0)])
def test_les_case(name, ny, re_tau):
    z = np.load(os.path.join(ROOT, "tests", "golden", name))
    for k in ("cu1", "cu2", "cu3"):
        assert z[k].shape == (ny + 2, 66, 33) and np.all(np.isfinite(z[k]))
    for pre in ("dns", "filtered"):
        yp = z[f"{pre}_y_plus"]
        assert np.all(np.diff(yp) >= 0.0) and yp[0] > 0.0 and yp[-1] < 1.01 * re_tau
        for k in ("U_plus", "urms_plus", "vrms_plus", "wrms_plus"):
            assert z[f"{pre}_{k}"].shape == yp.shape and np.all(z[f"{pre}_{k}"] >= 0.0)
    # the mean profile of a turbulent channel: U+ = y+ at the wall (within the wall shear's distance
    # from the forcing's u_tau = 1 over the averaging: 2% at Re_tau = 180, 7% at 395), about 18
    # (Re_tau = 180) or 19-20 (Re_tau = 395) at the centre
    assert abs(z["dns_U_plus"][0] / z["dns_y_plus"][0] - 1.0) < 0.1 and 16.0 < z["dns_U_plus"][-1] < 23.0
