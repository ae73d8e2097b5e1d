"""Writes the golden fixtures of the channel timestep from the oracle (the
reference itself cannot be built here and ships no outputs):

- small_case.npz: the full de-aliased state of the 16x17x8 deck (every term:
  scalar, rotation, oscillating forcing, mixed BCs) at its seeded initial
  condition and after 3 RK3 steps;
- kh_shear_layer.npz, channel_re180.npz: fingerprints of the state after a
  few steps (per-plane L2 norms and 64 seeded random projections of each
  field), since the full 128x129x128 state is ~70 MB.

Run from the repository root: python -m tests.golden.make_golden
"""
import os

import numpy as np

from oracle import dor
from tests import helpers

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3")
STEPS = {"small_case": 3, "kh_shear_layer": 3, "channel_re180": 1}
NPROJ = 64


def fingerprint(st, nkx, tnkz, n_th, seed=2024):
    """Per-plane L2 norms and seeded complex projections of each field."""
    out = {}
    rng = np.random.default_rng(seed)
    arrays = {k: helpers.dealiased(st[k], nkx, tnkz) for k in FIELDS}
    for n in range(n_th):
        arrays[f"cth{n}"] = helpers.dealiased(st["cth"][n], nkx, tnkz)
    for k, a in arrays.items():
        out[f"{k}_plane"] = np.sqrt((np.abs(a) ** 2).sum(axis=(1, 2)))
        r = rng.standard_normal((NPROJ,) + a.shape) + 1j * rng.standard_normal((NPROJ,) + a.shape)
        out[f"{k}_proj"] = np.real(np.tensordot(np.conj(r), a, axes=a.ndim))
    return out


def run(name):
    dor.load()
    o, p, gy, ic = helpers.make_oracle(dor, name)
    o.set_clock(0.0, p["delta_t"], 0)
    ic_state = helpers.oracle_state(dor, o)
    o.time_step(STEPS[name])
    st = helpers.oracle_state(dor, o)
    t, dt, ts = o.clock()
    return o, p, ic_state, st, (t, dt, ts)


def main():
    for name in STEPS:
        o, p, ic_state, st, clk = run(name)
        meta = dict(steps=STEPS[name], time=clk[0], dt=clk[1], time_step=clk[2])
        path = os.path.join(HERE, f"{name}.npz")
        if name == "small_case":
            data = {f"ic_{k}": helpers.dealiased(ic_state[k], o.nkx, o.tnkz) for k in FIELDS}
            data.update({k: helpers.dealiased(st[k], o.nkx, o.tnkz) for k in FIELDS})
            data["cth"] = helpers.dealiased(st["cth"], o.nkx, o.tnkz)
            data["th"] = st["th"][:, :, : o.nz, : o.nx]
        else:
            data = fingerprint(st, o.nkx, o.tnkz, o.n_th)
        np.savez_compressed(path, **data, **{k: np.float64(v) for k, v in meta.items()})
        print(path, os.path.getsize(path), "bytes", flush=True)


if __name__ == "__main__":
    main()
