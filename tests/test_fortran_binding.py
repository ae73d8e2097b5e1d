"""The ISO_C_BINDING module and a Fortran driver built on it compile with
amdflang and link against libdiablo_hip.so (CPU); on the GPU the driver runs
a time step from Fortran and must reproduce the Python-driven result
exactly (same library, same inputs)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from diablo_amd import channel
from tests import helpers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
MOD = os.path.join(ROOT, "diablo_amd", "fortran", "diablo_hip.f90")
DRV = os.path.join(ROOT, "tests", "fortran", "drive_substep.f90")


def _build(tmp):
    libdir = os.path.dirname(channel.LIB_PATH)
    exe = os.path.join(tmp, "drive_substep")
    cmd = [FC, "-O1", MOD, DRV, "-o", exe, f"-L{libdir}", "-ldiablo_hip", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_module_and_driver_compile_and_link(tmp_path):
    _build(str(tmp_path))


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_drop_in_rk_chan_1_compiles_and_links(tmp_path):
    """The fixed-form drop-in RK_CHAN_1 compiles against declarations of the
    reference header's names (tests/fortran/shim_include) and links with the
    binding module and the library."""
    tmp = str(tmp_path)
    libdir = os.path.dirname(channel.LIB_PATH)
    inc = os.path.join(ROOT, "tests", "fortran", "shim_include")
    shim = os.path.join(ROOT, "diablo_amd", "fortran", "rk_chan_1.f")
    main = os.path.join(ROOT, "tests", "fortran", "shim_main.f")
    exe = os.path.join(tmp, "shim_main")
    # fixed form ends a statement at column 72: the drop-in must fit
    for path in (shim, main):
        for n, line in enumerate(open(path), 1):
            assert line[:1] in "Cc!*" or len(line.rstrip("\n")) <= 72, f"{path}:{n} beyond column 72"
    steps = [[FC, "-O1", "-c", MOD, "-o", "diablo_hip.o"],
             [FC, "-O1", "-ffixed-form", f"-I{inc}", "-c", shim, "-o", "rk_chan_1.o"],
             [FC, "-O1", "-ffixed-form", f"-I{inc}", "-c", main, "-o", "shim_main.o"],
             [FC, "shim_main.o", "rk_chan_1.o", "diablo_hip.o", "-o", exe, f"-L{libdir}", "-ldiablo_hip",
              f"-Wl,-rpath,{libdir}"]]
    for cmd in steps:
        r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
        assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout + r.stderr
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_fortran_driver_matches_python(tmp_path):
    from oracle import dor
    dor.load()
    tmp = str(tmp_path)
    exe = _build(tmp)
    o, p, gy, ic = helpers.make_oracle(dor, "small_case")
    st = helpers.oracle_state(dor, o)
    nsteps = 2
    # Fortran input: dh_params bytes, clock, grid, state in Fortran order
    prm = channel.make_params(p)
    GY, GYF, DY, DYF = channel.grid_metrics(gy)
    with open(os.path.join(tmp, "fortran_in.bin"), "wb") as f:
        f.write(bytes(prm))
        f.write(np.array([nsteps, 0], np.int32).tobytes())
        f.write(np.array([0.0, p["delta_t"]]).tobytes())
        for a in (GY, GYF, DY, DYF):
            f.write(a.tobytes())
        for k in ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3", "cth", "cfth", "th"):
            f.write(np.ascontiguousarray(st[k]).tobytes())
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert f"sizeof(dh_params) = {ctypes.sizeof(channel.DhParams)}" in r.stdout
    raw = open(os.path.join(tmp, "fortran_out.bin"), "rb").read()
    t, dt = np.frombuffer(raw[:16], np.float64)
    ts = int(np.frombuffer(raw[16:20], np.int32)[0])
    off = 20
    cshape = st["cu1"].shape
    out = {}
    for k in ("cu1", "cu2", "cu3", "cp"):
        n = int(np.prod(cshape)) * 16
        out[k] = np.frombuffer(raw[off:off + n], np.complex128).reshape(cshape)
        off += n
    # the same run through the Python host mirror
    ch = channel.Channel(p, gy)
    ch.set_clock(0.0, p["delta_t"], 0)
    ch.upload(st)
    ch.time_step(nsteps)
    ref = ch.download()
    t2, dt2, _, ts2 = ch.clock()
    ch.close()
    assert ts == ts2 == nsteps and t == t2 and dt == dt2
    for k in ("cu1", "cu2", "cu3", "cp"):
        assert np.array_equal(out[k], ref[k]), k
