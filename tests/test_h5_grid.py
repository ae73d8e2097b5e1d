"""The library's HDF5 grid reader (dh_h5_read_f64, the subset ReadGridHDF5
reads, hdf5.f:548-710) on the reference's own grid file
(KH_test/grid.h5, copied to tests/golden/kh_test_grid.h5).  CPU only: the
reader is host code.

The file was written by pre_process/create_grid_h5.m; its /grids/y holds
GY(1..66) of a cubic grid refined in the centre (GRID_TYPE 2, CS = 1.5,
LY = 20).  configs.make_gy restates that generator independently, so the
two agreeing pins both the reader and the restated grid formula."""
import os

import numpy as np
import pytest

from diablo_amd import channel, configs

FIX = os.path.join(os.path.dirname(__file__), "golden", "kh_test_grid.h5")


def test_reads_reference_grid_file():
    y = channel.read_h5_f64(FIX, "/grids/y")
    assert y.shape == (66,)
    np.testing.assert_allclose(y, configs.make_gy(65, 20.0, 1.5, 2), rtol=0, atol=1e-14)
    # a different generator setting is far off: the comparison can fail
    assert np.abs(y - configs.make_gy(65, 20.0, 1.0, 2)).max() > 1e-2
    assert np.all(np.diff(y) > 0)


@pytest.mark.parametrize("ds,msg", [("/grids/x", "no member"), ("/grids", "not a dataset"),
                                    ("/nothing/y", "no member"), ("", "empty")])
def test_missing_members_are_errors(ds, msg):
    with pytest.raises(channel.DiabloError, match=msg):
        channel.read_h5_f64(FIX, ds)


def test_not_hdf5_and_truncated(tmp_path):
    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not an hdf5 file at all" * 10)
    with pytest.raises(channel.DiabloError, match="not an HDF5 file"):
        channel.read_h5_f64(bad, "/grids/y")
    data = open(FIX, "rb").read()
    cut = tmp_path / "cut.h5"
    cut.write_bytes(data[:2500])  # the dataset's storage ends at byte 2960
    with pytest.raises(channel.DiabloError, match="truncated"):
        channel.read_h5_f64(cut, "/grids/y")
    with pytest.raises(channel.DiabloError, match="cannot open"):
        channel.read_h5_f64(tmp_path / "missing.h5", "/grids/y")


def test_small_buffer_reports_size():
    import ctypes
    lib = channel.load_library()
    n = ctypes.c_int64(0)
    out = np.zeros(10)
    err = ctypes.create_string_buffer(128)
    rc = lib.dh_h5_read_f64(FIX.encode(), b"/grids/y", out.ctypes.data, 10, ctypes.byref(n), err, 128)
    assert rc != 0 and n.value == 66 and b"too small" in err.value


def test_corrupted_bytes_never_crash(tmp_path):
    """Single-byte corruptions anywhere in the file give an error or a
    result, never a crash: every read is bounds-checked."""
    data = bytearray(open(FIX, "rb").read())
    rng = np.random.default_rng(7)
    f = tmp_path / "fz.h5"
    outcomes = {"ok": 0, "err": 0}
    for trial in range(400):
        d = bytearray(data)
        for _ in range(int(rng.integers(1, 4))):
            d[int(rng.integers(0, len(d)))] = int(rng.integers(0, 256))
        f.write_bytes(bytes(d))
        try:
            channel.read_h5_f64(f, "/grids/y")
            outcomes["ok"] += 1
        except channel.DiabloError:
            outcomes["err"] += 1
    assert outcomes["err"] > 0 and outcomes["ok"] + outcomes["err"] == 400
