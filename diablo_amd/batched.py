"""Host side of the batched microbenchmark kernels (no channel context):
the Crank-Nicolson systems of the wall-normal solves and the 512 x 512
plane transforms, over device arrays held in torch tensors.

The workload is the reference's inner work at a larger scale: one complex
tridiagonal system per (kx, kz) of a 512 x 512 plane (257 x 512 systems of
NY = 257 rows), as THOMAS_FORWARD/BACKWARD_COMPLEX_MPI (``mpi.f:708-865``)
sweep them with the Crank-Nicolson matrices of ``channel.f:914-963``, and
the x/z transforms of ``FFT_XZ_TO_FOURIER/PHYSICAL`` (``mpi.f:1477-1606``)
on 257 planes of 512 x 512.
"""
import math

import numpy as np

from . import channel, configs

NX = NZ = 512
NK = NX // 2 + 1


def _check(rc):
    if rc != 0:
        lib = channel.load_library()
        raise RuntimeError(f"libdiablo_hip: {lib.dh_batched_last_error().decode()} (code {rc})")


def cn_rows(ny, ly=2.0, cs=1.75):
    """Row tables (sl, su) of the Crank-Nicolson operator on the tanh grid of
    ``create_grid_h5.m`` (GRID_TYPE 1): rows j = 0..ny-1 are the reference's
    J = 1..NY; J = 1 and J = NY are identity (Dirichlet) rows; interior rows
    carry h^2/(DY(J) DYF(J)) and h^2/(DY(J+1) DYF(J)) with h = LY/NY."""
    gy = configs.make_gy(ny, ly, cs, 1)
    _, _, dy, dyf = channel.grid_metrics(gy)
    h = ly / ny
    sl = np.zeros(ny)
    su = np.zeros(ny)
    for J in range(2, ny):
        sl[J - 1] = h * h / (dy[J] * dyf[J])
        su[J - 1] = h * h / (dy[J + 1] * dyf[J])
    return sl, su


def cn_inputs(ny=257, nkx=257, nkz=512, seed_alpha=2, seed_rhs=3):
    """alpha per system, log-uniform in [1e-3, 10] (seed 2), and a complex
    Gaussian right-hand side N(0,1) + i N(0,1) (seed 3), [ny][nkz*nkx]."""
    nsys = nkx * nkz
    alpha = np.exp(np.random.default_rng(seed_alpha).uniform(math.log(1e-3), math.log(10.0), nsys))
    rng = np.random.default_rng(seed_rhs)
    rhs = rng.standard_normal((ny, nsys)) + 1j * rng.standard_normal((ny, nsys))
    return alpha, rhs


def cn_matrix_apply(sl, su, alpha, x):
    """A x for the systems of dh_cn_solve_batched (numpy or torch arrays of
    shape [ny][nsys]); the couplings outside rows 0..ny-1 are dropped."""
    a = -alpha[None, :] * sl[:, None]
    c = -alpha[None, :] * su[:, None]
    b = 1.0 + alpha[None, :] * (sl + su)[:, None]
    y = b * x
    y[1:] += a[1:] * x[:-1]
    y[:-1] += c[:-1] * x[1:]
    return y


def cn_solve(sl, su, alpha, rhs, x, fp32=False, stream=None):
    """Solve on the device: sl, su [ny] float64, alpha [nsys] float64, rhs and
    x [ny][nsys] complex128 (complex64 with fp32) torch tensors on cuda."""
    lib = channel.load_library()
    ny, nsys = rhs.shape
    assert x.shape == rhs.shape and sl.numel() == ny and su.numel() == ny and alpha.numel() == nsys
    for t in (sl, su, alpha, rhs, x):
        assert t.is_cuda and t.is_contiguous()
    _check(lib.dh_cn_solve_batched_dev(sl.data_ptr(), su.data_ptr(), alpha.data_ptr(), rhs.data_ptr(),
                                       x.data_ptr(), ny, nsys, 1 if fp32 else 0, stream))


def fft_to_fourier(phys, four, fp32=False, stream=None):
    """phys [n][512][512] real -> four [n][512][257] complex (torch, cuda)."""
    lib = channel.load_library()
    n = phys.shape[0]
    assert phys.shape == (n, NZ, NX) and four.shape == (n, NZ, NK)
    assert phys.is_contiguous() and four.is_contiguous() and phys.is_cuda and four.is_cuda
    _check(lib.dh_fft512_xz_to_fourier_dev(phys.data_ptr(), four.data_ptr(), n, 1 if fp32 else 0, stream))


def fft_to_physical(four, phys, fp32=False, stream=None):
    """four [n][512][257] complex -> phys [n][512][512] real (unnormalised)."""
    lib = channel.load_library()
    n = phys.shape[0]
    assert phys.shape == (n, NZ, NX) and four.shape == (n, NZ, NK)
    assert phys.is_contiguous() and four.is_contiguous() and phys.is_cuda and four.is_cuda
    _check(lib.dh_fft512_xz_to_physical_dev(four.data_ptr(), phys.data_ptr(), n, 1 if fp32 else 0, stream))
