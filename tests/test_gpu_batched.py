"""The batched microbenchmark kernels (needs an MI355X): complex
Crank-Nicolson systems of NY = 257 rows, one per (kx, kz) of a 512 x 512
plane, and 2-D transforms of 512 x 512 planes.

Comparators: the oracle's restatement of THOMAS_COMPLEX (mpi.f:708-865,
itself pinned against scipy's banded solver) on a sample of systems, the
residual ||Ax - b|| / ||b|| in FP64 on every system, and numpy.fft
(pocketfft) with the reference's normalisation for the transforms.
Bounds: FP64 solutions within 1e-13 of the oracle and residuals <= 1e-14
(the systems are diagonally dominant, condition number < 50); FP32
residuals <= 2e-6; FP64 transforms within 1e-14 (relative L2) of numpy,
FP32 within 2e-6; the inverse of the forward transform returns the planes
to 1e-15 (FP64)."""
import numpy as np
import pytest
import torch

from diablo_amd import batched

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


def _oracle_solve(dor, sl, su, alpha, rhs):
    """THOMAS_COMPLEX of the oracle on rows 1..ny of (0:ny+1) arrays."""
    ny, ns = rhs.shape
    A = np.zeros((ny + 2, ns + 1))
    B = np.ones((ny + 2, ns + 1))
    Cm = np.zeros((ny + 2, ns + 1))
    G = np.zeros((ny + 2, ns + 1), np.complex128)
    A[1: ny + 1, :ns] = -alpha[None, :] * sl[:, None]
    Cm[1: ny + 1, :ns] = -alpha[None, :] * su[:, None]
    B[1: ny + 1, :ns] = 1.0 + alpha[None, :] * (sl + su)[:, None]
    G[1: ny + 1, :ns] = rhs
    dor.thomas_complex(A, B, Cm, G)
    return G[1: ny + 1, :ns]


def _rel_residual(sl, su, alpha, x, rhs):
    r = batched.cn_matrix_apply(sl, su, alpha, x) - rhs
    return float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(rhs))


@pytest.mark.parametrize("ny,nkx,nkz", [(257, 257, 16), (129, 64, 16), (40, 33, 8), (2, 5, 3)])
def test_cn_solve_fp64_matches_oracle(oracle_mod, ny, nkx, nkz):
    _say(f"[batched] cn solve fp64 ny={ny} systems={nkx * nkz}")
    sl, su = batched.cn_rows(ny)
    alpha, rhs = batched.cn_inputs(ny, nkx, nkz)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    x = torch.empty_like(d(rhs))
    batched.cn_solve(d(sl), d(su), d(alpha), d(rhs), x)
    torch.cuda.synchronize()
    xh = x.cpu().numpy()
    ref = _oracle_solve(oracle_mod, sl, su, alpha, rhs)
    err = np.linalg.norm(xh - ref) / np.linalg.norm(ref)
    res = _rel_residual(d(sl), d(su), d(alpha), x, d(rhs))
    _say(f"  rel err vs oracle {err:.2e}, residual {res:.2e}")
    assert err <= 1e-13 and res <= 1e-14


def test_cn_solve_in_place_and_fp32():
    ny, nkx, nkz = 257, 257, 32
    _say(f"[batched] cn solve in place + fp32 ny={ny} systems={nkx * nkz}")
    sl, su = batched.cn_rows(ny)
    alpha, rhs = batched.cn_inputs(ny, nkx, nkz)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    sl_d, su_d, al_d, b = d(sl), d(su), d(alpha), d(rhs)
    x = torch.empty_like(b)
    batched.cn_solve(sl_d, su_d, al_d, b, x)
    y = b.clone()
    batched.cn_solve(sl_d, su_d, al_d, y, y)  # x may equal rhs
    torch.cuda.synchronize()
    assert torch.equal(x, y)
    b32 = b.to(torch.complex64)
    x32 = torch.empty_like(b32)
    batched.cn_solve(sl_d, su_d, al_d, b32, x32, fp32=True)
    torch.cuda.synchronize()
    res32 = _rel_residual(sl_d, su_d, al_d, x32.to(torch.complex128), b)
    diff = float(torch.linalg.vector_norm(x32.to(torch.complex128) - x) / torch.linalg.vector_norm(x))
    _say(f"  fp32 residual (in fp64) {res32:.2e}, rel diff from fp64 {diff:.2e}")
    assert res32 <= 2e-6 and diff <= 2e-6


def test_cn_solve_full_size_residual():
    """The whole set: 257 x 512 systems of 257 rows."""
    ny, nkx, nkz = 257, 257, 512
    _say(f"[batched] cn solve full size ny={ny} systems={nkx * nkz}")
    sl, su = batched.cn_rows(ny)
    alpha, rhs = batched.cn_inputs(ny, nkx, nkz)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    sl_d, su_d, al_d, b = d(sl), d(su), d(alpha), d(rhs)
    x = torch.empty_like(b)
    batched.cn_solve(sl_d, su_d, al_d, b, x)
    torch.cuda.synchronize()
    res = _rel_residual(sl_d, su_d, al_d, x, b)
    _say(f"  residual {res:.2e}")
    assert res <= 1e-14


def _numpy_forward(p):
    """FFT_XZ_TO_FOURIER without truncation: rfft along x / NX, fft along z / NZ."""
    return np.fft.fft(np.fft.rfft(p, axis=-1) / p.shape[-1], axis=-2) / p.shape[-2]


@pytest.mark.parametrize("fp32", [False, True])
def test_fft512_against_numpy(fp32):
    n = 3
    _say(f"[batched] fft512 forward/inverse n={n} fp32={fp32}")
    rng = np.random.default_rng(4)
    ph = rng.standard_normal((n, 512, 512))
    rdt, cdt = (torch.float32, torch.complex64) if fp32 else (torch.float64, torch.complex128)
    phys = torch.from_numpy(ph).to(DEV).to(rdt)
    four = torch.empty((n, 512, 257), dtype=cdt, device=DEV)
    batched.fft_to_fourier(phys, four, fp32=fp32)
    torch.cuda.synchronize()
    ref = _numpy_forward(ph)
    f = four.cpu().numpy().astype(np.complex128)
    ef = np.linalg.norm(f - ref) / np.linalg.norm(ref)
    # inverse of the numpy spectrum (unnormalised) against numpy's
    four_ref = torch.from_numpy(ref).to(DEV).to(cdt)
    keep = four_ref.clone()
    back = torch.empty_like(phys)
    batched.fft_to_physical(four_ref, back, fp32=fp32)
    torch.cuda.synchronize()
    assert torch.equal(four_ref, keep)  # the input is left unchanged
    inv_ref = np.fft.irfft(np.fft.ifft(ref, axis=-2) * 512, n=512, axis=-1) * 512
    b = back.cpu().numpy().astype(np.float64)
    ei = np.linalg.norm(b - inv_ref) / np.linalg.norm(inv_ref)
    # round trip of the device transforms
    rt = torch.empty_like(phys)
    batched.fft_to_physical(four, rt, fp32=fp32)
    torch.cuda.synchronize()
    er = float(torch.linalg.vector_norm((rt - phys).double()) / torch.linalg.vector_norm(phys.double()))
    _say(f"  forward {ef:.2e}, inverse {ei:.2e}, round trip {er:.2e}")
    tol = 2e-6 if fp32 else 1e-14
    assert ef <= tol and ei <= tol and er <= (2e-6 if fp32 else 1e-15)


def test_fft512_single_modes():
    """cos(3x) -> 1/2 at kx = 3; sin(5z) -> -i/2 at kz = 5 and +i/2 at kz = -5."""
    _say("[batched] fft512 single modes")
    x = np.arange(512) * 2 * np.pi / 512
    p = np.zeros((2, 512, 512))
    p[0] = np.cos(3 * x)[None, :]
    p[1] = np.sin(5 * x)[:, None]
    phys = torch.from_numpy(p).to(DEV)
    four = torch.empty((2, 512, 257), dtype=torch.complex128, device=DEV)
    batched.fft_to_fourier(phys, four)
    torch.cuda.synchronize()
    f = four.cpu().numpy()
    e0 = np.zeros((512, 257), np.complex128)
    e0[0, 3] = 0.5
    e1 = np.zeros((512, 257), np.complex128)
    e1[5, 0] = -0.5j
    e1[512 - 5, 0] = 0.5j
    assert np.abs(f[0] - e0).max() < 1e-15 and np.abs(f[1] - e1).max() < 1e-15


def test_fft512_chunk_boundaries():
    """Plane counts around the 16-plane chunk: every plane equals its solo
    transform."""
    n = 17
    _say(f"[batched] fft512 {n} planes vs one at a time")
    rng = np.random.default_rng(6)
    phys = torch.from_numpy(rng.standard_normal((n, 512, 512))).to(DEV)
    four = torch.empty((n, 512, 257), dtype=torch.complex128, device=DEV)
    batched.fft_to_fourier(phys, four)
    one = torch.empty((1, 512, 257), dtype=torch.complex128, device=DEV)
    back = torch.empty_like(phys)
    batched.fft_to_physical(four, back)
    back1 = torch.empty((1, 512, 512), dtype=torch.float64, device=DEV)
    for i in (0, 15, 16):
        batched.fft_to_fourier(phys[i: i + 1].contiguous(), one)
        batched.fft_to_physical(four[i: i + 1].contiguous(), back1)
        torch.cuda.synchronize()
        assert torch.equal(one[0], four[i]) and torch.equal(back1[0], back[i])
