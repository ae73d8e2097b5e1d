"""The allocation-free entry point dh_fft512_xz_to_physical_noalloc_dev
against dh_fft512_xz_to_physical_dev (needs an MI355X): two calls in a row,
the second on fewer planes, each bit-identical to the allocating entry
point's result (the same kernels; the planes pass through the static
scratch 8 at a time instead of the library's 16-plane staging buffer)."""
import numpy as np
import pytest
import torch

from diablo_amd import batched, channel

pytestmark = pytest.mark.gpu


def _noalloc():
    return channel.load_library().dh_fft512_xz_to_physical_noalloc_dev


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_matches_allocating_entry_point(fp32):
    fn = _noalloc()
    rdt, cdt = (torch.float32, torch.complex64) if fp32 else (torch.float64, torch.complex128)
    rng = np.random.default_rng(11)
    for n in (19, 5):  # more planes than the scratch holds, then fewer
        print(f"[no_alloc] fft512 inverse n={n} fp32={fp32}", flush=True)
        phys = torch.from_numpy(rng.standard_normal((n, 512, 512))).to("cuda:0").to(rdt)
        four = torch.empty((n, 512, 257), dtype=cdt, device="cuda:0")
        batched.fft_to_fourier(phys, four, fp32=fp32)
        ref = torch.empty_like(phys)
        batched.fft_to_physical(four, ref, fp32=fp32)
        out = torch.full_like(phys, float("nan"))
        rc = fn(four.data_ptr(), out.data_ptr(), n, 1 if fp32 else 0, None)
        assert rc == 0, channel.load_library().dh_batched_last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
