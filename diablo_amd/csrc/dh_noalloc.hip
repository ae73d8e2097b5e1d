// An entry point that allocates nothing: dh_fft512_xz_to_physical with its
// staging buffer in a file-scope device array instead of memory the library
// allocates.  For callers that cannot allocate device memory (code under
// stream capture, tight serving loops, harnesses that forbid it).
#include <hip/hip_runtime.h>

#include <mutex>

#include "dh_internal.h"

namespace dh {
hipError_t fft512_inverse(const void *four, void *phys, long long nplanes, bool fp32, hipStream_t st, void *stage,
                          long long chunk);
const char *batched_set_error(int code, const char *msg);
}  // namespace dh

namespace {
// 8 planes of [512][257] FP64 complex values: 16.8 MB on each device the
// library runs on, held until the process ends.  Host code uses the address
// hipGetSymbolAddress returns for the current device, never the array's name.
constexpr long long kScratchPlanes = 8;
constexpr size_t kScratchElems = (size_t)kScratchPlanes * 512 * 257;
__device__ __align__(256) double2 scratch[kScratchElems];

std::mutex g_mu;
void *g_scratch[64] = {nullptr};  // address of `scratch` on each device
}  // namespace

extern "C" int dh_fft512_xz_to_physical_noalloc_dev(const void *four, void *phys, int64_t nplanes, int32_t fp32,
                                                    void *stream) {
  if (!phys || !four) return (dh::batched_set_error(DH_EARG, "null pointer"), DH_EARG);
  if (nplanes <= 0) return (dh::batched_set_error(DH_EARG, "nplanes must be positive"), DH_EARG);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (dh::batched_set_error(DH_EHIP, hipGetErrorString(e)), DH_EHIP);
  if (dev < 0 || dev >= 64) return (dh::batched_set_error(DH_EARG, "device index beyond 63"), DH_EARG);
  void *sp = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return (dh::batched_set_error(DH_EHIP, hipGetErrorString(e)), DH_EHIP);
      g_scratch[dev] = p;
    }
    sp = g_scratch[dev];
  }
  // the planes pass through the scratch kScratchPlanes at a time, so any
  // number of planes fits it
  e = dh::fft512_inverse(four, phys, nplanes, fp32 != 0, static_cast<hipStream_t>(stream), sp, kScratchPlanes);
  if (e != hipSuccess) return (dh::batched_set_error(DH_EHIP, hipGetErrorString(e)), DH_EHIP);
  return DH_OK;
}
