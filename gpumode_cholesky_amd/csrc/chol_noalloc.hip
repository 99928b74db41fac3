// chol_noalloc.hip — chol_potrf_f32 for callers that cannot allocate (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): the same kernels and launches, with the blocked route's
// workspace taken from a device array this library holds, and all work on
// the caller's stream (no side stream or events are created).
//
// The array (`scratch`, 64 MiB per device, held until the process ends)
// carries the blocked and fused routes' per-matrix 128 x 128 block
// inverses, 64 KiB a matrix, so one call takes at most 1024 matrices of
// order > 64; orders <= 64 and the whole-matrix LDS route use none of it.  Every call on a
// device shares that device's array: calls that overlap on one device (two
// streams, two host threads) race on it and must be serialised by the
// caller.  Nothing needs to start at zero: each diagonal-block launch
// writes a matrix's inverse before its panel launch reads it.
#include <mutex>

#include "chol_common.h"

namespace chol {

constexpr size_t kScratchBytes = size_t(CHOL_NOALLOC_SCRATCH_BYTES);
static_assert(kScratchBytes % 256 == 0, "scratch size");

// In host code `scratch` names a host-side stand-in: only its address from
// hipGetSymbolAddress may reach a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];  // per device, set on the first call there

// Device address of this device's `scratch`, looked up once per device with
// that device current.
static int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= 64) {
    set_error("chol_potrf_f32_noalloc: hipGetDevice failed (%d)", (int)e);
    return e != hipSuccess ? (int)e : CHOL_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("chol_potrf_f32_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return e != hipSuccess ? (int)e : CHOL_E_WORKSPACE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace chol

using namespace chol;

extern "C" int chol_potrf_f32_noalloc(const float* A, float* L, int64_t batch, int64_t n, int64_t batch_stride,
                                      int64_t ld, int32_t mode, int32_t* info, void* stream) {
  clear_error();
  const size_t need = potrf_workspace(batch, n);
  if (need > kScratchBytes) {
    set_error("chol_potrf_f32_noalloc: %lld matrices of order %lld need %zu bytes of scratch; the library holds "
              "%zu (at most %zu matrices of order > 64 per call)",
              (long long)batch, (long long)n, need, kScratchBytes, kScratchBytes / (kNB * kNB * sizeof(float)));
    return CHOL_E_WORKSPACE;
  }
  void* ws = nullptr;
  if (need > 0) {
    const int rc = scratch_address(&ws);
    if (rc) return rc;
  }
  return potrf_dispatch("chol_potrf_f32_noalloc", A, L, batch, n, batch_stride, ld, mode, info, ws,
                        ws ? kScratchBytes : 0, reinterpret_cast<hipStream_t>(stream), false);
}
