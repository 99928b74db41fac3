// chol_common.h — shared helpers of the gfx950 Cholesky kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "chol_amd.h"

namespace chol {

// Per-thread message for chol_last_error().
void set_error(const char* fmt, ...);
void clear_error();

// Wave size of CDNA: every kernel here is written for 64-lane wavefronts.
constexpr int kWave = 64;

// Arguments of one batched factorisation (row-major, see chol_amd.h).
struct Batch {
  const float* A;
  float* L;
  int64_t batch;
  int n;
  int64_t bs;  // batch stride (elements)
  int64_t ld;  // leading dimension (elements)
  int32_t* info;
};

// Compiler-level barrier between LDS traffic of lanes of one wavefront.
// LDS instructions of a wave execute in order, so lanes exchanging values
// through LDS inside one wave need no s_barrier, only this fence against
// compiler reordering.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

inline int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

// Launchers implemented in the kernel translation units.
int launch_small(const Batch& p, hipStream_t s);                  // n <= 64
int launch_lds(const Batch& p, hipStream_t s);                    // n <= 256
int launch_copy_lower(const Batch& p, hipStream_t s);             // L = tril(A)
int launch_diag_factor(const Batch& p, int k0, float* winv,
                       hipStream_t s);                            // 64x64 block
int launch_panel(const Batch& p, int k0, const float* winv,
                 hipStream_t s);                                  // L21 = A21 W^T
// C -= P P^T over {col0 <= j < col1, j <= i < n}, K = [k0, k0+kl);
// split = BF16x3 products (mixed mode) instead of FP32.
int launch_syrk(const Batch& p, int col0, int col1, int k0, int kl, bool split,
                hipStream_t s);
int launch_zero_info(int32_t* info, int64_t batch, hipStream_t s);

// Diagonal block width of the blocked path.
constexpr int kNB = 64;

}  // namespace chol
