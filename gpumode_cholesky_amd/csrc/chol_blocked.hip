// chol_blocked.hip — blocked right-looking Cholesky for n > 64.
//
// The factor is built in place in the output L (the input A is only read by
// the first copy).  The matrix is walked in 64-wide block columns:
//   diag  : factor the 64x64 diagonal block in one wavefront's registers and
//           form its inverse W = L11^-1 (kept in the workspace);
//   panel : L21 = A21 W^T for every row below the block (a GEMM, K = 64);
//   syrk  : C -= P P^T on the lower trailing tiles.
// The two products run on the matrix cores (chol_gemm.hip).
// Inner 64-blocks are grouped into an outer block of `nbo` columns: inner
// steps update only the outer block's columns, and one SYRK with K = nbo then
// updates the rest of the trailing matrix, so the dominant update streams
// its C tile once per nbo columns instead of once per 64.
//
// Reference counterparts: the 128-wide diag+inverse kernel e62_diag128
// (submission.py:4694-4997), the panel-times-inverse bmm and trailing
// baddbmm_ of _exp062_factor (submission.py:5152-5196), the two-level
// micro/panel blocking of _split32_launch (submission.py:2375-2567) and the
// trsm-free inverse panels of the large-n drivers (submission.py:3606-3676).
#include "chol_common.h"

namespace chol {
namespace {

// ---------------------------------------------------------------- copy ----
// L = tril(A) with exact zeros above the diagonal.
__global__ __launch_bounds__(256) void k_copy_lower(Batch p) {
  const int n = p.n;
  const int64_t b = blockIdx.y;
  const float* src = p.A + b * p.bs;
  float* dst = p.L + b * p.bs;
  const int64_t total = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n);
    const int j = (int)(e - (int64_t)i * n);
    dst[(int64_t)i * p.ld + j] = (j <= i) ? src[(int64_t)i * p.ld + j] : 0.f;
  }
}

// Vector form for n % 4 == 0 and 16-byte aligned rows.
__global__ __launch_bounds__(256) void k_copy_lower4(Batch p) {
  const int n = p.n;
  const int n4 = n / 4;
  const int64_t b = blockIdx.y;
  const float* src = p.A + b * p.bs;
  float* dst = p.L + b * p.bs;
  const int64_t total = (int64_t)n * n4;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n4);
    const int j = 4 * (int)(e - (int64_t)i * n4);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j <= i) v = *reinterpret_cast<const float4*>(src + (int64_t)i * p.ld + j);
    if (j + 1 > i) v.y = 0.f;
    if (j + 2 > i) v.z = 0.f;
    if (j + 3 > i) v.w = 0.f;
    *reinterpret_cast<float4*>(dst + (int64_t)i * p.ld + j) = v;
  }
}

__global__ void k_zero_info(int32_t* info, int64_t batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch) info[i] = 0;
}

// ---------------------------------------------------------------- diag ----
// One wavefront per matrix: factor L[k0:k0+64, k0:k0+64] (already fully
// updated) in registers, write it back, and write W = L11^-1 (64x64,
// row-major, zeros above the diagonal) to winv[b].
// Rows past n are padded with the identity, so the partial last block needs
// no special case.
__global__ __launch_bounds__(kWave) void k_diag_factor(Batch p, int k0, float* __restrict__ winv) {
  constexpr int NB = kNB;
  constexpr int PITCH = NB + 4;
  __shared__ __align__(16) float Ls[NB * PITCH];
  __shared__ __align__(16) float cb[NB];
  const int r = threadIdx.x;
  const int64_t b = blockIdx.x;
  const int nb = min(NB, p.n - k0);
  float* Lb = p.L + b * p.bs;
  float* row = Lb + (int64_t)(k0 + (r < nb ? r : 0)) * p.ld + k0;

  float a[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    float v;
    if (r < nb)
      v = (j <= r) ? row[j] : 0.f;
    else
      v = (j == r) ? 1.f : 0.f;
    a[j] = v;
  }

  int32_t bad = 0;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    cb[r] = a[k];
    wave_lds_sync();
    float c[NB];
#pragma unroll
    for (int q = (k / 4) * 4; q < NB; q += 4) {
      const float4 v = *reinterpret_cast<const float4*>(cb + q);
      c[q] = v.x;
      c[q + 1] = v.y;
      c[q + 2] = v.z;
      c[q + 3] = v.w;
    }
    const float piv = c[k];
    if (bad == 0 && !(piv > 0.f)) bad = k + 1;
    const float s = sqrtf(piv);
    const float is = 1.0f / s;
    const float lk = a[k] * is;
    const float t = (r > k) ? lk * is : 0.f;
#pragma unroll
    for (int j = k + 1; j < NB; ++j) a[j] = fmaf(-t, c[j], a[j]);
    a[k] = (r > k) ? lk : ((r == k) ? s : 0.f);
    wave_lds_sync();
  }

  if (r < nb) {
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j <= r) row[j] = a[j];
  }
  if (p.info && r == 0 && bad != 0 && bad <= nb) {
    if (p.info[b] == 0) p.info[b] = k0 + bad;
  }

  // Inverse: stage L11 in LDS; lane c forward-substitutes column c of W.
#pragma unroll
  for (int j = 0; j < NB; j += 4)
    *reinterpret_cast<float4*>(&Ls[r * PITCH + j]) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
  __syncthreads();
  const int c = r;
  float w[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    float s = (i == c) ? 1.f : 0.f;
    const float* Li = &Ls[i * PITCH];
#pragma unroll
    for (int q = 0; q < i; q += 4) {
      const float4 v = *reinterpret_cast<const float4*>(Li + q);
      s = fmaf(-v.x, w[q], s);
      if (q + 1 < i) s = fmaf(-v.y, w[q + 1], s);
      if (q + 2 < i) s = fmaf(-v.z, w[q + 2], s);
      if (q + 3 < i) s = fmaf(-v.w, w[q + 3], s);
    }
    w[i] = s / Li[i];
  }
  float* Wb = winv + b * (NB * NB);
#pragma unroll
  for (int i = 0; i < NB; ++i) Wb[i * NB + c] = w[i];
}

}  // namespace

int launch_zero_info(int32_t* info, int64_t batch, hipStream_t s) {
  if (!info || batch <= 0) return 0;
  hipLaunchKernelGGL(k_zero_info, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, info, batch);
  return launch_status("k_zero_info");
}

int launch_copy_lower(const Batch& p, hipStream_t s) {
  const bool vec = ((reinterpret_cast<uintptr_t>(p.A) | reinterpret_cast<uintptr_t>(p.L)) & 15) == 0 &&
                   (p.ld % 4) == 0 && (p.bs % 4) == 0 && (p.n % 4) == 0;
  const int64_t elems = vec ? (int64_t)p.n * (p.n / 4) : (int64_t)p.n * p.n;
  int64_t gx = (elems + 255) / 256;
  if (gx > 4096) gx = 4096;
  if (vec)
    hipLaunchKernelGGL(k_copy_lower4, dim3((unsigned)gx, (unsigned)p.batch), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL(k_copy_lower, dim3((unsigned)gx, (unsigned)p.batch), dim3(256), 0, s, p);
  return launch_status("k_copy_lower");
}

int launch_diag_factor(const Batch& p, int k0, float* winv, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_factor, dim3((unsigned)p.batch), dim3(kWave), 0, s, p, k0, winv);
  return launch_status("k_diag_factor");
}

}  // namespace chol
