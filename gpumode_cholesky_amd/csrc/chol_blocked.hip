// chol_blocked.hip — blocked right-looking Cholesky for n > 64.
//
// The factor is built in place in the output L (the input A is only read by
// the first copy).  The matrix is walked in 64-wide block columns:
//   diag  : factor the 64x64 diagonal block in one wavefront's registers and
//           form its inverse W = L11^-1 (kept in the workspace);
//   panel : L21 = A21 W^T for every row below the block (a GEMM, K = 64);
//   syrk  : C -= P P^T on the lower trailing tiles.
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

// --------------------------------------------------------------- panel ----
// L[i, k0:k0+64] = A21[i, :] W^T for rows i in [k0+64, n), 64 rows per block.
// W is lower-triangular, so out[i][j] = sum_{c <= j} A21[i][c] W[j][c].
__global__ __launch_bounds__(256) void k_panel(Batch p, int k0, const float* __restrict__ winv) {
  constexpr int NB = kNB;
  constexpr int P = 64 + 4;
  __shared__ __align__(16) float As[NB * P];  // As[c][row]
  __shared__ __align__(16) float Ws[NB * P];  // Ws[c][col] = W[col][c]
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.y;
  const int row0 = k0 + NB + blockIdx.x * 64;
  const int n = p.n;
  float* Lb = p.L + b * p.bs;
  const float* Wb = winv + b * (NB * NB);

  // Load: thread t covers row t/4 and 16 columns (t%4)*16 .. +15.
  {
    const int rr = tid >> 2;
    const int c0 = (tid & 3) * 16;
    const int gi = row0 + rr;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int cc = c0 + u;
      As[cc * P + rr] = (gi < n) ? Lb[(int64_t)gi * p.ld + k0 + cc] : 0.f;
      Ws[cc * P + rr] = Wb[rr * NB + cc];
    }
  }
  __syncthreads();
  const int tx = tid & 15;  // output cols tx*4 .. +3
  const int ty = tid >> 4;  // output rows ty*4 .. +3
  float acc[4][4] = {};
#pragma unroll 8
  for (int c = 0; c < NB; ++c) {
    const float4 av = *reinterpret_cast<const float4*>(&As[c * P + ty * 4]);
    const float4 wv = *reinterpret_cast<const float4*>(&Ws[c * P + tx * 4]);
    const float ar[4] = {av.x, av.y, av.z, av.w};
    const float wr[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(ar[i], wr[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = row0 + ty * 4 + i;
    if (gi >= n) continue;
    float* dst = Lb + (int64_t)gi * p.ld + k0 + tx * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[j] = acc[i][j];
  }
}

// ---------------------------------------------------------------- syrk ----
// C[i][j] -= sum_{c in [k0, k0+kl)} L[i][c] L[j][c]
// for j in [col0, col1), i in [j, n).  64x64 tiles (ti >= tj) over the
// lower region; tile (ti, tj) covers rows col0+64*ti.., cols col0+64*tj...
// Grid.x enumerates tiles column-major: column tj holds TI - tj tiles.
__device__ __forceinline__ void decode_tile(int idx, int TI, int& ti, int& tj) {
  // Tiles before column t: S(t) = t*TI - t*(t-1)/2.  Solve S(t) <= idx.
  const float twoTI1 = 2.f * TI + 1.f;
  const float disc = fmaxf(twoTI1 * twoTI1 - 8.f * (float)idx, 0.f);
  int t = (int)floorf((twoTI1 - sqrtf(disc)) * 0.5f);
  if (t < 0) t = 0;
  while (t > 0 && (int64_t)t * TI - (int64_t)t * (t - 1) / 2 > idx) --t;
  while ((int64_t)(t + 1) * TI - (int64_t)(t + 1) * t / 2 <= idx) ++t;
  tj = t;
  ti = t + (idx - (int)((int64_t)t * TI - (int64_t)t * (t - 1) / 2));
}

__global__ __launch_bounds__(256) void k_syrk(Batch p, int col0, int col1, int k0, int kl, int TI) {
  constexpr int KC = 16;
  constexpr int P = 64 + 4;
  __shared__ __align__(16) float As[KC * P];  // As[c][i]
  __shared__ __align__(16) float Bs[KC * P];  // Bs[c][j]
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.y;
  int ti, tj;
  decode_tile(blockIdx.x, TI, ti, tj);
  const int i0 = col0 + 64 * ti;
  const int j0 = col0 + 64 * tj;
  const int n = p.n;
  float* Lb = p.L + b * p.bs;

  const int tx = tid & 15;
  const int ty = tid >> 4;
  float acc[4][4] = {};
  const int lr = tid >> 2;        // load row 0..63
  const int lc = (tid & 3) * 4;   // load cols lc..lc+3 within the chunk
  const int gi = i0 + lr;
  const int gj = j0 + lr;
  for (int kc = 0; kc < kl; kc += KC) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = kc + lc + u;
      const bool kin = c < kl;
      As[(lc + u) * P + lr] = (gi < n && kin) ? Lb[(int64_t)gi * p.ld + k0 + c] : 0.f;
      Bs[(lc + u) * P + lr] = (gj < col1 && kin) ? Lb[(int64_t)gj * p.ld + k0 + c] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      const float4 av = *reinterpret_cast<const float4*>(&As[c * P + ty * 4]);
      const float4 bv = *reinterpret_cast<const float4*>(&Bs[c * P + tx * 4]);
      const float ar[4] = {av.x, av.y, av.z, av.w};
      const float br[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(ar[i], br[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int ri = i0 + ty * 4 + i;
    if (ri >= n) continue;
    float* dst = Lb + (int64_t)ri * p.ld;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int cj = j0 + tx * 4 + j;
      if (cj < col1 && cj <= ri) dst[cj] -= acc[i][j];
    }
  }
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

int launch_panel(const Batch& p, int k0, const float* winv, hipStream_t s) {
  const int rows = p.n - (k0 + kNB);
  if (rows <= 0) return 0;
  hipLaunchKernelGGL(k_panel, dim3((unsigned)((rows + 63) / 64), (unsigned)p.batch), dim3(256), 0, s, p,
                     k0, winv);
  return launch_status("k_panel");
}

int launch_syrk(const Batch& p, int row0, int col0, int col1, int k0, int kl, hipStream_t s) {
  (void)row0;  // the update region starts at its first column (row0 == col0)
  if (col1 <= col0 || kl <= 0) return 0;
  const int TI = (p.n - col0 + 63) / 64;
  const int TJ = (col1 - col0 + 63) / 64;
  // Tiles: sum over tj < TJ of (TI - tj).
  const int64_t tiles = (int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2;
  hipLaunchKernelGGL(k_syrk, dim3((unsigned)tiles, (unsigned)p.batch), dim3(256), 0, s, p, col0, col1,
                     k0, kl, TI);
  return launch_status("k_syrk");
}

}  // namespace chol
