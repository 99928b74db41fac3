// chol_lds.hip — whole-matrix Cholesky for 64 < n <= 256 in one workgroup,
// the matrix resident in LDS for the whole factorisation.
//
// Storage: the lower triangle, row-packed, each row padded to a multiple of
// four floats so every row starts 16-byte aligned.  Row r occupies
// ceil((r+1)/4)*4 floats; order 256 takes 33,280 floats (130 KiB) of the
// 160 KiB LDS of a gfx950 CU, order 128 takes 33 KiB.  Orders n < N are padded
// with an identity block (its factor is the identity).
//
// Per 16-wide panel p (one s_barrier between phases):
//   1. wavefront 0, lanes 0..15: factor the 16x16 diagonal block in
//      registers (pivot columns broadcast with v_readlane, reg_factor_rl)
//      and form its inverse W = L11^-1;
//   2. all 4 wavefronts: L21 = A21 W^T, one 16-row block per MFMA tile
//      (v_mfma_f32_16x16x4_f32, K = 16);
//   3. all 4 wavefronts: trailing update C -= L21 L21^T over the lower 16x16
//      tiles, read-modify-write in LDS.
// The input is read from HBM once and the factor written once.
//
// Reference counterparts: K3 cholesky128_block16 (submission.py:316-411, a
// blocked-16 factorisation of a 128x129 shared-memory tile) and K5
// cholesky256_wmma16 (submission.py:629-1077, packed lower 16x16 tiles with a
// tensor-core trailing update and 174 KB of shared memory, which would not
// fit a gfx950 CU).
#include "chol_common.h"

namespace chol {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kPB = 16;  // panel width
constexpr int kThreads = 256;

// Offset (floats) of packed row i: 4*sum_{r<i} (r/4 + 1).
__host__ __device__ constexpr int row_off(int i) {
  const int q = i >> 2, s = i & 3;
  return 4 * i + 8 * q * (q - 1) + 4 * s * q;
}

template <int N>
__device__ __forceinline__ void lds_factor(const Batch& p, int64_t b) {
  constexpr int kFloats = row_off(N);
  constexpr int WP = kPB + 4;
  __shared__ __align__(16) float S[kFloats];
  __shared__ __align__(16) float Wt[kPB * WP];  // W = L11^-1, row-major
  __shared__ int bad_s;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const Mat M = mat_at(p, b);
  const int n = M.n;
  const float* A = M.A;

  // ---- load the lower triangle (rows >= n: identity) ----
  for (int i = wave; i < N; i += kThreads / 64) {
    float* row = S + row_off(i);
    const int len = ((i >> 2) + 1) * 4;
    for (int j = lane; j < len; j += 64) {
      float v;
      if (i < n)
        v = (j <= i) ? gld_nt(A + (int64_t)i * M.lda + j) : 0.f;
      else
        v = (j == i) ? 1.f : 0.f;
      row[j] = v;
    }
  }
  if (tid == 0) bad_s = 0;
  __syncthreads();

  const int li = lane & 15;
  const int lg = lane >> 4;

  for (int p0 = 0; p0 < N; p0 += kPB) {
    // ---- 1. diagonal block: wave 0, lane r < 16 holds row p0 + r ----
    if (wave == 0) {
      const int r = li;
      float a[kPB];
      // Packed row p0+r holds columns up to p0+r (padded to a multiple of
      // 4): only the float4 chunks starting at j <= r belong to it.
      const float* rowp = S + row_off(p0 + r) + p0;
#pragma unroll
      for (int j = 0; j < kPB; j += 4) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j <= r) v = *reinterpret_cast<const float4*>(rowp + j);
        a[j] = v.x;
        a[j + 1] = v.y;
        a[j + 2] = v.z;
        a[j + 3] = v.w;
      }
      const int bad = reg_factor_rl<kPB>(a, r, 0);
      if (lane < kPB) {
        // Chunks j <= r only; their entries past r are zeros that land in
        // the row's own padding.
        float* rw = S + row_off(p0 + r) + p0;
#pragma unroll
        for (int j = 0; j < kPB; j += 4)
          if (j <= r) *reinterpret_cast<float4*>(rw + j) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
      }
      if (lane == 0 && bad != 0 && bad_s == 0 && p0 + bad <= n) bad_s = p0 + bad;
      wave_lds_sync();
      // Inverse: lane c < 16 forward-substitutes column c of W, reading the
      // rows of L11 just written (broadcast reads).
      float w[kPB];
      const int c = li;
#pragma unroll
      for (int i = 0; i < kPB; ++i) {
        const float* Li = S + row_off(p0 + i) + p0;
        float s = (i == c) ? 1.f : 0.f;
#pragma unroll
        for (int q = 0; q < i; q += 4) {
          const float4 v = *reinterpret_cast<const float4*>(Li + q);
          s = fmaf(-v.x, w[q], s);
          if (q + 1 < i) s = fmaf(-v.y, w[q + 1], s);
          if (q + 2 < i) s = fmaf(-v.z, w[q + 2], s);
          if (q + 3 < i) s = fmaf(-v.w, w[q + 3], s);
        }
        w[i] = s * __builtin_amdgcn_rcpf(Li[i]);
      }
      if (lane < kPB) {
#pragma unroll
        for (int i = 0; i < kPB; ++i) Wt[i * WP + c] = w[i];
      }
    }
    __syncthreads();

    const int r0 = p0 + kPB;  // first row below the diagonal block
    const int T = (N - r0) / kPB;
    if (T > 0) {
      // ---- 2. panel: L21 = A21 W^T, one 16-row block per wave step ----
      const float4 wf = *reinterpret_cast<const float4*>(&Wt[li * WP + 4 * lg]);  // B[k][j] = W[j][k]
      for (int blk = wave; blk < T; blk += kThreads / 64) {
        const int i0 = r0 + blk * kPB;
        const float4 af = *reinterpret_cast<const float4*>(S + row_off(i0 + li) + p0 + 4 * lg);
        floatx4 d = {0.f, 0.f, 0.f, 0.f};
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.x, wf.x, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.y, wf.y, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.z, wf.z, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.w, wf.w, d, 0, 0, 0);
        // All lanes of this wave have read the block before any writes it.
        wave_lds_sync();
#pragma unroll
        for (int v = 0; v < 4; ++v) S[row_off(i0 + 4 * lg + v) + p0 + li] = d[v];
      }
      __syncthreads();

      // ---- 3. trailing update over lower 16x16 tiles (I >= J) ----
      const int tiles = T * (T + 1) / 2;
      for (int t = wave; t < tiles; t += kThreads / 64) {
        // column-major enumeration: column J holds T - J tiles
        int J = 0, rem = t;
        while (rem >= T - J) {
          rem -= T - J;
          ++J;
        }
        const int I = J + rem;
        const int i0 = r0 + I * kPB;
        const int j0 = r0 + J * kPB;
        const float4 af = *reinterpret_cast<const float4*>(S + row_off(i0 + li) + p0 + 4 * lg);
        const float4 bf = *reinterpret_cast<const float4*>(S + row_off(j0 + li) + p0 + 4 * lg);
        floatx4 d = {0.f, 0.f, 0.f, 0.f};
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.x, bf.x, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.y, bf.y, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.z, bf.z, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(af.w, bf.w, d, 0, 0, 0);
        const int j = j0 + li;
        if (I != J) {
          // below the diagonal tile: every element, old values loaded together
          float o[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) o[v] = S[row_off(i0 + 4 * lg + v) + j];
#pragma unroll
          for (int v = 0; v < 4; ++v) S[row_off(i0 + 4 * lg + v) + j] = o[v] - d[v];
        } else {
          // diagonal tile: only j <= i (in the packed rows, a row's
          // elements past its diagonal group belong to the next rows)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = i0 + 4 * lg + v;
            if (j <= i) S[row_off(i) + j] -= d[v];
          }
        }
      }
    }
    __syncthreads();
  }

  // ---- write L (zeros above the diagonal) ----
  if (p.info && tid == 0) p.info[M.slot] = bad_s;
  float* Lout = M.L;
  for (int i = wave; i < n; i += kThreads / 64) {
    const float* row = S + row_off(i);
    float* dst = Lout + (int64_t)i * M.ld;
    for (int j = lane; j < n; j += 64) gst_nt(dst + j, (j <= i) ? row[j] : 0.f);
  }
}

template <int N>
__global__ __launch_bounds__(kThreads) void k_potrf_lds(Batch p) {
  lds_factor<N>(p, blockIdx.x);
}

}  // namespace

int launch_lds(const Batch& p, hipStream_t s) {
  if (p.n <= 128)
    hipLaunchKernelGGL(k_potrf_lds<128>, dim3((unsigned)p.batch), dim3(kThreads), 0, s, p);
  else
    hipLaunchKernelGGL(k_potrf_lds<256>, dim3((unsigned)p.batch), dim3(kThreads), 0, s, p);
  return launch_status("k_potrf_lds");
}

}  // namespace chol
