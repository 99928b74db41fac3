// chol_diag.hip — factor one 64x64 diagonal block per matrix and form its
// inverse, the serial link of the blocked path (one launch per 64 columns).
//
// One workgroup of 4 wavefronts per matrix:
//   1. all waves stage the (already updated) block into LDS with coalesced
//      row loads; rows past n are the identity;
//   2. wavefront 0 factors it in registers, one row per lane (reg_factor);
//   3. the inverse W = L11^-1 by 16x16 blocks: the 4 wavefronts invert the
//      four diagonal 16x16 blocks at once (lane = column, forward
//      substitution); then block rows 1..3 in turn,
//         W_ij = -W_ii (sum_{k=j}^{i-1} L_ik W_kj),   j < i,
//      one 16x16 block per wavefront on v_mfma_f32_16x16x4_f32;
//   4. all waves store L11 and W with coalesced row stores.
//
// Reference counterpart: e62_diag128 (submission.py:4694-4997), which
// factors a 128x128 block and builds its inverse in one CTA with warp 0
// forming 32x32 inverses while the other warps update.
#include "chol_common.h"

namespace chol {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int NB = kNB;
constexpr int P = NB + 4;

// C(16x16) += X Y for 16x16 row-major LDS blocks (pitch P): X rows as float4,
// Y columns as scalars.  Lane (li, lg) ends with C[4lg+v][li] in acc[v].
__device__ __forceinline__ floatx4 mma16(const float* X, const float* Y, floatx4 acc, int li, int lg) {
  const float4 xa = *reinterpret_cast<const float4*>(X + li * P + 4 * lg);
  const float y0 = Y[(4 * lg + 0) * P + li];
  const float y1 = Y[(4 * lg + 1) * P + li];
  const float y2 = Y[(4 * lg + 2) * P + li];
  const float y3 = Y[(4 * lg + 3) * P + li];
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, y0, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, y1, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, y2, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, y3, acc, 0, 0, 0);
  return acc;
}

__global__ __launch_bounds__(256) void k_diag_factor(Batch p, int k0, float* __restrict__ winv) {
  __shared__ __align__(16) float Ls[NB * P];
  __shared__ __align__(16) float Ws[NB * P];
  __shared__ __align__(16) float Ts[4][16 * P];  // per-wave 16x16 staging
  __shared__ __align__(16) float cb[NB];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int li = lane & 15;
  const int lg = lane >> 4;
  const int64_t b = blockIdx.x;
  const Mat M = mat_at(p, b);
  const int nb = min(NB, M.n - k0);
  if (nb <= 0) return;  // (ragged) this matrix has no block at k0
  float* Lb = M.L + (int64_t)k0 * M.ld + k0;
  const bool vec = ((M.ld & 3) == 0) && ((reinterpret_cast<uintptr_t>(M.L) & 15) == 0);

  // ---- 1. stage the block: thread covers rows tid/16 + 16u, cols 4*(tid%16) ----
  const int c4 = (tid & 15) * 4;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = (tid >> 4) + 16 * u;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (r < nb) {
      if (c4 <= r) {
        const float* src = Lb + (int64_t)r * M.ld + c4;
        if (vec && c4 + 3 < nb) {
          v = *reinterpret_cast<const float4*>(src);
        } else {
          v.x = src[0];
          if (c4 + 1 < nb) v.y = src[1];
          if (c4 + 2 < nb) v.z = src[2];
          if (c4 + 3 < nb) v.w = src[3];
        }
        if (c4 + 1 > r) v.y = 0.f;
        if (c4 + 2 > r) v.z = 0.f;
        if (c4 + 3 > r) v.w = 0.f;
      }
    } else {
      v.x = (c4 == r) ? 1.f : 0.f;
      v.y = (c4 + 1 == r) ? 1.f : 0.f;
      v.z = (c4 + 2 == r) ? 1.f : 0.f;
      v.w = (c4 + 3 == r) ? 1.f : 0.f;
    }
    *reinterpret_cast<float4*>(&Ls[r * P + c4]) = v;
    *reinterpret_cast<float4*>(&Ws[r * P + c4]) = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __syncthreads();

  // ---- 2. wavefront 0 factors in registers ----
  if (wave == 0) {
    float a[NB];
#pragma unroll
    for (int j = 0; j < NB; j += 4) {
      const float4 v = *reinterpret_cast<const float4*>(&Ls[lane * P + j]);
      a[j] = v.x;
      a[j + 1] = v.y;
      a[j + 2] = v.z;
      a[j + 3] = v.w;
    }
    const int bad = reg_factor<NB>(a, lane, cb, true);
#pragma unroll
    for (int j = 0; j < NB; j += 4)
      *reinterpret_cast<float4*>(&Ls[lane * P + j]) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
    if (p.info && lane == 0 && bad != 0 && bad <= nb && p.info[M.slot] == 0) p.info[M.slot] = k0 + bad;
  }
  __syncthreads();

  // ---- 3a. diagonal 16x16 inverses, wave w -> block w, lane c < 16 ----
  {
    const int d0 = 16 * wave;
    const int c = li;
    float w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float* Li = &Ls[(d0 + i) * P + d0];
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
    if (lane < 16) {
#pragma unroll
      for (int i = 0; i < 16; ++i) Ws[(d0 + i) * P + d0 + c] = w[i];
    }
  }
  __syncthreads();

  // ---- 3b. off-diagonal blocks, block row i = 1..3, wave j < i ----
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    if (wave < i) {
      const int j = wave;
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int k = j; k < i; ++k)
        acc = mma16(&Ls[(16 * i) * P + 16 * k], &Ws[(16 * k) * P + 16 * j], acc, li, lg);
      float* T = Ts[wave];
#pragma unroll
      for (int v = 0; v < 4; ++v) T[(4 * lg + v) * P + li] = acc[v];
      wave_lds_sync();
      floatx4 r = {0.f, 0.f, 0.f, 0.f};
      r = mma16(&Ws[(16 * i) * P + 16 * i], T, r, li, lg);
#pragma unroll
      for (int v = 0; v < 4; ++v) Ws[(16 * i + 4 * lg + v) * P + 16 * j + li] = -r[v];
    }
    __syncthreads();
  }

  // ---- 4. store L11 (rows < nb, columns < nb) and W ----
  float* Wb = winv + b * (NB * NB);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = (tid >> 4) + 16 * u;
    const float4 w4 = *reinterpret_cast<const float4*>(&Ws[r * P + c4]);
    *reinterpret_cast<float4*>(&Wb[r * NB + c4]) = w4;
    if (r < nb && c4 <= r) {
      const float4 v = *reinterpret_cast<const float4*>(&Ls[r * P + c4]);
      float* dst = Lb + (int64_t)r * M.ld + c4;
      if (vec && c4 + 3 < nb) {
        *reinterpret_cast<float4*>(dst) = v;
      } else {
        dst[0] = v.x;
        if (c4 + 1 < nb) dst[1] = v.y;
        if (c4 + 2 < nb) dst[2] = v.z;
        if (c4 + 3 < nb) dst[3] = v.w;
      }
    }
  }
}

}  // namespace

int launch_diag_factor(const Batch& p, int k0, float* winv, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_factor, dim3((unsigned)p.batch), dim3(256), 0, s, p, k0, winv);
  return launch_status("k_diag_factor");
}

}  // namespace chol
