// chol_diag.hip — factor one 128x128 diagonal block per matrix and form its
// inverse: the serial link of the blocked path (one launch per 128 columns).
//
// One workgroup of 4 wavefronts per matrix, the block resident in LDS:
//   1. all waves stage the (already updated) block with coalesced row loads;
//      rows past n are the identity;
//   2. wavefront 0 factors the top-left 64x64 block D1 (factor64: 16-column
//      panels in registers, lane = row, pivots broadcast with v_readlane,
//      MFMA updates between panels);
//   3. W1 = D1^-1 by 16x16 blocks: the 4 wavefronts invert the four diagonal
//      16x16 blocks at once (lane = column, forward substitution), then
//      block rows 1..3 in turn, W_ij = -W_ii sum_{k=j}^{i-1} L_ik W_kj;
//   4. L21 = A21 W1^T and D2 -= L21 L21^T on v_mfma_f32_16x16x4_f32;
//   5. wavefront 0 factors D2 while waves 1-3 form T1 = L21 W1 and store the
//      parts that are final (L rows 0..63, L21, W1); W2 = D2^-1 as in 3;
//   6. W21 = -W2 T1, so W = L11^-1 is complete;
//   7. all waves store the rest of L11 and W with coalesced row stores.
// The two 64x64 factorisations (~15k cycles each, s_memtime probe) are the
// serial part; staging, inverses and the MFMA steps add about as much again.
//
// Reference counterpart: e62_diag128 (submission.py:4694-4997), which
// factors a 128x128 block and builds its inverse in one CTA.
#include "chol_common.h"
#include "chol_block.h"

namespace chol {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int NB = kNB;     // 128
constexpr int H = NB / 2;   // 64
constexpr int P = NB + 4;   // LDS pitch of the block buffers
constexpr int TP = 16 + 4;  // pitch of the per-wave 16x16 staging

// C(16x16) += X Y, X rows read as float4 (pitch px), Y columns read as
// scalars (pitch py), K = 16.  Lane (li, lg) holds C[4lg+v][li] in acc[v].
__device__ __forceinline__ floatx4 mma_nn(const float* X, int px, const float* Y, int py, floatx4 acc, int li,
                                          int lg) {
  const float4 xa = *reinterpret_cast<const float4*>(X + li * px + 4 * lg);
  const float y0 = Y[(4 * lg + 0) * py + li];
  const float y1 = Y[(4 * lg + 1) * py + li];
  const float y2 = Y[(4 * lg + 2) * py + li];
  const float y3 = Y[(4 * lg + 3) * py + li];
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, y0, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, y1, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, y2, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, y3, acc, 0, 0, 0);
  return acc;
}

// C(16x16) += X Y^T, both X and Y rows read as float4 (pitch P), K = 16.
__device__ __forceinline__ floatx4 mma_nt(const float* X, const float* Y, floatx4 acc, int li, int lg) {
  const float4 xa = *reinterpret_cast<const float4*>(X + li * P + 4 * lg);
  const float4 yb = *reinterpret_cast<const float4*>(Y + li * P + 4 * lg);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, yb.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, yb.y, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, yb.z, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, yb.w, acc, 0, 0, 0);
  return acc;
}

// Ws(d..d+64, d..d+64) = inverse of the lower-triangular Ls(d.., d..) block,
// all 4 wavefronts.  Ends with a workgroup barrier.
__device__ __forceinline__ void inverse64(const float* Ls, float* Ws, float (*Ts)[16 * TP], int d, int wave,
                                          int lane, int li, int lg) {
  {  // four diagonal 16x16 inverses at once, wave w -> block w
    const int d0 = d + 16 * wave;
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
#pragma unroll
  for (int i = 1; i < 4; ++i) {  // block row i, wave j < i
    if (wave < i) {
      const int j = wave;
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int k = j; k < i; ++k)
        acc = mma_nn(&Ls[(d + 16 * i) * P + d + 16 * k], P, &Ws[(d + 16 * k) * P + d + 16 * j], P, acc, li, lg);
      float* T = Ts[wave];
#pragma unroll
      for (int v = 0; v < 4; ++v) T[(4 * lg + v) * TP + li] = acc[v];
      wave_lds_sync();
      floatx4 r = {0.f, 0.f, 0.f, 0.f};
      r = mma_nn(&Ws[(d + 16 * i) * P + d + 16 * i], P, T, TP, r, li, lg);
#pragma unroll
      for (int v = 0; v < 4; ++v) Ws[(d + 16 * i + 4 * lg + v) * P + d + 16 * j + li] = -r[v];
    }
    __syncthreads();
  }
}

// LDS of the kernel: the staging tiles first, then the block and its
// inverse (Ls starts below 64 KiB, so most of its accesses encode their
// offsets as immediates).
struct __align__(16) DiagLds {
  float Ts[4][16 * TP];  // per-wave 16x16 staging
  float Ls[NB * P];
  float Ws[NB * P];
};

template <bool FROM_A, bool LAST, bool FULL>
__global__ __launch_bounds__(256) void k_diag_factor(Batch p, int k0, float* __restrict__ winv) {
  __shared__ DiagLds sm;
  float(*const Ts)[16 * TP] = sm.Ts;
  float* const Ls = sm.Ls;
  float* const Ws = sm.Ws;
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
  // FROM_A: the first block, staged from A (nothing has written L yet).
  // FULL (the whole matrix is this one block): every row is stored in full,
  // zeros above the diagonal included, so nothing else writes L.
  const float* Sb = FROM_A ? M.A + (int64_t)k0 * M.ld + k0 : Lb;
  const bool vec = ((M.ld & 3) == 0) && ((reinterpret_cast<uintptr_t>(M.L) & 15) == 0) &&
                   (!FROM_A || (reinterpret_cast<uintptr_t>(M.A) & 15) == 0);

  // ---- 1. stage: thread covers rows tid/32 + 8u, columns 4*(tid%32).
  //         All 16 loads are issued before the first LDS store, so the
  //         block costs one memory latency, not one per few rows. ----
  const int c4 = (tid & 31) * 4;
  float4 stg[NB / 8];
  if (vec && nb == NB) {
    // Full aligned block: every load unconditional (chunks above the
    // diagonal read column 0 of their row and are zeroed afterwards).
#pragma unroll
    for (int u = 0; u < NB / 8; ++u) {
      const int r = (tid >> 5) + 8 * u;
      stg[u] = gld4_nt(Sb + (int64_t)r * M.ld + (c4 <= r ? c4 : 0));
    }
#pragma unroll
    for (int u = 0; u < NB / 8; ++u) {
      const int r = (tid >> 5) + 8 * u;
      float4 v = stg[u];
      if (c4 > r) v.x = 0.f;
      if (c4 + 1 > r) v.y = 0.f;
      if (c4 + 2 > r) v.z = 0.f;
      if (c4 + 3 > r) v.w = 0.f;
      stg[u] = v;
    }
  } else {
#pragma unroll
  for (int u = 0; u < NB / 8; ++u) {
    const int r = (tid >> 5) + 8 * u;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (r < nb) {
      if (c4 <= r) {
        const float* src = Sb + (int64_t)r * M.ld + c4;
        if (vec && c4 + 3 < nb) {
          v = gld4(src);
        } else {
          v.x = gld(src);
          if (c4 + 1 < nb) v.y = gld(src + 1);
          if (c4 + 2 < nb) v.z = gld(src + 2);
          if (c4 + 3 < nb) v.w = gld(src + 3);
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
    stg[u] = v;
  }
  }
#pragma unroll
  for (int u = 0; u < NB / 8; ++u) {
    const int r = (tid >> 5) + 8 * u;
    *reinterpret_cast<float4*>(&Ls[r * P + c4]) = stg[u];
    *reinterpret_cast<float4*>(&Ws[r * P + c4]) = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __syncthreads();

  // ---- 2-3. D1 and W1 ----
  int bad = 0;
  if (wave == 0) bad = factor64<16, P>(Ls, 0, lane);
  __syncthreads();
  inverse64(Ls, Ws, Ts, 0, wave, lane, li, lg);

  // ---- 4a. L21 = A21 W1^T: 16 output tiles, 4 per wave, K = 64 ----
  {
    floatx4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int ti = wave, tj = t;  // tile row ti, tile column tj (16x16)
      acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < H; kk += 16)
        acc[t] = mma_nt(&Ls[(H + 16 * ti) * P + kk], &Ws[(16 * tj) * P + kk], acc[t], li, lg);
    }
    __syncthreads();  // every read of A21 done before it is overwritten
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) Ls[(H + 16 * wave + 4 * lg + v) * P + 16 * t + li] = acc[t][v];
  }
  __syncthreads();
  // ---- 4b. D2 -= L21 L21^T on the 10 lower 16x16 tiles ----
  for (int t = wave; t < 10; t += 4) {
    int ti = 0, tj = t;  // t -> (ti >= tj): rows of lengths 1, 2, 3, 4
    while (tj > ti) {
      tj -= ti + 1;
      ++ti;
    }
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < H; kk += 16)
      acc = mma_nt(&Ls[(H + 16 * ti) * P + kk], &Ls[(H + 16 * tj) * P + kk], acc, li, lg);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 16 * ti + 4 * lg + v, j = 16 * tj + li;
      if (ti != tj || j <= i) Ls[(H + i) * P + H + j] -= acc[v];
    }
  }
  __syncthreads();

  // Final stores of one row chunk: W row r (all 128 columns) and the
  // lower-triangle chunk of L row r (rows < nb, columns < nb; the factor is
  // final, so streaming stores).
  float* Wb = winv + b * (NB * NB);
  auto store_w = [&](int r, int cc) {
    gst4(&Wb[r * NB + cc], *reinterpret_cast<const float4*>(&Ws[r * P + cc]));
  };
  auto store_l = [&](int r, int cc) {
    if (r >= nb || cc >= nb || (!FULL && cc > r)) return;
    const float4 v = *reinterpret_cast<const float4*>(&Ls[r * P + cc]);
    float* dst = Lb + (int64_t)r * M.ld + cc;
    if (vec && cc + 3 < nb) {
      gst4_nt(dst, v);
    } else {
      gst_nt(dst, v.x);
      if (cc + 1 < nb) gst_nt(dst + 1, v.y);
      if (cc + 2 < nb) gst_nt(dst + 2, v.z);
      if (cc + 3 < nb) gst_nt(dst + 3, v.w);
    }
  };

  // ---- 5. wavefront 0 factors D2; meanwhile waves 1-3 form T1 = L21 W1
  //         in the unused upper-right 64x64 block of Ls (never stored) and
  //         store what is final: W rows 0..63 (W1, zeros to its right), L
  //         rows 0..63 and L21.  LAST (no panel follows this block): no W,
  //         so neither T1 nor W2 / W21 is formed. ----
  static_assert(!FULL || (FROM_A && LAST), "a full block is the whole matrix");
  if (wave == 0) {
    const int bad2 = factor64<16, P>(Ls, H, lane);
    if (bad == 0 && bad2 != 0) bad = H + bad2;
  } else {
    if constexpr (!LAST) {
      for (int t = wave - 1; t < 16; t += 3) {
        const int ti = t >> 2, tj = t & 3;
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < H; kk += 16)
          acc = mma_nn(&Ls[(H + 16 * ti) * P + kk], P, &Ws[kk * P + 16 * tj], P, acc, li, lg);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ls[(16 * ti + 4 * lg + v) * P + H + 16 * tj + li] = acc[v];
      }
    }
    const int t2 = tid - 64;  // 0..191
    for (int e = t2; e < H * (NB / 4); e += 192) {  // rows 0..63, 32 chunks each
      const int r = e >> 5, cc = (e & 31) * 4;
      if constexpr (!LAST) store_w(r, cc);
      if (cc < H || FULL) store_l(r, cc);
    }
    for (int e = t2; e < H * (H / 4); e += 192) {  // rows 64..127, columns 0..63
      const int r = H + (e >> 4), cc = (e & 15) * 4;
      store_l(r, cc);
    }
  }
  __syncthreads();
  if constexpr (!LAST) {
    inverse64(Ls, Ws, Ts, H, wave, lane, li, lg);  // W2

    // ---- 6. W21 = -W2 T1, so W = L11^-1 is complete ----
    floatx4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < H; kk += 16)
        acc[t] = mma_nn(&Ws[(H + 16 * wave) * P + H + kk], P, &Ls[kk * P + H + 16 * t], P, acc[t], li, lg);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) Ws[(H + 16 * wave + 4 * lg + v) * P + 16 * t + li] = -acc[t][v];
    __syncthreads();
  }
  if (p.info && tid == 0 && bad != 0 && bad <= nb && p.info[M.slot] == 0) p.info[M.slot] = k0 + bad;

  // ---- 7. the rest: W rows 64..127, L rows 64..127 from column 64 on ----
  for (int e = tid; e < H * (NB / 4); e += 256) {
    const int r = H + (e >> 5), cc = (e & 31) * 4;
    if constexpr (!LAST) store_w(r, cc);
    if (cc >= H) store_l(r, cc);
  }
}

}  // namespace

int launch_diag_factor(const Batch& p, int k0, float* winv, hipStream_t s) {
  // The block that reaches column n needs no inverse (no panel follows).
  // p.first (dense, k0 = 0): staged from A; a matrix that is one block is
  // then stored in full.  (Ragged batches: every matrix must end at this
  // block for LAST; they never take p.first.)
  const bool last = k0 + kNB >= p.n;
  const bool from_a = p.first && !p.Ap;
  const dim3 g((unsigned)p.batch), b(256);
  if (from_a && last)
    hipLaunchKernelGGL((k_diag_factor<true, true, true>), g, b, 0, s, p, k0, winv);
  else if (from_a)
    hipLaunchKernelGGL((k_diag_factor<true, false, false>), g, b, 0, s, p, k0, winv);
  else if (last)
    hipLaunchKernelGGL((k_diag_factor<false, true, false>), g, b, 0, s, p, k0, winv);
  else
    hipLaunchKernelGGL((k_diag_factor<false, false, false>), g, b, 0, s, p, k0, winv);
  return launch_status("k_diag_factor");
}

}  // namespace chol
