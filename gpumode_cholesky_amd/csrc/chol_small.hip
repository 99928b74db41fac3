// chol_small.hip — whole-matrix Cholesky for n <= 64, one matrix per
// 64-lane wavefront slice, held in VGPRs.
//
// Replaces the reference's K1 chol32_register_rank2 (submission.py:100-171)
// and K2 cholesky64_rank2 (submission.py:204-279), which give one 32-lane
// warp (n=32) or two warps (n=64) a matrix and broadcast pivot columns with
// __shfl_sync.  Here a 64-lane wavefront holds 64/NP matrices of padded order
// NP in {16, 32, 64}: lane l owns row (l % NP) of matrix (l / NP).  Orders
// n < NP are padded with an identity block, whose factor is the identity, so
// the unrolled loop needs no masking by n.
//
// Step k (right-looking, fully unrolled so every register index is static):
//   every lane writes its column-k entry a[k] to a per-wave LDS column;
//   every lane reads back the pivot c[k] and the entries c[k+1..NP) with
//   broadcast ds_read_b128 (all lanes of a matrix read one address);
//   lanes below the pivot fold the rank-1 update a[j] -= (a[k]/c[k]) * c[j];
//   a[k] becomes L[r][k] = a[k]/sqrt(c[k]) (r > k), sqrt(c[k]) (r == k), or 0.
// Only the lower triangle of A is loaded; the upper part of each lane's row
// never feeds another lane, and is zeroed as the loop passes it.
#include "chol_common.h"

namespace chol {
namespace {

constexpr int kWavesPerBlock = 4;

template <int NP, bool VEC>
__global__ __launch_bounds__(kWavesPerBlock* kWave) void k_potrf_small(Batch p) {
  constexpr int MPW = kWave / NP;  // matrices per wavefront
  constexpr int PITCH = NP + 4;    // padded column stride (bank spread)
  __shared__ __align__(16) float colbuf[kWavesPerBlock][MPW * PITCH];

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int m = lane / NP;
  const int r = lane - m * NP;
  const int64_t b = ((int64_t)blockIdx.x * kWavesPerBlock + wave) * MPW + m;
  const bool live = b < p.batch;
  const int n = p.n;

  float a[NP];
  const float* src = p.A + (live ? b : 0) * p.bs + (int64_t)(r < n ? r : 0) * p.ld;
  if constexpr (VEC) {
    // n == NP and every row 16-byte aligned: float4 loads of the lower part.
#pragma unroll
    for (int j = 0; j < NP; j += 4) {
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (live && j <= r) v = *reinterpret_cast<const float4*>(src + j);
      a[j] = v.x;
      a[j + 1] = v.y;
      a[j + 2] = v.z;
      a[j + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      float v;
      if (r < n)
        v = (live && j <= r) ? src[j] : 0.f;
      else
        v = (j == r) ? 1.f : 0.f;
      a[j] = v;
    }
  }

  float* cb = &colbuf[wave][m * PITCH];
  const int32_t bad = reg_factor<NP>(a, r, cb, true);

  if (!live) return;
  if (p.info && r == 0) p.info[b] = (bad > n) ? 0 : bad;
  if (r >= n) return;
  float* dst = p.L + b * p.bs + (int64_t)r * p.ld;
  if constexpr (VEC) {
#pragma unroll
    for (int j = 0; j < NP; j += 4)
      *reinterpret_cast<float4*>(dst + j) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (j < n) dst[j] = a[j];
  }
}

template <int NP, bool VEC>
int launch_np(const Batch& p, hipStream_t s) {
  constexpr int per_block = kWavesPerBlock * (kWave / NP);
  const int64_t blocks = (p.batch + per_block - 1) / per_block;
  hipLaunchKernelGGL((k_potrf_small<NP, VEC>), dim3((unsigned)blocks),
                     dim3(kWavesPerBlock * kWave), 0, s, p);
  return launch_status("k_potrf_small");
}

}  // namespace

int launch_small(const Batch& p, hipStream_t s) {
  const bool aligned = ((reinterpret_cast<uintptr_t>(p.A) | reinterpret_cast<uintptr_t>(p.L)) & 15) == 0 &&
                       (p.ld % 4) == 0 && (p.bs % 4) == 0;
  if (p.n <= 16) return launch_np<16, false>(p, s);
  if (p.n <= 32) {
    if (p.n == 32 && aligned) return launch_np<32, true>(p, s);
    return launch_np<32, false>(p, s);
  }
  if (p.n == 64 && aligned) return launch_np<64, true>(p, s);
  return launch_np<64, false>(p, s);
}

}  // namespace chol
