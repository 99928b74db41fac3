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
#include "chol_block.h"

namespace chol {
namespace {

constexpr int kWavesPerBlock = 4;

// Row r of matrix b into a[] (lower part; rows past the matrix's order are
// the identity, so a padded NP x NP block factors as diag(L, I)).
template <int NP>
__device__ __forceinline__ void slot_rows(const Batch& p, int64_t b, bool live, int r, float (&a)[NP]) {
  const Mat M = mat_at(p, live ? b : 0);
  const int n = M.n;
  const float* src = M.A + (int64_t)(r < n ? r : 0) * M.lda;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    float v;
    if (r < n)
      v = (live && j <= r) ? gld_nt(src + j) : 0.f;
    else
      v = (j == r) ? 1.f : 0.f;
    a[j] = v;
  }
}

template <int NP>
__device__ __forceinline__ void slot_store(const Batch& p, int64_t b, bool live, int r, const float (&a)[NP]) {
  if (!live) return;
  const Mat M = mat_at(p, b);
  if (r >= M.n) return;
  float* dst = M.L + (int64_t)r * M.ld;
#pragma unroll
  for (int j = 0; j < NP; ++j)
    if (j < M.n) gst_nt(dst + j, a[j]);
}

template <int NP, bool VEC>
__global__ __launch_bounds__(kWavesPerBlock* kWave) void k_potrf_small(Batch p) {
  constexpr int MPW = kWave / NP;  // matrices per wavefront
  constexpr int PITCH = NP + 4;    // padded row / column stride (bank spread)
  __shared__ __align__(16) float colbuf[kWavesPerBlock][MPW * PITCH];
  // VEC, and always for NP >= 32: per-wave staging of its MPW matrices, one
  // padded row per lane.
  constexpr bool STAGED = VEC || NP >= 32;
  __shared__ __align__(16) float stage[STAGED ? kWavesPerBlock : 1][STAGED ? kWave * PITCH : 4];

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int m = lane / NP;
  const int r = lane - m * NP;
  const int64_t b0 = ((int64_t)blockIdx.x * kWavesPerBlock + wave) * MPW;  // first matrix of the wave
  const int64_t b = b0 + m;
  const bool live = b < p.batch;
  const int n = p.n;

  float a[NP];
  if constexpr (VEC) {
    // n == NP, contiguous batch (ld == n, stride n*n), 16-byte aligned: the
    // wave's MPW matrices are one contiguous run of 64*NP floats.  Lanes read
    // consecutive float4s (coalesced), skipping chunks above the diagonal,
    // and the transpose to one row per lane goes through LDS.
    float* st = stage[wave];
    const float* src = p.A + b0 * (int64_t)NP * NP;
#pragma unroll
    for (int u = 0; u < NP / 4; ++u) {
      const int e = (u * kWave + lane) * 4;  // element within the run
      const int row = e / NP;                // 0 .. 63 (matrix-major rows)
      const int col = e % NP;
      const int rr = row % NP;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (b0 + row / NP < p.batch && col <= rr) v = gld4_nt(src + e);
      *reinterpret_cast<float4*>(st + row * PITCH + col) = v;
    }
    wave_lds_sync();
    if constexpr (NP < 32) {
#pragma unroll
      for (int j = 0; j < NP; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(st + lane * PITCH + j);
        a[j] = v.x;
        a[j + 1] = v.y;
        a[j + 2] = v.z;
        a[j + 3] = v.w;
      }
    }
  } else {
    slot_rows<NP>(p, b, live, r, a);
    if constexpr (NP >= 32) {
      float* st = stage[wave];
#pragma unroll
      for (int j = 0; j < NP; j += 4)
        *reinterpret_cast<float4*>(st + lane * PITCH + j) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
      wave_lds_sync();
    }
  }

  int32_t bad;
  if constexpr (NP >= 32) {
    // staged in LDS, 16-column register panels with MFMA trailing updates
    // (the factor overwrites the staged rows): one matrix per wave for
    // NP = 64, two for NP = 32
    if constexpr (NP == 64)
      bad = factor64<16, PITCH>(stage[wave], 0, lane);
    else
      bad = factor32_pair<PITCH>(stage[wave], colbuf[wave], lane);
    wave_lds_sync();
    if constexpr (!VEC) {
#pragma unroll
      for (int j = 0; j < NP; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(stage[wave] + lane * PITCH + j);
        a[j] = v.x;
        a[j + 1] = v.y;
        a[j + 2] = v.z;
        a[j + 3] = v.w;
      }
    }
  } else {
    float* cb = &colbuf[wave][m * PITCH];
    bad = reg_factor<NP>(a, r, cb, true);
  }
  if (live && p.info && r == 0) p.info[VEC ? b : mat_at(p, b).slot] = bad;

  if constexpr (VEC) {
    // Rows back through LDS, then coalesced float4 stores (zeros above the
    // diagonal included).
    float* st = stage[wave];
    if constexpr (NP < 32) {
#pragma unroll
      for (int j = 0; j < NP; j += 4)
        *reinterpret_cast<float4*>(st + lane * PITCH + j) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
      wave_lds_sync();
    }
    float* dst = p.L + b0 * (int64_t)NP * NP;
#pragma unroll
    for (int u = 0; u < NP / 4; ++u) {
      const int e = (u * kWave + lane) * 4;
      const int row = e / NP;
      const int col = e % NP;
      if (b0 + row / NP < p.batch)
        gst4_nt(dst + e, *reinterpret_cast<const float4*>(st + row * PITCH + col));
    }
  } else {
    slot_store<NP>(p, b, live, r, a);
  }
}

// Persistent form for ragged batches: wavefronts pop slots of MPW matrices
// from a work queue (an atomic counter) over the sorted range
// [first, first + count); every wave leaves when the queue is exhausted.
template <int NP>
__global__ __launch_bounds__(kWavesPerBlock* kWave) void k_potrf_small_queue(Batch p, int64_t first, int64_t count,
                                                                            int* queue) {
  constexpr int MPW = kWave / NP;
  constexpr int PITCH = NP + 4;
  __shared__ __align__(16) float colbuf[kWavesPerBlock][MPW * PITCH];
  // NP >= 32: the same LDS-staged blocked factorisation as k_potrf_small
  __shared__ __align__(16) float stage[NP >= 32 ? kWavesPerBlock : 1][NP >= 32 ? kWave * PITCH : 4];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int m = lane / NP;
  const int r = lane - m * NP;
  float* cb = &colbuf[wave][m * PITCH];
  for (;;) {
    int slot = 0;
    if (lane == 0) slot = atomicAdd(queue, 1);
    slot = __builtin_amdgcn_readfirstlane(slot);
    const int64_t idx = (int64_t)slot * MPW;
    if (idx >= count) break;
    const int64_t b = first + idx + m;
    const bool live = idx + m < count;
    float a[NP];
    slot_rows<NP>(p, b, live, r, a);
    int32_t bad;
    if constexpr (NP >= 32) {
      float* st = stage[wave];
#pragma unroll
      for (int j = 0; j < NP; j += 4)
        *reinterpret_cast<float4*>(st + lane * PITCH + j) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
      wave_lds_sync();
      if constexpr (NP == 64)
        bad = factor64<16, PITCH>(st, 0, lane);
      else
        bad = factor32_pair<PITCH>(st, colbuf[wave], lane);
      wave_lds_sync();
#pragma unroll
      for (int j = 0; j < NP; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(st + lane * PITCH + j);
        a[j] = v.x;
        a[j + 1] = v.y;
        a[j + 2] = v.z;
        a[j + 3] = v.w;
      }
      wave_lds_sync();  // the next slot's rows overwrite the stage
    } else {
      bad = reg_factor<NP>(a, r, cb, true);
    }
    if (live && p.info && r == 0) p.info[mat_at(p, b).slot] = bad;
    slot_store<NP>(p, b, live, r, a);
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
  // The coalesced form needs n == NP and a contiguous, 16-byte aligned batch.
  const bool aligned = ((reinterpret_cast<uintptr_t>(p.A) | reinterpret_cast<uintptr_t>(p.L)) & 15) == 0 &&
                       p.ld == p.n && (p.batch == 1 || p.bs == (int64_t)p.n * p.n);
  if (p.n <= 16) return launch_np<16, false>(p, s);
  if (p.n <= 32) {
    if (p.n == 32 && aligned) return launch_np<32, true>(p, s);
    return launch_np<32, false>(p, s);
  }
  if (p.n == 64 && aligned) return launch_np<64, true>(p, s);
  return launch_np<64, false>(p, s);
}

int launch_small_queue(const Batch& p, int64_t first, int64_t count, int np, int* queue, int grid,
                       hipStream_t s) {
  if (count <= 0) return 0;
  const dim3 g((unsigned)grid), blk(kWavesPerBlock * kWave);
  if (np <= 16)
    hipLaunchKernelGGL((k_potrf_small_queue<16>), g, blk, 0, s, p, first, count, queue);
  else if (np <= 32)
    hipLaunchKernelGGL((k_potrf_small_queue<32>), g, blk, 0, s, p, first, count, queue);
  else
    hipLaunchKernelGGL((k_potrf_small_queue<64>), g, blk, 0, s, p, first, count, queue);
  return launch_status("k_potrf_small_queue");
}

}  // namespace chol
