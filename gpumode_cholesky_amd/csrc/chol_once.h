// chol_once.h — the one-shot (K = 128) MFMA tile of the blocked path and
// its staging helpers, shared by chol_gemm.hip (the SYRK and panel kernels)
// and chol_diag.hip (whose extra workgroups run SYRK tiles on the CUs its
// diagonal blocks leave idle).  The caller provides the tile's LDS.
#pragma once

#include "chol_common.h"

namespace chol {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int KC = 32;        // K chunk staged in LDS
constexpr int KP = KC + 4;    // LDS row pitch (floats)
constexpr int kThreads = 256;

// LOWER: the rows are those of a lower-triangular matrix whose row and
// column indices are the global ones (the inverse W of the panel): chunks
// wholly right of its diagonal hold zeros and are not read.
template <int ROWS, bool VEC, int KW = KC, bool LOWER = false>
__device__ __forceinline__ void load_rows(float4 (&reg)[ROWS * KW / 4 / kThreads], const float* base,
                                          int64_t ld, int row0, int rlim, int k, int tid) {
  // ROWS x KW floats as float4 chunks; thread t takes chunks t, t+256, ...
  constexpr int PER = ROWS * KW / 4 / kThreads;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int c = tid + u * kThreads;
    const int r = c / (KW / 4);
    const int q = (c % (KW / 4)) * 4;
    const int gr = row0 + r;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (gr < rlim && !(LOWER && k + q > gr)) {
      const float* src = base + (int64_t)gr * ld + k + q;
      if constexpr (VEC) {
        v = gld4(src);
      } else {
        v.x = gld(src);
        v.y = gld(src + 1);
        v.z = gld(src + 2);
        v.w = gld(src + 3);
      }
    }
    reg[u] = v;
  }
}

template <int ROWS, int KW = KC, int PITCH = KP>
__device__ __forceinline__ void store_rows(float* __restrict__ lds, const float4 (&reg)[ROWS * KW / 4 / kThreads],
                                           int tid) {
  constexpr int PER = ROWS * KW / 4 / kThreads;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int c = tid + u * kThreads;
    const int r = c / (KW / 4);
    const int q = (c % (KW / 4)) * 4;
    *reinterpret_cast<float4*>(lds + r * PITCH + q) = reg[u];
  }
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int KPH = KC + 8;  // bf16 LDS row pitch (80 bytes)

template <int ROWS, int KW = KC, int PITCH = KPH>
__device__ __forceinline__ void store_rows_split(__bf16* __restrict__ hi, __bf16* __restrict__ lo,
                                                 const float4 (&reg)[ROWS * KW / 4 / kThreads], int tid) {
  constexpr int PER = ROWS * KW / 4 / kThreads;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int c = tid + u * kThreads;
    const int r = c / (KW / 4);
    const int q = (c % (KW / 4)) * 4;
    const float4 v = reg[u];
    uint32_t h01, l01, h23, l23;
    split_bf16_pair(v.x, v.y, h01, l01);
    split_bf16_pair(v.z, v.w, h23, l23);
    uint32_t* ph = reinterpret_cast<uint32_t*>(hi + r * PITCH + q);
    uint32_t* pl = reinterpret_cast<uint32_t*>(lo + r * PITCH + q);
    ph[0] = h01;
    ph[1] = h23;
    pl[0] = l01;
    pl[1] = l23;
  }
}

// ------------------------------------------------------ one-shot K = 128 ----
// Latency form of the per-step launches (K = 128, too few tiles to fill the
// GPU, as for one or a few large matrices): a workgroup requests all of its
// A and B rows (64 KiB for a 64x64 tile) and its C tile at once, stages A
// and B into LDS in one pass and runs the whole K extent without further
// barriers.  A chunked pipeline pays one global-memory round trip per
// 32-wide chunk plus one for C, and with a few short tiles per CU those
// round trips, not the matrix cores, set the time.
constexpr int K1 = kNB;       // K extent of the one-shot form
constexpr int K1P = K1 + 4;   // FP32 LDS row pitch (floats)
constexpr int K1H = K1 + 8;   // BF16 LDS row pitch (elements)

// LDS of one gemm_tile_once call, in floats (the caller provides it).
template <int BM, int BN, bool SPLIT>
constexpr int once_lds_floats() {
  return SPLIT ? (BM + BN) * K1H : (BM + BN) * K1P;
}

// MODE 0: C -= A B^T on the lower part (i >= j), i < ilim, j < jlim, the old
//         values read from Csrc with row stride ldcs (A's on a first touch);
// MODE 1: C  = A B^T, i < ilim (panel; A and C alias row-for-row).
// SPLIT: BF16x3 products (x = hi + lo, hi*hi + hi*lo + lo*hi).

template <int BM, int BN, int MODE, bool VEC, bool SPLIT>
__device__ __forceinline__ void gemm_tile_once(const float* Abase, int64_t lda, int a_row0, int a_lim, int ka,
                                               const float* Bbase, int64_t ldb, int b_row0, int b_lim, int kb,
                                               float* Cbase, const float* Csrc, int64_t ldc, int64_t ldcs,
                                               int c_row0, int c_col0, int c_ilim, int c_jlim,
                                               float* __restrict__ lds) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int PA = BM * K1 / 4 / kThreads;
  constexpr int PB = BN * K1 / 4 / kThreads;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * WM;
  const int wn = (wave & 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  float4 ra[PA], rb[PB];
  load_rows<BM, VEC, K1>(ra, Abase, lda, a_row0, a_lim, ka, tid);
  load_rows<BN, VEC, K1, MODE == 1>(rb, Bbase, ldb, b_row0, b_lim, kb, tid);

  // C tile of this wave: offsets, masks and (MODE 0) the old values, whose
  // loads fly together with the operand loads.  A masked element reads
  // C[0][0], always valid; only the stores are masked.
  const int r0 = c_row0 + wm, cq = c_col0 + wn;
  int64_t off[FM][FN][4];
  bool ok[FM][FN][4];
  float old[FM][FN][4];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = r0 + a * 16 + 4 * lg + v;
        const int j = cq + b * 16 + li;
        const bool m = (MODE == 0) ? (i < c_ilim && j < c_jlim && j <= i) : (i < c_ilim);
        ok[a][b][v] = m;
        off[a][b][v] = m ? (int64_t)i * ldc + j : 0;
        if constexpr (MODE == 0) old[a][b][v] = gld(Csrc + (m ? (int64_t)i * ldcs + j : 0));
      }

  floatx4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};

  if constexpr (SPLIT) {
    __bf16* const Ah = reinterpret_cast<__bf16*>(lds);
    __bf16* const Al = Ah + BM * K1H;
    __bf16* const Bh = Al + BM * K1H;
    __bf16* const Bl = Bh + BN * K1H;
    store_rows_split<BM, K1, K1H>(Ah, Al, ra, tid);
    store_rows_split<BN, K1, K1H>(Bh, Bl, rb, tid);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < K1; ks += 32) {
      bf16x8 fah[FM], fal[FM], fbh[FN], fbl[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        const int o = (wm + a * 16 + li) * K1H + ks + 8 * lg;
        fah[a] = *reinterpret_cast<const bf16x8*>(&Ah[o]);
        fal[a] = *reinterpret_cast<const bf16x8*>(&Al[o]);
      }
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        const int o = (wn + b * 16 + li) * K1H + ks + 8 * lg;
        fbh[b] = *reinterpret_cast<const bf16x8*>(&Bh[o]);
        fbl[b] = *reinterpret_cast<const bf16x8*>(&Bl[o]);
      }
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fal[a], fbh[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah[a], fbl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah[a], fbh[b], acc[a][b], 0, 0, 0);
        }
    }
  } else {
    float* const As = lds;
    float* const Bs = lds + BM * K1P;
    store_rows<BM, K1, K1P>(As, ra, tid);
    store_rows<BN, K1, K1P>(Bs, rb, tid);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < K1; ks += 16) {
      float4 fa[FM], fb[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a)
        fa[a] = *reinterpret_cast<const float4*>(As + (wm + a * 16 + li) * K1P + ks + 4 * lg);
#pragma unroll
      for (int b = 0; b < FN; ++b)
        fb[b] = *reinterpret_cast<const float4*>(Bs + (wn + b * 16 + li) * K1P + ks + 4 * lg);
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].x, fb[b].x, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].y, fb[b].y, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].z, fb[b].z, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].w, fb[b].w, acc[a][b], 0, 0, 0);
    }
  }

  // The panel (MODE 1) overwrites its own A rows: every wave staged them
  // into LDS before the barrier above, so these stores cannot race.
  if constexpr (MODE == 0) {
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (ok[a][b][v]) gst(Cbase + off[a][b][v], old[a][b][v] - acc[a][b][v]);
  } else {
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (ok[a][b][v]) gst(Cbase + off[a][b][v], acc[a][b][v]);
  }
}

// Tile `bid` of the one-shot SYRK on matrix b: C -= P P^T, P = the 128
// columns from k0, C = the lower tiles of columns [col0, col1) (col1 < 0:
// to n); first: old values of C from A.
template <int T, bool VEC, bool SPLIT>
__device__ __forceinline__ void syrk_once_tile(const Batch& p, int bid, int64_t b, int col0, int col1, int k0, bool first,
                               float* lds) {
  const Mat M = mat_at(p, b);
  const int c1 = (col1 < 0 || col1 > M.n) ? M.n : col1;
  if (c1 <= col0) return;
  const int TI = (M.n - col0 + T - 1) / T;
  const int TJ = (c1 - col0 + T - 1) / T;
  const int64_t T64 = (int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2;
  if ((int64_t)bid >= T64) return;
  int ti, tj;
  decode_tile(xcd_tile(bid, (int)T64), TI, ti, tj);
  float* Lb = M.L;
  const float* Cs = first ? M.A : M.L;
  const int i0 = col0 + T * ti;
  const int j0 = col0 + T * tj;
  gemm_tile_once<T, T, 0, VEC, SPLIT>(Lb, M.ld, i0, M.n, k0, Lb, M.ld, j0, c1, k0, Lb, Cs, M.ld,
                                      first ? M.lda : M.ld, i0, j0, M.n, c1, lds);
}

// Rectangular one-shot tiles of the same update, TM rows x TN columns
// (TN = 2 TM): column tile t covers columns col0 + TN t .. + TN and the
// rows from its diagonal to n in TM-row tiles.  A workgroup stages its
// TN-row operand once for two TM x TM tiles' worth of work, so the extra
// workgroups of a diagonal-block launch take half as many rounds of the
// idle CUs (on its LDS).  rect_tiles() counts them per matrix; skip_diag:
// the 128 x 128 block at col0 is left out (the diagonal-block kernel
// updates it itself).
template <int TM, int TN>
__host__ __device__ __forceinline__ int rect_tiles(int n, int col0, int col1) {
  const int TJ = (col1 - col0 + TN - 1) / TN;
  int total = 0;
  for (int t = 0; t < TJ; ++t) total += (n - col0 - TN * t + TM - 1) / TM;
  return total;
}

template <int TM, int TN, bool VEC, bool SPLIT>
__device__ __forceinline__ void syrk_once_rect(const Batch& p, int bid, int64_t b, int col0, int col1, int k0, bool first,
                                               float* lds, bool skip_diag) {
  const Mat M = mat_at(p, b);
  const int c1 = (col1 < 0 || col1 > M.n) ? M.n : col1;
  if (c1 <= col0) return;
  const int TJ = (c1 - col0 + TN - 1) / TN;
  int t = 0, rem = bid;
  for (; t < TJ; ++t) {
    const int R = (M.n - col0 - TN * t + TM - 1) / TM;
    if (rem < R) break;
    rem -= R;
  }
  if (t >= TJ) return;
  const int i0 = col0 + TN * t + TM * rem;
  const int j0 = col0 + TN * t;
  if (skip_diag && i0 < col0 + 128 && j0 < col0 + 128) return;
  float* Lb = M.L;
  const float* Cs = first ? M.A : M.L;
  gemm_tile_once<TM, TN, 0, VEC, SPLIT>(Lb, M.ld, i0, M.n, k0, Lb, M.ld, j0, c1, k0, Lb, Cs, M.ld,
                                        first ? M.lda : M.ld, i0, j0, M.n, c1, lds);
}
}  // namespace
}  // namespace chol
