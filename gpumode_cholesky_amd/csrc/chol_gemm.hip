// chol_gemm.hip — MFMA products of the blocked path (gfx950).
//
//   syrk  : C[i][j] -= sum_k P[i][k] P[j][k]   over the lower trailing tiles
//   panel : L21 = A21 W^T                        (in place, K = 128)
//
// Both are C (op)= A B^T with A, B row-major along K, run on
// v_mfma_f32_16x16x4_f32 (plain) or as BF16x3 split products on
// v_mfma_f32_16x16x32_bf16 (mixed).  A workgroup of 4 wavefronts computes a BM x BN
// tile; each wavefront a (BM/2) x (BN/2) quarter as 16x16 MFMA blocks.  K is
// staged through LDS in chunks of 32 (rows padded to 36 floats), double
// buffered: the global loads of chunk c+1 are in flight while chunk c is
// multiplied.  Within a 16-wide K slice lane group g = lane/16 owns
// k = 4g..4g+3, so one ds_read_b128 feeds four MFMA steps (the sum over k is
// the same set of products in a different order).
//
// MFMA 16x16x4 f32 operand layout (checked on gfx950): A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16*(i/4), element i%4.
//
// Reference counterparts: the trailing baddbmm_ / addmm_ SYRK-as-GEMM and
// the panel-times-inverse bmm of the reference's blocked drivers
// (submission.py:5163-5184, :2529-2534, :3362-3401), which ran on TF32/FP16
// tensor cores; here FP32 or BF16x3 on the matrix cores.
#include "chol_common.h"
#include "chol_once.h"

namespace chol {
namespace {

// Write back a wave's FM x FN blocks of 16x16 accumulators (lane holds
// D[4*lg + v][li] of each block) at rows r0.., columns c0..:
//   MODE 0: C -= D where i < ilim, j < jlim, j <= i (lower trailing part);
//   MODE 1: C  = D where i < ilim.
// Every C load is issued unconditionally (a masked element reads C[0][0],
// always valid) so the loads stay in flight together; only the stores are
// masked.  The old values come from Cs with row stride ldcs (A's on a first
// touch, which may differ from L's in a ragged batch).
template <int FM, int FN, int MODE>
__device__ __forceinline__ void epilogue(const floatx4 (&acc)[FM][FN], float* C, const float* Cs, int64_t ldc,
                                         int64_t ldcs, int r0, int c0, int ilim, int jlim, int li, int lg) {
  int64_t off[FM][FN][4];
  bool ok[FM][FN][4];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = r0 + a * 16 + 4 * lg + v;
        const int j = c0 + b * 16 + li;
        const bool m = (MODE == 0) ? (i < ilim && j < jlim && j <= i) : (i < ilim);
        ok[a][b][v] = m;
        off[a][b][v] = m ? (int64_t)i * ldc + j : 0;
      }
  if constexpr (MODE == 0) {
    float old[FM][FN][4];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = r0 + a * 16 + 4 * lg + v;
          const int j = c0 + b * 16 + li;
          old[a][b][v] = gld(Cs + (ok[a][b][v] ? (int64_t)i * ldcs + j : 0));
        }
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (ok[a][b][v]) gst(C + off[a][b][v], old[a][b][v] - acc[a][b][v]);
  } else {
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (ok[a][b][v]) gst(C + off[a][b][v], acc[a][b][v]);
  }
}

// MODE 0: C -= A B^T on the lower part (i >= j), j < jlim, i < ilim.
// MODE 1: C  = A B^T, i < ilim (panel; A and C alias row-for-row).
template <int BM, int BN, int MODE, bool VEC>
__device__ __forceinline__ void gemm_tile(const float* Abase, int64_t lda, int a_row0, int a_lim,
                                          int ka, const float* Bbase, int64_t ldb, int b_row0,
                                          int b_lim, int kb, int K, float* Cbase, const float* Csrc,
                                          int64_t ldc, int64_t ldcs, int c_row0, int c_col0, int c_ilim,
                                          int c_jlim) {
  constexpr int WM = BM / 2, WN = BN / 2;      // per-wave tile
  constexpr int FM = WM / 16, FN = WN / 16;    // MFMA blocks per wave
  constexpr int PA = BM * KC / 4 / kThreads;
  constexpr int PB = BN * KC / 4 / kThreads;
  __shared__ __align__(16) float As[2][BM * KP];
  __shared__ __align__(16) float Bs[2][BN * KP];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * WM;
  const int wn = (wave & 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  floatx4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};

  float4 ra[PA], rb[PB];
  load_rows<BM, VEC>(ra, Abase, lda, a_row0, a_lim, ka, tid);
  load_rows<BN, VEC, KC, MODE == 1>(rb, Bbase, ldb, b_row0, b_lim, kb, tid);
  store_rows<BM>(As[0], ra, tid);
  store_rows<BN>(Bs[0], rb, tid);
  __syncthreads();

  const int nchunks = K / KC;
  for (int ch = 0; ch < nchunks; ++ch) {
    const int cur = ch & 1;
    const bool more = ch + 1 < nchunks;
    if (more) {
      load_rows<BM, VEC>(ra, Abase, lda, a_row0, a_lim, ka + (ch + 1) * KC, tid);
      load_rows<BN, VEC, KC, MODE == 1>(rb, Bbase, ldb, b_row0, b_lim, kb + (ch + 1) * KC, tid);
    }
    const float* as = As[cur];
    const float* bs = Bs[cur];
#pragma unroll
    for (int ks = 0; ks < KC; ks += 16) {
      float4 fa[FM], fb[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a)
        fa[a] = *reinterpret_cast<const float4*>(as + (wm + a * 16 + li) * KP + ks + 4 * lg);
#pragma unroll
      for (int b = 0; b < FN; ++b)
        fb[b] = *reinterpret_cast<const float4*>(bs + (wn + b * 16 + li) * KP + ks + 4 * lg);
      // k-step outermost: consecutive MFMAs write different accumulators.
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
    if (more) {
      store_rows<BM>(As[cur ^ 1], ra, tid);
      store_rows<BN>(Bs[cur ^ 1], rb, tid);
    }
    __syncthreads();
  }

  // Epilogue: lane holds D[4*lg + v][li] of each 16x16 block.
  epilogue<FM, FN, MODE>(acc, Cbase, Csrc, ldc, ldcs, c_row0 + wm, c_col0 + wn, c_ilim, c_jlim, li, lg);
}

// ------------------------------------------------------------ BF16x3 ----
// Mixed form of the SYRK: each FP32 operand x is split when staged into LDS
// as x = hi + lo (hi = bf16(x), lo = bf16(x - hi), v_cvt_pk_bf16_f32
// rounding to nearest even), and A B^T is accumulated in FP32 from three
// v_mfma_f32_16x16x32_bf16 products hiA hiB^T + hiA loB^T + loA hiB^T.
// The dropped loA loB^T term and the split remainder are ~2^-16 of |a b|.
// Operand layout of 16x16x32 bf16: A[i][k] for k = 8g..8g+7 in lane i + 16g,
// B likewise by column; D as for 16x16x4 f32.
// MODE 0: C -= A B^T on the lower part (i >= j), j < jlim, i < ilim;
// MODE 1: C = A B^T, i < ilim (panel); BF16x3 products.
template <int BM, int BN, bool VEC, int MODE = 0>
__device__ __forceinline__ void gemm_tile_bf16x3(const float* Abase, int64_t lda, int a_row0, int a_lim, int ka,
                                                 const float* Bbase, int64_t ldb, int b_row0, int b_lim, int kb,
                                                 int K, float* Cbase, const float* Csrc, int64_t ldc,
                                                 int64_t ldcs, int c_row0, int c_col0, int c_ilim, int c_jlim) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int PA = BM * KC / 4 / kThreads;
  constexpr int PB = BN * KC / 4 / kThreads;
  __shared__ __align__(16) __bf16 Ah[2][BM * KPH];
  __shared__ __align__(16) __bf16 Al[2][BM * KPH];
  __shared__ __align__(16) __bf16 Bh[2][BN * KPH];
  __shared__ __align__(16) __bf16 Bl[2][BN * KPH];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * WM;
  const int wn = (wave & 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  floatx4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};

  float4 ra[PA], rb[PB];
  load_rows<BM, VEC>(ra, Abase, lda, a_row0, a_lim, ka, tid);
  load_rows<BN, VEC, KC, MODE == 1>(rb, Bbase, ldb, b_row0, b_lim, kb, tid);
  store_rows_split<BM>(Ah[0], Al[0], ra, tid);
  store_rows_split<BN>(Bh[0], Bl[0], rb, tid);
  __syncthreads();

  const int nchunks = K / KC;
  for (int ch = 0; ch < nchunks; ++ch) {
    const int cur = ch & 1;
    const bool more = ch + 1 < nchunks;
    if (more) {
      load_rows<BM, VEC>(ra, Abase, lda, a_row0, a_lim, ka + (ch + 1) * KC, tid);
      load_rows<BN, VEC, KC, MODE == 1>(rb, Bbase, ldb, b_row0, b_lim, kb + (ch + 1) * KC, tid);
    }
    bf16x8 fah[FM], fal[FM], fbh[FN], fbl[FN];
#pragma unroll
    for (int a = 0; a < FM; ++a) {
      const int off = (wm + a * 16 + li) * KPH + 8 * lg;
      fah[a] = *reinterpret_cast<const bf16x8*>(&Ah[cur][off]);
      fal[a] = *reinterpret_cast<const bf16x8*>(&Al[cur][off]);
    }
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      const int off = (wn + b * 16 + li) * KPH + 8 * lg;
      fbh[b] = *reinterpret_cast<const bf16x8*>(&Bh[cur][off]);
      fbl[b] = *reinterpret_cast<const bf16x8*>(&Bl[cur][off]);
    }
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fal[a], fbh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah[a], fbl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah[a], fbh[b], acc[a][b], 0, 0, 0);
      }
    if (more) {
      store_rows_split<BM>(Ah[cur ^ 1], Al[cur ^ 1], ra, tid);
      store_rows_split<BN>(Bh[cur ^ 1], Bl[cur ^ 1], rb, tid);
    }
    __syncthreads();
  }

  epilogue<FM, FN, MODE>(acc, Cbase, Csrc, ldc, ldcs, c_row0 + wm, c_col0 + wn, c_ilim, c_jlim, li, lg);
}

// SYRK tile T x T (128, or 64 when a launch has too few 128-tiles to fill
// the GPU); panel row tile PT (same rule).
template <int T, bool VEC, bool SPLIT>
__global__ __launch_bounds__(kThreads) void k_syrk_mfma(Batch p, int col0, int col1, int k0, int kl) {
  const int64_t b = blockIdx.y;
  const Mat M = mat_at(p, b);
  const int c1 = (col1 < 0 || col1 > M.n) ? M.n : col1;  // col1 < 0: up to this matrix's order
  if (c1 <= col0) return;
  const int TI = (M.n - col0 + T - 1) / T;
  const int TJ = (c1 - col0 + T - 1) / T;
  const int64_t T64 = (int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2;
  if ((int64_t)blockIdx.x >= T64) return;
  int ti, tj;
  decode_tile(xcd_tile(blockIdx.x, (int)T64), TI, ti, tj);
  float* Lb = M.L;
  const float* Cs = p.first ? M.A : M.L;  // first touch: old values from A (its own row stride)
  const int64_t ldcs = p.first ? M.lda : M.ld;
  const int i0 = col0 + T * ti;
  const int j0 = col0 + T * tj;
  if constexpr (SPLIT)
    gemm_tile_bf16x3<T, T, VEC>(Lb, M.ld, i0, M.n, k0, Lb, M.ld, j0, c1, k0, kl, Lb, Cs, M.ld, ldcs, i0, j0, M.n,
                                c1);
  else
    gemm_tile<T, T, 0, VEC>(Lb, M.ld, i0, M.n, k0, Lb, M.ld, j0, c1, k0, kl, Lb, Cs, M.ld, ldcs, i0, j0, M.n, c1);
}

// Panel: L21 = A21 W^T (SPLIT: BF16x3 products, the mixed form).
template <int PT, bool VEC, bool SPLIT>
__global__ __launch_bounds__(kThreads) void k_panel_mfma(Batch p, int k0, const float* __restrict__ winv) {
  const int64_t b = blockIdx.y;
  const Mat M = mat_at(p, b);
  const int i0 = k0 + kNB + PT * blockIdx.x;
  if (i0 >= M.n) return;  // (ragged) rows of a smaller matrix
  float* Lb = M.L;
  const float* As = p.first ? M.A : M.L;  // first touch: A21 from A (its own row stride)
  const int64_t lda = p.first ? M.lda : M.ld;
  const float* Wb = winv + b * (kNB * kNB);
  if constexpr (SPLIT)
    gemm_tile_bf16x3<PT, kNB, VEC, 1>(As, lda, i0, M.n, k0, Wb, kNB, 0, kNB, 0, kNB, Lb, Lb, M.ld, M.ld, i0, k0,
                                      M.n, k0 + kNB);
  else
    gemm_tile<PT, kNB, 1, VEC>(As, lda, i0, M.n, k0, Wb, kNB, 0, kNB, 0, kNB, Lb, Lb, M.ld, M.ld, i0, k0, M.n,
                               k0 + kNB);
}

// One-shot forms (K = 128) of the two kernels above (tiles in chol_once.h).

template <int T, bool VEC, bool SPLIT>
__global__ __launch_bounds__(kThreads) void k_syrk_once(Batch p, int col0, int col1, int k0) {
  __shared__ __align__(16) float lds[once_lds_floats<T, T, SPLIT>()];
  syrk_once_tile<T, VEC, SPLIT>(p, blockIdx.x, blockIdx.y, col0, col1, k0, p.first, lds);
}

template <int PT, bool VEC, bool SPLIT>
__global__ __launch_bounds__(kThreads) void k_panel_once(Batch p, int k0, const float* __restrict__ winv) {
  const int64_t b = blockIdx.y;
  const Mat M = mat_at(p, b);
  const int i0 = k0 + kNB + PT * blockIdx.x;
  if (i0 >= M.n) return;
  float* Lb = M.L;
  const float* As = p.first ? M.A : M.L;
  const int64_t lda = p.first ? M.lda : M.ld;
  const float* Wb = winv + b * (kNB * kNB);
  __shared__ __align__(16) float lds[once_lds_floats<PT, kNB, SPLIT>()];
  gemm_tile_once<PT, kNB, 1, VEC, SPLIT>(As, lda, i0, M.n, k0, Wb, kNB, 0, kNB, 0, Lb, Lb, M.ld, M.ld, i0, k0, M.n,
                                         k0 + kNB, lds);
}


// Workgroups that keep every CU busy with at least two (256 CUs x 2).
constexpr int64_t kFillWGs = 512;

template <int T, bool SPLIT>
void syrk_launch(const Batch& p, int col0, int col1, int k0, int kl, bool vec, hipStream_t s) {
  // grid.x covers the largest matrix (p.n); smaller ragged ones exit early
  const int TI = (p.n - col0 + T - 1) / T;
  const int TJ = ((col1 < 0 ? p.n : col1) - col0 + T - 1) / T;
  const int64_t tiles = (int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2;
  dim3 grid((unsigned)tiles, (unsigned)p.batch);
  if (vec)
    hipLaunchKernelGGL((k_syrk_mfma<T, true, SPLIT>), grid, dim3(kThreads), 0, s, p, col0, col1, k0, kl);
  else
    hipLaunchKernelGGL((k_syrk_mfma<T, false, SPLIT>), grid, dim3(kThreads), 0, s, p, col0, col1, k0, kl);
}

}  // namespace

bool vec_ok(const Batch& p) {
  if (p.Ap) return p.ld % 4 == 0;  // ragged: the host sets ld = 4 when every matrix is aligned
  // a first-touch launch also reads A with float4 loads
  const uintptr_t base = reinterpret_cast<uintptr_t>(p.L) | (p.first ? reinterpret_cast<uintptr_t>(p.A) : 0);
  return (base & 15) == 0 && (p.ld % 4) == 0 && (p.bs % 4) == 0;
}

int launch_panel(const Batch& p, int k0, const float* winv, bool split, hipStream_t s) {
  const int rows = p.n - (k0 + kNB);
  if (rows <= 0) return 0;
  // One-shot 32-row tiles (W, 64 KiB, and the rows staged in one pass)
  // while four times the 128-row tile count still fits the GPU about once;
  // beyond that the pipelined 128-row tiles, which stage W once per 128 rows.
  const bool small = (int64_t)((rows + 127) / 128) * p.batch <= 64;
  const bool vec = vec_ok(p);
#define CHOL_PANEL(KERNEL, PT, ROWS_PER)                                                                        \
  {                                                                                                            \
    dim3 grid((unsigned)((rows + (ROWS_PER)-1) / (ROWS_PER)), (unsigned)p.batch);                              \
    if (vec && split)                                                                                          \
      hipLaunchKernelGGL((KERNEL<PT, true, true>), grid, dim3(kThreads), 0, s, p, k0, winv);                   \
    else if (vec)                                                                                              \
      hipLaunchKernelGGL((KERNEL<PT, true, false>), grid, dim3(kThreads), 0, s, p, k0, winv);                  \
    else if (split)                                                                                            \
      hipLaunchKernelGGL((KERNEL<PT, false, true>), grid, dim3(kThreads), 0, s, p, k0, winv);                  \
    else                                                                                                       \
      hipLaunchKernelGGL((KERNEL<PT, false, false>), grid, dim3(kThreads), 0, s, p, k0, winv);                 \
  }
  if (small)
    CHOL_PANEL(k_panel_once, 32, 32)
  else
    CHOL_PANEL(k_panel_mfma, 128, 128)
#undef CHOL_PANEL
  return launch_status("k_panel_mfma");
}

int launch_syrk(const Batch& p, int col0, int col1, int k0, int kl, bool split, hipStream_t s) {
  if ((col1 >= 0 && col1 <= col0) || col0 >= p.n || kl <= 0) return 0;
  if (kl % KC != 0) {
    set_error("launch_syrk: K=%d is not a multiple of %d", kl, KC);
    return CHOL_E_ARG;
  }
  const int TI = (p.n - col0 + 127) / 128;
  const int TJ = ((col1 < 0 ? p.n : col1) - col0 + 127) / 128;
  const int64_t tiles128 = ((int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2) * p.batch;
  const bool small = tiles128 < kFillWGs;
  const bool vec = vec_ok(p);
  if (!small && kl >= 256 && syrk256_tiles(p, col0, col1) >= 256)
    return launch_syrk256(p, col0, col1, k0, kl, split, vec, s);
  if (small && kl == K1) {
    // one-shot 64x64 tiles (the per-step update of one or a few matrices)
    const int TI = (p.n - col0 + 63) / 64;
    const int TJ = ((col1 < 0 ? p.n : col1) - col0 + 63) / 64;
    dim3 grid((unsigned)((int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2), (unsigned)p.batch);
    if (split) {
      if (vec)
        hipLaunchKernelGGL((k_syrk_once<64, true, true>), grid, dim3(kThreads), 0, s, p, col0, col1, k0);
      else
        hipLaunchKernelGGL((k_syrk_once<64, false, true>), grid, dim3(kThreads), 0, s, p, col0, col1, k0);
    } else {
      if (vec)
        hipLaunchKernelGGL((k_syrk_once<64, true, false>), grid, dim3(kThreads), 0, s, p, col0, col1, k0);
      else
        hipLaunchKernelGGL((k_syrk_once<64, false, false>), grid, dim3(kThreads), 0, s, p, col0, col1, k0);
    }
    return launch_status("k_syrk_once");
  }
  if (split) {
    if (small)
      syrk_launch<64, true>(p, col0, col1, k0, kl, vec, s);
    else
      syrk_launch<128, true>(p, col0, col1, k0, kl, vec, s);
  } else {
    if (small)
      syrk_launch<64, false>(p, col0, col1, k0, kl, vec, s);
    else
      syrk_launch<128, false>(p, col0, col1, k0, kl, vec, s);
  }
  return launch_status("k_syrk_mfma");
}

}  // namespace chol
