// chol_diag.hip — factor one 128x128 diagonal block per matrix and form its
// inverse: the serial link of the blocked path (one launch per 128 columns).
//
// One workgroup of 4 wavefronts per matrix, the block resident in LDS.
// Wavefront 0 runs the pivot chains: the 64x64 blocks D1 (top left) and then
// D2 (bottom right), each as four 16-column panels (factor64_step: MFMA
// update of the panel by the earlier ones, then the register chain, lane =
// row, pivots broadcast with v_readlane).  Everything else runs on wavefronts 1-3 in the
// shadow of those chains, one panel behind, with a workgroup barrier after
// each panel:
//   * waves 1 and 2: L21 of the finished panel (rows 64..127 of its 16
//     columns, a triangular solve in registers, lane = row; both waves solve,
//     each keeps its copy in its own LDS slab) and the panel's rank-16 update
//     of the later A21 columns and of D2;
//   * wave 3: the next block row of W1 = D1^-1 (and, during D2, of W2);
//   * during D2: L21 W1 = T1, W21 = -W2 T1 block row by block row, and the
//     stores of what is final.
// What is left after the last D2 panel is one block row of W2, one of W21
// and the final stores.
//
// Reference counterpart: e62_diag128 (submission.py:4694-4997), which
// factors a 128x128 block and builds its inverse in one CTA.
#include "chol_common.h"
#include "chol_block.h"
#include "chol_once.h"

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

constexpr int XP = 16 + 4;  // pitch of the L21 panel slabs
constexpr int PH = NB + 8;  // BF16 pitch of a staged previous-panel block (PRE)

// LDS of the kernel: the staging tiles first, then the block and its
// inverse (Ls starts below 64 KiB, so most of its accesses encode their
// offsets as immediates), then the two L21 panel slabs.
struct __align__(16) DiagLds {
  float Ts[4][16 * TP];  // per-wave 16x16 staging
  float Ls[NB * P];
  float Ws[NB * P];
  float Xp[2][H * XP];  // waves 1, 2: their copy of the current L21 panel
};
// PRE: a previous-panel block as BF16 hi / lo fills Ws and runs into Xp,
// which is unused until then.
static_assert(2 * NB * PH * 2 <= (NB * P + 2 * H * XP) * 4, "staged BF16 block exceeds Ws + Xp");
static_assert(once_lds_floats<64, 128, true>() * 4 <= sizeof(DiagLds) &&
                  once_lds_floats<64, 128, false>() * 4 <= sizeof(DiagLds),
              "bulk update tile exceeds the kernel's LDS");

// C(16x16) += X Y^T, rows of X (pitch px) and Y (pitch py) read as float4, K = 16.
__device__ __forceinline__ floatx4 mma_nt2(const float* X, int px, const float* Y, int py, floatx4 acc, int li,
                                           int lg) {
  const float4 xa = *reinterpret_cast<const float4*>(X + li * px + 4 * lg);
  const float4 yb = *reinterpret_cast<const float4*>(Y + li * py + 4 * lg);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, yb.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, yb.y, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, yb.z, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, yb.w, acc, 0, 0, 0);
  return acc;
}

// The lower triangle of the 16x16 block of Ls at (d0, d0), read whole
// into registers (uniform addresses: LDS broadcasts), so the triangular
// chains that follow never wait on LDS.
__device__ __forceinline__ void load_tri16(const float* Ls, int d0, float (&L)[16][16]) {
#pragma unroll
  for (int k = 0; k < 16; ++k)
#pragma unroll
    for (int q = 0; q <= k; q += 4) {
      const float4 v = *reinterpret_cast<const float4*>(&Ls[(d0 + k) * P + d0 + q]);
      L[k][q] = v.x;
      L[k][q + 1] = v.y;
      L[k][q + 2] = v.z;
      L[k][q + 3] = v.w;
    }
}

// Block row R of W = inverse of the lower-triangular 64x64 block Ls(d.., d..),
// one wavefront, block rows 0..R-1 of W already in Ws:
//   W_Rm = -W_RR Q_m, Q_m = sum_{t=m}^{R-1} L_Rt W_tm   (m < R),
//   W_RR = L_RR^-1.
// w_row_sums forms Q_0..Q_{R-1} in T[0..R-1]; w_row_finish forms W_RR
// (lane = column, column-oriented forward substitution: each solved entry
// is applied to all later ones at once, so the dependent chain is two
// operations per column) and the products.
template <int R>
__device__ __forceinline__ void w_row_sums(const float* Ls, const float* Ws, float (*T)[16 * TP], int d, int li,
                                           int lg) {
  const int d0 = d + 16 * R;
  floatx4 acc[R > 0 ? R : 1];
#pragma unroll
  for (int m = 0; m < R; ++m) acc[m] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < R; ++t)
#pragma unroll
    for (int m = 0; m <= t; ++m)
      acc[m] = mma_nn(&Ls[d0 * P + d + 16 * t], P, &Ws[(d + 16 * t) * P + d + 16 * m], P, acc[m], li, lg);
#pragma unroll
  for (int m = 0; m < R; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) T[m][(4 * lg + v) * TP + li] = acc[m][v];
}

// W_RR = L_RR^-1 for the 16x16 diagonal block at (d0, d0) into Ws.
__device__ __forceinline__ void w_diag_inv(const float* Ls, float* Ws, int d0, int lane, int li) {
  float L[16][16];
  load_tri16(Ls, d0, L);
  float rinv[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) rinv[i] = __builtin_amdgcn_rcpf(L[i][i]);
  const int c = li;
  float w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = (i == c) ? 1.f : 0.f;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    w[q] *= rinv[q];
#pragma unroll
    for (int i = q + 1; i < 16; ++i) w[i] = fmaf(-L[i][q], w[q], w[i]);
  }
  // every lane group holds the same columns: all store (same values), so
  // the code stays one basic block and interleaves with the caller's MFMAs
#pragma unroll
  for (int i = 0; i < 16; ++i) Ws[(d0 + i) * P + d0 + c] = w[i];
}

// W_Rm = -W_RR T[m] for m in [M0, M1) (W_RR already in Ws).
template <int M0, int M1>
__device__ __forceinline__ void w_row_products(float* Ws, const float (*T)[16 * TP], int d, int R, int li, int lg) {
  const int d0 = d + 16 * R;
  floatx4 r[M1 - M0];
#pragma unroll
  for (int m = M0; m < M1; ++m)
    r[m - M0] = mma_nn(&Ws[d0 * P + d0], P, T[m], TP, floatx4{0.f, 0.f, 0.f, 0.f}, li, lg);
#pragma unroll
  for (int m = M0; m < M1; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) Ws[(d0 + 4 * lg + v) * P + d + 16 * m + li] = -r[m - M0][v];
}

template <int R>
__device__ __forceinline__ void w_row_finish(const float* Ls, float* Ws, float (*T)[16 * TP], int d, int lane,
                                             int li, int lg) {
  w_diag_inv(Ls, Ws, d + 16 * R, lane, li);
  wave_lds_sync();
  if constexpr (R > 0) w_row_products<0, R>(Ws, T, d, R, li, lg);
  wave_lds_sync();
}

template <int R>
__device__ __forceinline__ void w_block_row(const float* Ls, float* Ws, float (*T)[16 * TP], int d, int lane, int li,
                                            int lg) {
  w_row_sums<R>(Ls, Ws, T, d, li, lg);
  w_row_finish<R>(Ls, Ws, T, d, lane, li, lg);
}

// L21 of D1 panel Q (columns 16Q..16Q+15, rows 64..127; lane = row 64 +
// lane): x = a L_QQ^-T by column-oriented forward substitution, a = the
// panel's A21 columns (already brought up to date by panels < Q).  The
// result goes to x and to the wave's slab Xq (row lane, pitch XP).
template <int Q>
__device__ __forceinline__ void l21_panel(const float* Ls, float* Xq, int lane, float (&x)[16]) {
  const float* arow = &Ls[(H + lane) * P + 16 * Q];
#pragma unroll
  for (int j = 0; j < 16; j += 4) {
    const float4 v = *reinterpret_cast<const float4*>(arow + j);
    x[j] = v.x;
    x[j + 1] = v.y;
    x[j + 2] = v.z;
    x[j + 3] = v.w;
  }
  float L[16][16];
  load_tri16(Ls, 16 * Q, L);
  float rinv[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) rinv[k] = __builtin_amdgcn_rcpf(L[k][k]);
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    x[m] *= rinv[m];
#pragma unroll
    for (int k = m + 1; k < 16; ++k) x[k] = fmaf(-x[m], L[k][m], x[k]);
  }
#pragma unroll
  for (int j = 0; j < 16; j += 4)
    *reinterpret_cast<float4*>(&Xq[lane * XP + j]) = make_float4(x[j], x[j + 1], x[j + 2], x[j + 3]);
  wave_lds_sync();
}

// Rank-16 update by L21 panel Q (in Xq) of the later A21 columns
// (A21[:, 16tj..] -= X L[16tj.., panel Q]^T, tj > Q: read-modify-write, the
// next panel's solve reads them) and of the lower tiles of D2 (D2 -= X X^T:
// accumulated in registers over the four panels, subtracted once after the
// last).  Wave HH (0 or 1) of the two takes every other tile of each kind.
template <int Q, int HH>
__device__ __forceinline__ void l21_updates(float* Ls, const float* Xq, floatx4 (&d2)[5], int li, int lg) {
  constexpr int NA = 4 * (3 - Q);  // A21 tiles
  constexpr int NM = (NA - HH + 1) / 2;
  floatx4 acc[NM > 0 ? NM : 1];
  int off[NM > 0 ? NM : 1];
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const int t = HH + 2 * k;
    const int ti = t & 3, tj = Q + 1 + (t >> 2);
    acc[k] = mma_nt2(&Xq[16 * ti * XP], XP, &Ls[16 * tj * P + 16 * Q], P, floatx4{0.f, 0.f, 0.f, 0.f}, li, lg);
    off[k] = (H + 16 * ti + 4 * lg) * P + 16 * tj + li;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const int u = HH + 2 * k;  // lower tile u of D2, row by row: (0,0) (1,0) (1,1) (2,0) ...
    const int ti = u >= 6 ? 3 : u >= 3 ? 2 : u >= 1 ? 1 : 0;
    const int tj = u - ti * (ti + 1) / 2;
    d2[k] = mma_nt2(&Xq[16 * ti * XP], XP, &Xq[16 * tj * XP], XP, d2[k], li, lg);
  }
  if constexpr (NM > 0) {
    float old[NM][4];
#pragma unroll
    for (int k = 0; k < NM; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) old[k][v] = Ls[off[k] + v * P];
#pragma unroll
    for (int k = 0; k < NM; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) Ls[off[k] + v * P] = old[k][v] - acc[k][v];
  }
  if constexpr (Q == 3) {
    float old[5][4];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int u = HH + 2 * k;
      const int ti = u >= 6 ? 3 : u >= 3 ? 2 : u >= 1 ? 1 : 0;
      const int tj = u - ti * (ti + 1) / 2;
#pragma unroll
      for (int v = 0; v < 4; ++v) old[k][v] = Ls[(H + 16 * ti + 4 * lg + v) * P + H + 16 * tj + li];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int u = HH + 2 * k;
      const int ti = u >= 6 ? 3 : u >= 3 ? 2 : u >= 1 ? 1 : 0;
      const int tj = u - ti * (ti + 1) / 2;
#pragma unroll
      for (int v = 0; v < 4; ++v)  // diagonal tile: only j <= i changes
        Ls[(H + 16 * ti + 4 * lg + v) * P + H + 16 * tj + li] =
            (ti != tj || li <= 4 * lg + v) ? old[k][v] - d2[k][v] : old[k][v];
    }
  }
}

// T1 = L21 W1 (into the upper-right 64x64 block of Ls, which L never uses),
// column tiles TA and TB, all four row tiles; W1 is lower triangular, so
// tile column tj takes K = 16 (4 - tj).
template <int TA, int TB>
__device__ __forceinline__ void t1_tiles(float* Ls, const float* Ws, int li, int lg) {
  floatx4 acc[2][4];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    constexpr int tjs[2] = {TA, TB};
    const int tj = tjs[u];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
      acc[u][ti] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k >= tj)
          acc[u][ti] = mma_nn(&Ls[(H + 16 * ti) * P + 16 * k], P, &Ws[16 * k * P + 16 * tj], P, acc[u][ti], li, lg);
    }
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    constexpr int tjs[2] = {TA, TB};
    const int tj = tjs[u];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v) Ls[(16 * ti + 4 * lg + v) * P + H + 16 * tj + li] = acc[u][ti][v];
  }
}

// C(16x16) += X Y with Y held in registers in the accumulator layout
// (lane (li, lg) holds Y[4lg+v][li] in y[v]), X rows read as float4, K = 16.
__device__ __forceinline__ floatx4 mma_nr(const float* X, int px, const floatx4& y, floatx4 acc, int li, int lg) {
  const float4 xa = *reinterpret_cast<const float4*>(X + li * px + 4 * lg);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, y[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, y[1], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, y[2], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, y[3], acc, 0, 0, 0);
  return acc;
}

// W21 block row R (R < 3), column tiles TA and TB: -sum_{m<=R} W2_Rm T1_m,tj,
// then the same tiles of U = T1_3 + sum_{t<3} L22_3t W21_t (kept in
// registers; U is loaded from T1_3 at R = 0), from which W21's block row 3
// is -W2_33 U: block row 3 of L22 W21 = -T1.
template <int R, int TA, int TB>
__device__ __forceinline__ void w21_tiles(const float* Ls, float* Ws, floatx4 (&U)[2], int li, int lg) {
  floatx4 acc[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
  for (int m = 0; m <= R; ++m) {
    acc[0] = mma_nn(&Ws[(H + 16 * R) * P + H + 16 * m], P, &Ls[16 * m * P + H + 16 * TA], P, acc[0], li, lg);
    acc[1] = mma_nn(&Ws[(H + 16 * R) * P + H + 16 * m], P, &Ls[16 * m * P + H + 16 * TB], P, acc[1], li, lg);
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) acc[u] = -acc[u];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    Ws[(H + 16 * R + 4 * lg + v) * P + 16 * TA + li] = acc[0][v];
    Ws[(H + 16 * R + 4 * lg + v) * P + 16 * TB + li] = acc[1][v];
  }
  if constexpr (R == 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      U[0][v] = Ls[(48 + 4 * lg + v) * P + H + 16 * TA + li];
      U[1][v] = Ls[(48 + 4 * lg + v) * P + H + 16 * TB + li];
    }
  }
  const float* L3 = &Ls[(H + 48) * P + H + 16 * R];  // L22_3R
  U[0] = mma_nr(L3, P, acc[0], U[0], li, lg);
  U[1] = mma_nr(L3, P, acc[1], U[1], li, lg);
}

// L = 0 above the diagonal in rows [r0, r1) of every matrix, in whole
// 4-column chunks right of each row's diagonal chunk (the diagonal-block
// kernels store the chunks that hold a diagonal element), by `nwg`
// workgroups, streaming stores.
__device__ __forceinline__ void zero_upper_rows(const Batch& p, int r0, int r1, int wg, int nwg) {
  const int n = p.n;
  const int n4 = (n + 3) / 4;
  const bool vec = ((reinterpret_cast<uintptr_t>(p.L) & 15) == 0) && (p.ld % 4) == 0 && (p.bs % 4) == 0 &&
                   (n % 4) == 0;
  const int64_t per = (int64_t)(r1 - r0) * n4;
  const int64_t total = per * p.batch;
  for (int64_t e = (int64_t)wg * 256 + threadIdx.x; e < total; e += (int64_t)nwg * 256) {
    const int64_t b = e / per;
    const int64_t r = e - b * per;
    const int i = r0 + (int)(r / n4);
    const int c = 4 * (int)(r - (int64_t)(i - r0) * n4);
    if (c <= i) continue;
    float* q = p.L + b * p.bs + (int64_t)i * p.ld + c;
    if (vec) {
      gst4_nt(q, make_float4(0.f, 0.f, 0.f, 0.f));
    } else {
      for (int j = 0; j < 4 && c + j < n; ++j) gst_nt(q + j, 0.f);
    }
  }
}

// PRE (mixed form): D -= P P^T on lower 16x16 tiles of the block in Ls, P
// (128 x 128, the previous panel's rows of this block) staged as BF16 hi /
// lo (pitch PH) over Ws and Xp; BF16x3 products as in the SYRK tiles.  The
// tiles (TA, j), JA0 <= j < JA1, and (TB, j), JB0 <= j < JB1: independent
// accumulators, so the matrix-core chains of the tiles interleave.
template <int TA, int JA0, int JA1, int TB, int JB0, int JB1>
__device__ __forceinline__ void pre_tiles(float* Ls, const __bf16* Ph, int li, int lg) {
  constexpr int NA = JA1 - JA0;
  constexpr int NT = NA + (JB1 - JB0);
  const __bf16* Pl = Ph + NB * PH;
  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int ko = 32 * kk + 8 * lg;
    const bf16x8 aah = *reinterpret_cast<const bf16x8*>(Ph + (16 * TA + li) * PH + ko);
    const bf16x8 aal = *reinterpret_cast<const bf16x8*>(Pl + (16 * TA + li) * PH + ko);
    const bf16x8 abh = *reinterpret_cast<const bf16x8*>(Ph + (16 * TB + li) * PH + ko);
    const bf16x8 abl = *reinterpret_cast<const bf16x8*>(Pl + (16 * TB + li) * PH + ko);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int tj = t < NA ? JA0 + t : JB0 + t - NA;
      const bf16x8 ah = t < NA ? aah : abh;
      const bf16x8 al = t < NA ? aal : abl;
      const bf16x8 bh = *reinterpret_cast<const bf16x8*>(Ph + (16 * tj + li) * PH + ko);
      const bf16x8 bl = *reinterpret_cast<const bf16x8*>(Pl + (16 * tj + li) * PH + ko);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[t], 0, 0, 0);
    }
  }
  float old[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int ti = t < NA ? TA : TB, tj = t < NA ? JA0 + t : JB0 + t - NA;
#pragma unroll
    for (int v = 0; v < 4; ++v) old[t][v] = Ls[(16 * ti + 4 * lg + v) * P + 16 * tj + li];
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int ti = t < NA ? TA : TB, tj = t < NA ? JA0 + t : JB0 + t - NA;
#pragma unroll
    for (int v = 0; v < 4; ++v)  // diagonal tile: only j <= i changes
      Ls[(16 * ti + 4 * lg + v) * P + 16 * tj + li] = (ti != tj || li <= 4 * lg + v) ? old[t][v] - acc[t][v] : old[t][v];
  }
}

// The diagonal block at (k0, k0) of matrix b, one workgroup of 256 threads
// (every thread calls it): L11 into L, its inverse W (unless LAST) into Wb
// (128 x 128, row-major).  The body of k_diag_factor, and of each step of
// the fused per-matrix kernel below.
// PRE: the previous panel (columns bk0..bk0+127) has not yet updated this
// block: the block's update is applied to it in LDS first (bfirst: that
// panel is the first, the block's old values come from A).  from_a_rt: the
// block is staged from A, as for FROM_A, chosen at run time.  opaque: the
// thread index is hidden from the optimiser, for a caller that runs this
// in a loop with other large code (the fused kernel): otherwise addresses
// derived from it are hoisted out of the loop, stay live across the other
// phases and spill.
template <bool FROM_A, bool LAST, bool PRE>
__device__ __forceinline__ void diag_block(const Batch& p, int64_t b, int k0, float* __restrict__ Wb, DiagLds& sm,
                                           int bk0, int bfirst, bool from_a_rt = false, bool opaque = false) {
  // Where wave 0's panels switch from right- to left-looking updates
  // (factor64_step): D1 with PRE from its second panel (the other
  // wavefronts update D21 and D2 during the first); D1 without PRE from the
  // first (they are idle then); D2 right-looking, except in the last block,
  // where wave 0 factors it alone.
  constexpr int D1LF = PRE ? 1 : 0;
  constexpr int D2LF = LAST ? 0 : 4;
  float(*const Ts)[16 * TP] = sm.Ts;
  float* const Ls = sm.Ls;
  float* const Ws = sm.Ws;
  int tid = threadIdx.x;
  if (opaque) asm volatile("" : "+v"(tid));
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15;
  const int lg = lane >> 4;
  const Mat M = mat_at(p, b);
  const int nb = min(NB, M.n - k0);
  if (nb <= 0) return;  // (ragged) this matrix has no block at k0
  float* Lb = M.L + (int64_t)k0 * M.ld + k0;
  // FROM_A: the first block, staged from A (nothing has written L yet),
  // with A's own row stride (a ragged batch pads L's rows, not A's).
  const bool from_a = FROM_A || (PRE && bfirst != 0) || from_a_rt;
  const int64_t lds_ = from_a ? M.lda : M.ld;  // row stride of the staging source
  const float* Sb = from_a ? M.A + (int64_t)k0 * lds_ + k0 : Lb;
  const bool vec = ((M.ld & 3) == 0) && ((reinterpret_cast<uintptr_t>(M.L) & 15) == 0) &&
                   (!from_a || ((lds_ & 3) == 0 && (reinterpret_cast<uintptr_t>(M.A) & 15) == 0));
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
      stg[u] = gld4_nt(Sb + (int64_t)r * lds_ + (c4 <= r ? c4 : 0));
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
        const float* src = Sb + (int64_t)r * lds_ + c4;
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
  // PRE: the previous panel's rows of this block, requested right behind
  // the block (rows 0..63 are needed first: u < 8)
  float4 pst[PRE ? NB / 8 : 1];
  __bf16* const Ph = reinterpret_cast<__bf16*>(Ws);
  if constexpr (PRE) {
    const float* Pb = M.L + (int64_t)k0 * M.ld + bk0;
#pragma unroll
    for (int u = 0; u < NB / 8; ++u) {
      const int r = (tid >> 5) + 8 * u;
      pst[u] = r < nb ? gld4(Pb + (int64_t)r * M.ld + c4) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  auto put_p = [&](int u) {  // P row tid/32 + 8u as BF16 hi / lo
    const int r = (tid >> 5) + 8 * u;
    const float4 v = pst[u];
    uint2 hi, lo;
    split_bf16_pair(v.x, v.y, hi.x, lo.x);
    split_bf16_pair(v.z, v.w, hi.y, lo.y);
    *reinterpret_cast<uint2*>(Ph + r * PH + c4) = hi;
    *reinterpret_cast<uint2*>(Ph + NB * PH + r * PH + c4) = lo;
  };
#pragma unroll
  for (int u = 0; u < NB / 8; ++u) {
    const int r = (tid >> 5) + 8 * u;
    *reinterpret_cast<float4*>(&Ls[r * P + c4]) = stg[u];
    if constexpr (PRE) {
      if (u < NB / 16) put_p(u);
    }
  }
  __syncthreads();
  if constexpr (PRE) {
    // D1's ten tiles (P rows 0..63) by all four wavefronts, then the rest
    // of P into LDS; D21 and D2 (rows 64..127) follow in D1's first step.
    if (wave == 0)
      pre_tiles<0, 0, 1, 1, 0, 2>(Ls, Ph, li, lg);
    else if (wave == 1)
      pre_tiles<2, 0, 3, 0, 0, 0>(Ls, Ph, li, lg);
    else if (wave == 2)
      pre_tiles<3, 0, 2, 0, 0, 0>(Ls, Ph, li, lg);
    else
      pre_tiles<3, 2, 4, 0, 0, 0>(Ls, Ph, li, lg);
#pragma unroll
    for (int u = NB / 16; u < NB / 8; ++u) put_p(u);
    __syncthreads();
  }

  // ---- 2. D1 by wavefront 0; one panel behind it, waves 1-2 form that
  //         panel's L21 and its rank-16 update of A21 and D2, wave 3 the
  //         next block row of W1.  Wave 1 moves its L21 panel into Ls one
  //         round later, when no wave reads those A21 columns any more. ----
  float* Xq = sm.Xp[wave == 2 ? 1 : 0];
  int bad1 = 0, bad2 = 0;
  float x[16];
  floatx4 d2[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) d2[k] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto put_x = [&](int q) {
#pragma unroll
    for (int j = 0; j < 16; j += 4)
      *reinterpret_cast<float4*>(&Ls[(H + lane) * P + 16 * q + j]) = make_float4(x[j], x[j + 1], x[j + 2], x[j + 3]);
  };
#define CHOL_L21_ROUND(Q)                            \
  {                                                  \
    if (wave == 1 && (Q) > 0) put_x((Q) - 1);        \
    l21_panel<Q>(Ls, Xq, lane, x);                   \
    if (wave == 1)                                   \
      l21_updates<Q, 0>(Ls, Xq, d2, li, lg);         \
    else                                             \
      l21_updates<Q, 1>(Ls, Xq, d2, li, lg);         \
  }
  if (wave == 0) {
    factor64_step<16, P, 0, D1LF>(Ls, 0, lane, bad1);
  } else if constexpr (PRE) {  // rows 64..127: 26 tiles
    if (wave == 1)
      pre_tiles<4, 0, 5, 5, 0, 4>(Ls, Ph, li, lg);
    else if (wave == 2)
      pre_tiles<5, 4, 6, 6, 0, 7>(Ls, Ph, li, lg);
    else
      pre_tiles<7, 0, 8, 0, 0, 0>(Ls, Ph, li, lg);
  }
  __syncthreads();
  if (wave == 0) {
    factor64_step<16, P, 1, D1LF>(Ls, 0, lane, bad1);
  } else if (wave < 3) {
    CHOL_L21_ROUND(0)
  } else if constexpr (!LAST) {
    w_block_row<0>(Ls, Ws, Ts, 0, lane, li, lg);
  }
  __syncthreads();
  if (wave == 0) {
    factor64_step<16, P, 2, D1LF>(Ls, 0, lane, bad1);
  } else if (wave < 3) {
    CHOL_L21_ROUND(1)
  } else if constexpr (!LAST) {
    w_block_row<1>(Ls, Ws, Ts, 0, lane, li, lg);
  }
  __syncthreads();
  if (wave == 0) {
    factor64_step<16, P, 3, D1LF>(Ls, 0, lane, bad1);
  } else if (wave < 3) {
    CHOL_L21_ROUND(2)
  } else if constexpr (!LAST) {
    w_block_row<2>(Ls, Ws, Ts, 0, lane, li, lg);
  }
  __syncthreads();
  if (wave == 1 || wave == 2) {
    CHOL_L21_ROUND(3)
  } else if constexpr (!LAST) {
    // wave 3, otherwise idle here: the sums of W1's block row 3 (its
    // operands, D1's rows 48..63 and W1's rows 0..47, are final); the
    // rest of that block row follows in the next round, which it used to
    // hold up
    if (wave == 3) w_row_sums<3>(Ls, Ws, Ts, 0, li, lg);
  }
#undef CHOL_L21_ROUND
  __syncthreads();

  // Final stores (one wavefront each call; all LDS reads of a call before
  // its global stores): L chunks of 32 rows x 64 columns (rows < nb,
  // columns < nb, streaming stores: the factor is final; chunks above the
  // diagonal hold the zeros that L has there) and W chunks of 16 rows x 128
  // columns.  The upper-right 64x64 quadrant of the block is stored (zeros)
  // only when the block is the whole matrix; otherwise the zero fill of L's
  // upper triangle covers it (every chunk right of a row's diagonal chunk).
  auto store_l32 = [&](int r0, int c0) {
    float4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = lane + 64 * u;
      v[u] = *reinterpret_cast<const float4*>(&Ls[(r0 + (e >> 4)) * P + c0 + (e & 15) * 4]);
    }
    if (vec && nb == NB) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = lane + 64 * u;
        gst4_nt(Lb + (int64_t)(r0 + (e >> 4)) * M.ld + c0 + (e & 15) * 4, v[u]);
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = lane + 64 * u;
      const int r = r0 + (e >> 4), cc = c0 + (e & 15) * 4;
      if (r >= nb || cc >= nb) continue;
      float* dst = Lb + (int64_t)r * M.ld + cc;
      if (vec && cc + 3 < nb) {
        gst4_nt(dst, v[u]);
      } else {
        gst_nt(dst, v[u].x);
        if (cc + 1 < nb) gst_nt(dst + 1, v[u].y);
        if (cc + 2 < nb) gst_nt(dst + 2, v[u].z);
        if (cc + 3 < nb) gst_nt(dst + 3, v[u].w);
      }
    }
  };
  // W is lower triangular; Ws holds it on and below the diagonal only.
  auto store_w16 = [&](int r0) {
    float4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = lane + 64 * u;
      const int r = r0 + (e >> 5), c = (e & 31) * 4;
      float4 w = *reinterpret_cast<const float4*>(&Ws[r * P + c]);
      if (c > r) w.x = 0.f;
      if (c + 1 > r) w.y = 0.f;
      if (c + 2 > r) w.z = 0.f;
      if (c + 3 > r) w.w = 0.f;
      v[u] = w;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = lane + 64 * u;
      gst4_nt(&Wb[(r0 + (e >> 5)) * NB + (e & 31) * 4], v[u]);
    }
  };

  // ---- 3. D2 by wavefront 0, a barrier after each panel (not for LAST:
  //         no W, so nothing waits on its panels).  Meanwhile: round 0
  //         completes W1 and stores L rows 0..63, L21 and W rows 0..63;
  //         round 1 forms T1 = L21 W1 and W2's block row 0; rounds 2-3
  //         W2's next block row and W21's block row behind it, with U
  //         (see w21_tiles) in registers of waves 1-2. ----
  floatx4 U[2];
  if (wave == 0) {
    factor64_step<16, P, 0, D2LF>(Ls, H, lane, bad2);
  } else if (wave == 1) {
    put_x(3);
    wave_lds_sync();
    store_l32(H, 0);  // L21
    store_l32(H + 32, 0);
    if constexpr (!LAST) {
      store_w16(0);
      store_w16(16);
    }
  } else if (wave == 2) {
    store_l32(0, 0);  // L rows 0..63
    store_l32(32, 0);
    if constexpr (FROM_A && LAST) {  // the whole matrix: no zero fill
      store_l32(0, H);
      store_l32(32, H);
    }
    if constexpr (!LAST) store_w16(32);
  } else if constexpr (!LAST) {
    w_row_finish<3>(Ls, Ws, Ts, 0, lane, li, lg);  // its sums were formed in the round before
    store_w16(48);
  }
  if constexpr (!LAST) __syncthreads();
  if (wave == 0) {
    factor64_step<16, P, 1, D2LF>(Ls, H, lane, bad2);
  } else if constexpr (!LAST) {
    if (wave == 1)
      t1_tiles<0, 3>(Ls, Ws, li, lg);  // column tiles {0, 3} and {1, 2}:
    else if (wave == 2)                // equal work, W1 being lower triangular
      t1_tiles<1, 2>(Ls, Ws, li, lg);
    else
      w_block_row<0>(Ls, Ws, Ts, H, lane, li, lg);
  }
  if constexpr (!LAST) __syncthreads();
  if (wave == 0) {
    factor64_step<16, P, 2, D2LF>(Ls, H, lane, bad2);
  } else if constexpr (!LAST) {
    if (wave == 1)
      w21_tiles<0, 0, 2>(Ls, Ws, U, li, lg);
    else if (wave == 2)
      w21_tiles<0, 1, 3>(Ls, Ws, U, li, lg);
    else
      w_block_row<1>(Ls, Ws, Ts, H, lane, li, lg);
  }
  if constexpr (!LAST) __syncthreads();
  if (wave == 0) {
    factor64_step<16, P, 3, D2LF>(Ls, H, lane, bad2);
  } else if constexpr (!LAST) {
    if (wave == 1)
      w21_tiles<1, 0, 2>(Ls, Ws, U, li, lg);
    else if (wave == 2)
      w21_tiles<1, 1, 3>(Ls, Ws, U, li, lg);
    else
      w_block_row<2>(Ls, Ws, Ts, H, lane, li, lg);
  }
  if constexpr (!LAST) __syncthreads();

  // ---- 4. the tail: waves 0-1 store D2's factor, wave 0 forms the sums
  //         Q_m of W2's block row 3, wave 3 its diagonal block W2_33, waves
  //         1-2 W21's block row 2 (and U); then W2_3m = -W2_33 Q_m (waves 0, 3)
  //         and W21's block row 3 = -W2_33 U (waves 1-2); then the stores
  //         of W rows 64..127. ----
  if (wave == 0) {
    const int bad = bad1 != 0 ? bad1 : (bad2 != 0 ? H + bad2 : 0);
    if (p.info && lane == 0 && bad != 0 && bad <= nb && p.info[M.slot] == 0) p.info[M.slot] = k0 + bad;
    if constexpr (!LAST) w_row_sums<3>(Ls, Ws, Ts, H, li, lg);
    store_l32(H, H);
    if constexpr (LAST) store_l32(H + 32, H);
  } else if constexpr (!LAST) {
    if (wave == 1) {
      w21_tiles<2, 0, 2>(Ls, Ws, U, li, lg);
      store_l32(H + 32, H);  // D2's lower half (wave 0 has the longer round)
    } else if (wave == 2) {
      w21_tiles<2, 1, 3>(Ls, Ws, U, li, lg);
    } else {
      w_diag_inv(Ls, Ws, H + 48, lane, li);
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
    if (wave == 0) {
      w_row_products<0, 2>(Ws, Ts, H, 3, li, lg);
    } else if (wave == 3) {
      w_row_products<2, 3>(Ws, Ts, H, 3, li, lg);
    } else {
      const float* W33 = &Ws[(H + 48) * P + H + 48];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int tj = (wave - 1) + 2 * u;
        const floatx4 r = mma_nr(W33, P, U[u], floatx4{0.f, 0.f, 0.f, 0.f}, li, lg);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ws[(H + 48 + 4 * lg + v) * P + 16 * tj + li] = -r[v];
      }
    }
    __syncthreads();
    store_w16(H + 16 * wave);
  }
}

// Workgroups past the batch: nz of them zero L's upper triangle in rows
// [zr0, zr1); then, BULK (1 FP32, 2 BF16x3), the previous step's
// trailing-update tiles (64x128, one-shot; bt per matrix).
// PRE (with BULK == 2): the previous panel (columns bk0..bk0+127) has not
// yet updated this block: the kernel applies that update to it first, and
// the workgroups past the batch (skipping this block) the rest of it.
template <bool FROM_A, bool LAST, int BULK, bool PRE = false>
__global__ __launch_bounds__(256) void k_diag_factor(Batch p, int k0, float* __restrict__ winv, int bc0, int bc1,
                                                     int bk0, int bt, int bfirst, int zr0, int zr1, int nz) {
  __shared__ DiagLds sm;
  // Workgroups past the batch run on the CUs that the diagonal blocks leave
  // idle: a share of the zero fill of L's upper triangle, and the previous
  // step's trailing-update tiles beyond the next column block (which no
  // diagonal block of this launch reads).
  if (blockIdx.x >= p.batch) {
    int id = blockIdx.x - (int)p.batch;
    if (id < nz) {
      zero_upper_rows(p, zr0, zr1, id, nz);
      return;
    }
    id -= nz;
    if constexpr (BULK != 0) {
      syrk_once_rect<64, 128, true, BULK == 2>(p, id % bt, id / bt, bc0, bc1, bk0, bfirst != 0,
                                               reinterpret_cast<float*>(&sm), PRE);
    }
    return;
  }
  diag_block<FROM_A, LAST, PRE>(p, blockIdx.x, k0, winv + (int64_t)blockIdx.x * (NB * NB), sm, bk0, bfirst);
}

// ------------------------------------------------------------- fused ----
// L = 0 in the whole 4-column chunks right of each row's diagonal chunk of
// one matrix, by one workgroup (wavefront w: rows w, w+4, ...; lanes along
// the row), streaming stores.  The chunks that hold a diagonal element are
// stored by the diagonal-block code.
__device__ __forceinline__ void zero_upper_one(const Mat& M) {
  const int n = M.n;
  const int n4 = (n + 3) >> 2;
  const bool vec = ((reinterpret_cast<uintptr_t>(M.L) & 15) == 0) && (M.ld & 3) == 0 && (n & 3) == 0;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int i = wave; i < n; i += 4) {
    float* row = M.L + (int64_t)i * M.ld;
    for (int q = (i >> 2) + 1 + lane; q < n4; q += 64) {
      const int c = 4 * q;
      if (vec) {
        gst4_nt(row + c, make_float4(0.f, 0.f, 0.f, 0.f));
      } else {
        for (int j = 0; j < 4 && c + j < n; ++j) gst_nt(row + c + j, 0.f);
      }
    }
  }
}

// A stream of one-shot (K = 128) tiles that share one operand, for the
// fused kernel: C[i0 .. i0+64, j0 .. j0+128) for the 64-row tiles
// i0 = ibeg, ibeg + 64, ... < n.  B (128 rows from brow0, the 128 columns
// from kb) is staged into LDS once; each tile's A rows (the 128 columns
// from ka) and old C values are loaded into registers while the previous
// tile's MFMAs run, so a tile costs its MFMA time plus one LDS staging,
// not a global-memory round trip.
//   MODE 0: C = Cold - A B^T on {j <= i, i < n, j < jlim} (Cold from Cs,
//           row stride ldcs);
//   MODE 1: C = A B^T for i < n, j < jlim (the panel: C may be A's own
//           rows; a tile's rows are in LDS before any of its stores, and
//           the prefetched rows are the next tile's).
// SPLIT: BF16x3 products (x = hi + lo; hi*hi + hi*lo + lo*hi), FP32
// accumulation.  tid: the thread index (opaque to the optimiser).
template <int MODE, bool VEC, bool SPLIT>
__device__ __forceinline__ void stream_tiles(const float* A, int64_t lda, int ka, int ibeg, int n, const float* B,
                                             int64_t ldb, int brow0, int blim, int kb, float* C, const float* Cs,
                                             int64_t ldc, int64_t ldcs, int j0, int jlim, float* __restrict__ lds,
                                             int tid) {
  constexpr int TM = 64, TN = 128;
  constexpr int WM = 32, WN = 64, FM = WM / 16, FN = WN / 16;
  constexpr int PA = TM * K1 / 4 / kThreads;
  constexpr int PB = TN * K1 / 4 / kThreads;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * WM;
  const int wn = (wave & 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;
  __bf16* const Bh = reinterpret_cast<__bf16*>(lds);
  __bf16* const Bl = Bh + TN * K1H;
  __bf16* const Ah = Bl + TN * K1H;
  __bf16* const Al = Ah + TM * K1H;
  float* const Bs = lds;
  float* const As = lds + TN * K1P;

  float4 ra[PA];
  float oldn[FM][FN][4];
  auto load_tile = [&](int i0) {
    load_rows<TM, VEC, K1>(ra, A, lda, i0, n, ka, tid);
    if constexpr (MODE == 0) {
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = i0 + wm + a * 16 + 4 * lg + v;
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            const int j = j0 + wn + b * 16 + li;
            const bool ok = i < n && j < jlim && j <= i;
            oldn[a][b][v] = gld(Cs + (ok ? (int64_t)i * ldcs + j : 0));  // masked: C[0][0], always valid
          }
        }
    }
  };
  __syncthreads();  // the LDS is free (the caller's previous stream or block)
  {
    float4 rb[PB];
    load_rows<TN, VEC, K1>(rb, B, ldb, brow0, blim, kb, tid);
    if (ibeg < n) load_tile(ibeg);
    if constexpr (SPLIT)
      store_rows_split<TN, K1, K1H>(Bh, Bl, rb, tid);
    else
      store_rows<TN, K1, K1P>(Bs, rb, tid);
  }
  for (int i0 = ibeg; i0 < n; i0 += TM) {
    __syncthreads();  // the previous tile's LDS reads are done
    if constexpr (SPLIT)
      store_rows_split<TM, K1, K1H>(Ah, Al, ra, tid);
    else
      store_rows<TM, K1, K1P>(As, ra, tid);
    float oldc[FM][FN][4];
    if constexpr (MODE == 0) {
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) oldc[a][b][v] = oldn[a][b][v];
    }
    __syncthreads();
    if (i0 + TM < n) load_tile(i0 + TM);  // in flight during the MFMAs below

    floatx4 acc[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    if constexpr (SPLIT) {
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
#pragma unroll
      for (int ks = 0; ks < K1; ks += 16) {
        float4 fa[FM], fb[FN];
#pragma unroll
        for (int a = 0; a < FM; ++a)
          fa[a] = *reinterpret_cast<const float4*>(As + (wm + a * 16 + li) * K1P + ks + 4 * lg);
#pragma unroll
        for (int b = 0; b < FN; ++b)
          fb[b] = *reinterpret_cast<const float4*>(Bs + (wn + b * 16 + li) * K1P + ks + 4 * lg);
        // one k at a time over all FM x FN accumulators: consecutive MFMAs
        // are independent
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
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wm + a * 16 + 4 * lg + v;
        if (i >= n) continue;
        float* crow = C + (int64_t)i * ldc;
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          const int j = j0 + wn + b * 16 + li;
          if constexpr (MODE == 0) {
            if (j < jlim && j <= i) gst(crow + j, oldc[a][b][v] - acc[a][b][v]);
          } else {
            if (j < jlim) gst(crow + j, acc[a][b][v]);
          }
        }
      }
  }
}
static_assert((2 * 128 * K1H + 2 * 64 * K1H) * 2 <= (int)sizeof(DiagLds) &&
                  (128 + 64) * K1P * 4 <= (int)sizeof(DiagLds),
              "stream_tiles staging exceeds the fused kernel's LDS");

// The whole factorisation of matrix b by one workgroup: for each 128
// columns, the diagonal block and its inverse (diag_block, W at Wb), the
// panel L21 = A21 W^T and the right-looking update of the trailing lower
// triangle, C -= L21 L21^T, as streams of 64-row tiles (stream_tiles; the
// update one stream per 128-column tile of C); then the zeros above the
// diagonal.  The matrix lives in L, where the caches keep the part the
// workgroup works on; every tile is staged through LDS.  Phases are
// separated by workgroup barriers only: all waves of a workgroup share one
// CU and its L1, so after a barrier a wave's loads see the global stores
// the others made before it.  p.first: step 0 reads the old values it
// first touches from A (at A's own row stride; no copy of tril(A) into L)
// and the zeros above the diagonal are stored at the end;
// p.trailing_from_a: the same, except that the first 128 columns are
// already in L; otherwise L already holds tril(A), zeros included.
template <bool VEC, bool SPLIT>
__device__ __forceinline__ void fused_matrix(const Batch& p, int64_t b, float* __restrict__ Wb, DiagLds& sm) {
  const Mat M = mat_at(p, b);
  const int n = M.n;
  float* const lds = reinterpret_cast<float*>(&sm);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const bool first = p.first && k0 == 0;
    const bool cfirst = (p.first || p.trailing_from_a) && k0 == 0;  // trailing old values from A
    __syncthreads();  // the previous step's tiles are done with the LDS
    // One form of the block code for every step (a second inlined copy
    // makes the kernel spill): the last block also forms its inverse,
    // unused, and its upper right quadrant is left to zero_upper_one.
    diag_block<false, false, false>(p, b, k0, Wb, sm, 0, 0, first, true);
    if (k0 + NB >= n) break;
    // The thread index hidden from the optimiser (as in diag_block): no
    // tile addresses hoisted out of this loop across the block code.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    // first touch: A21 and the old values of the trailing triangle from A,
    // with A's own row stride
    const float* As = first ? M.A : M.L;
    const int64_t lda = first ? M.lda : M.ld;
    stream_tiles<1, VEC, SPLIT>(As, lda, k0, k0 + NB, n, Wb, NB, 0, NB, 0, M.L, M.L, M.ld, M.ld, k0, k0 + NB, lds,
                                tid);
    const float* Cs = cfirst ? M.A : M.L;
    const int64_t ldcs = cfirst ? M.lda : M.ld;
    for (int j0 = k0 + NB; j0 < n; j0 += 128)
      stream_tiles<0, VEC, SPLIT>(M.L, M.ld, k0, j0, n, M.L, M.ld, j0, n, k0, M.L, Cs, M.ld, ldcs, j0, n, lds, tid);
  }
  if (p.first || p.trailing_from_a) zero_upper_one(M);
}

// One workgroup per matrix (grid.x = the batch; workgroups are dispatched
// in block order, so a ragged list sorted largest first starts its largest
// matrices first).  W of matrix b at winv + 128 * 128 * b.
template <bool VEC, bool SPLIT>
__global__ __launch_bounds__(256) void k_potrf_fused(Batch p, float* __restrict__ winv) {
  __shared__ DiagLds sm;
  const int64_t b = blockIdx.x;
  fused_matrix<VEC, SPLIT>(p, b, winv + b * (NB * NB), sm);
}

}  // namespace

int launch_fused(const Batch& p, float* winv, bool split, hipStream_t s) {
  if (p.batch <= 0) return 0;
  if (p.batch > 0x7fffffffLL) {
    set_error("k_potrf_fused: batch %lld exceeds the grid", (long long)p.batch);
    return CHOL_E_ARG;
  }
  const bool vec = vec_ok(p);
  const dim3 g((unsigned)p.batch), b(256);
  if (vec && split)
    hipLaunchKernelGGL((k_potrf_fused<true, true>), g, b, 0, s, p, winv);
  else if (vec)
    hipLaunchKernelGGL((k_potrf_fused<true, false>), g, b, 0, s, p, winv);
  else if (split)
    hipLaunchKernelGGL((k_potrf_fused<false, true>), g, b, 0, s, p, winv);
  else
    hipLaunchKernelGGL((k_potrf_fused<false, false>), g, b, 0, s, p, winv);
  return launch_status("k_potrf_fused");
}

int64_t diag_bulk_tiles(const Batch& p, const DiagBulk& u) {
  return (int64_t)rect_tiles<64, 128>(p.n, u.c0, u.c1) * p.batch;
}

int launch_diag_factor(const Batch& p, int k0, float* winv, hipStream_t s, const DiagBulk* bulk,
                       const DiagZero* zero) {
  // The block that reaches column n needs no inverse (no panel follows).
  // p.first (k0 = 0): staged from A, at A's own row stride.  (Ragged
  // batches: every matrix must end at this block for LAST.)
  const bool last = k0 + kNB >= p.n;
  const bool from_a = p.first;
  const dim3 b(256);
  dim3 g((unsigned)p.batch);
  int zr0 = 0, zr1 = 0, nz = 0;
  if (zero && zero->r1 > zero->r0 && zero->wgs > 0) {
    zr0 = zero->r0;
    zr1 = zero->r1;
    nz = zero->wgs;
    g.x += (unsigned)nz;
  }
  int bc0 = 0, bc1 = 0, bk0 = 0, bt = 1, bfirst = 0;
  if (bulk) {  // dense, float4-aligned (vec_ok), not the first block
    bt = rect_tiles<64, 128>(p.n, bulk->c0, bulk->c1);
    bc0 = bulk->c0;
    bc1 = bulk->c1;
    bk0 = bulk->k0;
    bfirst = bulk->first ? 1 : 0;
    g.x += (unsigned)(bt * p.batch);
  }
#define CHOL_DIAG_LAUNCH(FA, LA, BU) \
  hipLaunchKernelGGL((k_diag_factor<FA, LA, BU>), g, b, 0, s, p, k0, winv, bc0, bc1, bk0, bt, bfirst, zr0, zr1, nz)
#define CHOL_DIAG_LAUNCH_PRE(LA)                                                                                  \
  hipLaunchKernelGGL((k_diag_factor<false, LA, 2, true>), g, b, 0, s, p, k0, winv, bc0, bc1, bk0, bt, bfirst, zr0, \
                     zr1, nz)
  const int bu = bulk ? (bulk->split ? 2 : 1) : 0;
  if (bulk && bulk->pre) {
    if (!bulk->split || from_a || bulk->c0 != k0 || !vec_ok(p)) {
      set_error("k_diag_factor: in-kernel update needs the split form, c0 = k0 and float4 access");
      return CHOL_E_ARG;
    }
    if (last) {
      CHOL_DIAG_LAUNCH_PRE(true);
    } else {
      CHOL_DIAG_LAUNCH_PRE(false);
    }
  } else if (from_a && last) {
    CHOL_DIAG_LAUNCH(true, true, 0);
  } else if (from_a) {
    CHOL_DIAG_LAUNCH(true, false, 0);
  } else if (last) {
    if (bu == 2) {
      CHOL_DIAG_LAUNCH(false, true, 2);
    } else if (bu == 1) {
      CHOL_DIAG_LAUNCH(false, true, 1);
    } else {
      CHOL_DIAG_LAUNCH(false, true, 0);
    }
  } else {
    if (bu == 2) {
      CHOL_DIAG_LAUNCH(false, false, 2);
    } else if (bu == 1) {
      CHOL_DIAG_LAUNCH(false, false, 1);
    } else {
      CHOL_DIAG_LAUNCH(false, false, 0);
    }
  }
#undef CHOL_DIAG_LAUNCH
#undef CHOL_DIAG_LAUNCH_PRE
  return launch_status("k_diag_factor");
}

}  // namespace chol
