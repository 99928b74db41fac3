// chol_block.h — the 64x64 block factorisation shared by the diagonal-block
// kernel (chol_diag.hip) and the n = 64 whole-matrix kernel (chol_small.hip):
// 16-column register panels with v_readlane pivot broadcasts, each brought
// up to date by the earlier panels on MFMA first, on a block resident in LDS.
#pragma once

#include "chol_common.h"

namespace chol {

typedef float floatx4_b __attribute__((ext_vector_type(4)));

// C(16x16) += X Y^T, X and Y rows read as float4 at pitch P (floats), K = 16,
// on v_mfma_f32_16x16x4_f32.  Lane (li, lg) holds C[4lg+v][li] in acc[v].
template <int P>
__device__ __forceinline__ floatx4_b mma_nt_p(const float* X, const float* Y, floatx4_b acc, int li, int lg) {
  const float4 xa = *reinterpret_cast<const float4*>(X + li * P + 4 * lg);
  const float4 yb = *reinterpret_cast<const float4*>(Y + li * P + 4 * lg);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, yb.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, yb.y, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, yb.z, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, yb.w, acc, 0, 0, 0);
  return acc;
}

// One column panel S of factor64 (below): the register chain on rows
// c0..d+63 (c0 = d + PW S) and the MFMA updates around it.  Panels before
// LF update right-looking: after its chain, panel S < LF applies its rank-PW
// update to all later columns at once.  From LF on, updates are
// left-looking: before its chain, panel S takes the updates of panels
// LF..S-1 on its own columns only (K = PW (S - LF), one MFMA sum and one
// read-modify-write per 16x16 tile).
//   LF = 0 (all left-looking): each tile is written once in all and panel 0
//     updates nothing, where a right-looking panel 0 updates six tiles: the
//     shorter path for a wavefront that factors alone, or whose partners
//     are idle in the first panel.
//   LF = 1: panel 0's update of all later columns fills the first panel,
//     in which the diagonal-block kernel's other wavefronts are busy, and
//     panel 1's update moves out of the second panel into the last two.
//   LF = 4: all right-looking.
// Accumulates the 1-based failing pivot in bad.
template <int PW, int P, int S, int LF>
__device__ __forceinline__ void factor64_step(float* Ls, int d, int lane, int& bad) {
  const int li = lane & 15, lg = lane >> 4;
  const int c0 = d + PW * S;
  if constexpr (S > LF) {
    constexpr int t0 = PW * S / 16;  // the panel's first tile row
    constexpr int NT = 4 - t0;       // its tiles (t0 + t, t0), t < NT
    floatx4_b acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = floatx4_b{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = PW * LF; kk < PW * S; kk += 16)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[t] = mma_nt_p<P>(&Ls[(d + 16 * (t0 + t)) * P + d + kk], &Ls[c0 * P + d + kk], acc[t], li, lg);
    // all old values first, then every element written back (the diagonal
    // tile's upper part unchanged): no branch, one LDS round trip.  The
    // block is this wavefront's alone while it factors it.
    float old[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) old[t][v] = Ls[(d + 16 * (t0 + t) + 4 * lg + v) * P + c0 + li];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * t + 4 * lg + v;
        Ls[(d + 16 * t0 + i) * P + c0 + li] = (t != 0 || li <= i) ? old[t][v] - acc[t][v] : old[t][v];
      }
    wave_lds_sync();
  }
  float* row = &Ls[(d + lane) * P + c0];
  float a[PW];
#pragma unroll
  for (int j = 0; j < PW; j += 4) {
    const float4 v = *reinterpret_cast<const float4*>(row + j);
    a[j] = v.x;
    a[j + 1] = v.y;
    a[j + 2] = v.z;
    a[j + 3] = v.w;
  }
  const int b = reg_factor_rl<PW>(a, lane, PW * S);
  if (bad == 0 && b != 0) bad = PW * S + b;
  if (lane >= PW * S) {
#pragma unroll
    for (int j = 0; j < PW; j += 4)
      *reinterpret_cast<float4*>(row + j) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
  }
  wave_lds_sync();
  // right-looking: lower 16x16 tiles (ti >= tj) of rows / columns
  // [PW (S+1), 64); reads only panel columns, writes only later ones
  constexpr int t1 = PW * (S + 1) / 16;
  if constexpr (S < LF && t1 < 4) {
    floatx4_b acc[10];
    int nt = 0;
#pragma unroll
    for (int ti = t1; ti < 4; ++ti)
#pragma unroll
      for (int tj = t1; tj <= ti; ++tj) {
        floatx4_b c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < PW; kk += 16)
          c = mma_nt_p<P>(&Ls[(d + 16 * ti) * P + c0 + kk], &Ls[(d + 16 * tj) * P + c0 + kk], c, li, lg);
        acc[nt++] = c;
      }
    float old[10][4];
    nt = 0;
#pragma unroll
    for (int ti = t1; ti < 4; ++ti)
#pragma unroll
      for (int tj = t1; tj <= ti; ++tj) {
#pragma unroll
        for (int v = 0; v < 4; ++v) old[nt][v] = Ls[(d + 16 * ti + 4 * lg + v) * P + d + 16 * tj + li];
        ++nt;
      }
    nt = 0;
#pragma unroll
    for (int ti = t1; ti < 4; ++ti)
#pragma unroll
      for (int tj = t1; tj <= ti; ++tj) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * ti + 4 * lg + v, j = 16 * tj + li;
          Ls[(d + i) * P + d + j] = (ti != tj || j <= i) ? old[nt][v] - acc[nt][v] : old[nt][v];
        }
        ++nt;
      }
    wave_lds_sync();
  }
}

// One wavefront factors the 64x64 block of Ls (pitch P) at (d, d), blocked
// by PW columns: for each column panel [c0, c0+PW), lane = row,
//   1. the panel's lower 16x16 tiles take the update of all earlier panels
//      on v_mfma_f32_16x16x4_f32 (left-looking);
//   2. reg_factor_rl<PW> on rows c0..d+63 yields the panel's diagonal factor
//      and, for the rows below it, its part of L (the panel's own
//      triangular solve), broadcasting pivot columns with v_readlane.
// A pivot step then costs a few dependent VALU operations plus
// 2 (PW-1-k) issue slots, where a 64-wide register chain spends an LDS
// round trip and 63-k row updates.  Returns the 1-based first failing
// pivot (0 if none).
template <int PW, int P>
__device__ __forceinline__ int factor64(float* Ls, int d, int lane) {
  static_assert(PW == 16, "four 16-column panels");
  int bad = 0;
  factor64_step<PW, P, 0, 0>(Ls, d, lane, bad);
  factor64_step<PW, P, 1, 0>(Ls, d, lane, bad);
  factor64_step<PW, P, 2, 0>(Ls, d, lane, bad);
  factor64_step<PW, P, 3, 0>(Ls, d, lane, bad);
  return bad;
}

// Two 32x32 matrices per wavefront, staged in LDS (lane l holds row l % 32 of
// matrix l / 32 at st + l*P), factored as two 16-column panels:
//   1. reg_factor<16> on rows 0..31 of each matrix (pivot columns exchanged
//      through cb, 32 floats per matrix): L11 and L21 of the left panel;
//   2. A22 -= L21 L21^T for each matrix (one 16x16 tile, K = 16, MFMA);
//   3. reg_factor<16> on rows 16..31: L22.
// Against one 32-pivot chain this halves the row updates (each row spans
// 16 columns per chain, not up to 32) for one MFMA tile per matrix.
// Returns the lane's matrix's 1-based first failing pivot (0 if none).
template <int P>
__device__ __forceinline__ int factor32_pair(float* st, float* cb, int lane) {
  const int m = lane >> 5, r = lane & 31, li = lane & 15, lg = (lane >> 4) & 3;
  float* row = st + lane * P;
  float* cbm = cb + 32 * m;
  int bad;
  {
    float a[16];
#pragma unroll
    for (int j = 0; j < 16; j += 4) {
      const float4 v = *reinterpret_cast<const float4*>(row + j);
      a[j] = v.x;
      a[j + 1] = v.y;
      a[j + 2] = v.z;
      a[j + 3] = v.w;
    }
    bad = reg_factor<16>(a, r, cbm, true);
#pragma unroll
    for (int j = 0; j < 16; j += 4) *reinterpret_cast<float4*>(row + j) = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    float* t = st + (32 * q + 16) * P;
    const floatx4_b c = mma_nt_p<P>(t, t, floatx4_b{0.f, 0.f, 0.f, 0.f}, li, lg);
    float o[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = t[(4 * lg + v) * P + 16 + li];
#pragma unroll
    for (int v = 0; v < 4; ++v) t[(4 * lg + v) * P + 16 + li] = (li <= 4 * lg + v) ? o[v] - c[v] : o[v];
  }
  wave_lds_sync();
  {
    // rows 16..31 (r1 = r - 16); lanes with r < 16 take no update and do not write
    const int r1 = r - 16;
    float b[16];
#pragma unroll
    for (int j = 0; j < 16; j += 4) {
      const float4 v = *reinterpret_cast<const float4*>(row + 16 + j);
      b[j] = v.x;
      b[j + 1] = v.y;
      b[j + 2] = v.z;
      b[j + 3] = v.w;
    }
    const int bad2 = reg_factor<16>(b, r1, cbm, r1 >= 0);
    if (r1 >= 0) {
#pragma unroll
      for (int j = 0; j < 16; j += 4)
        *reinterpret_cast<float4*>(row + 16 + j) = make_float4(b[j], b[j + 1], b[j + 2], b[j + 3]);
    }
    if (bad == 0 && bad2 != 0) bad = 16 + bad2;
  }
  wave_lds_sync();
  return bad;
}

}  // namespace chol
