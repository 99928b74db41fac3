// chol_bwd.hip — triangular inverse and the adjoint (backward) of the
// factorisation, on FP32 matrix cores.
//
//   tri_inverse : X = L^-1 by recursive doubling.  64x64 diagonal blocks are
//                 inverted by one wavefront each (lane = column, forward
//                 substitution); then for h = 64, 128, ... each pair of
//                 h-blocks [A 0; B C] gets its off-diagonal block
//                 -C^-1 (B A^-1) from two batched GEMMs.
//   backward    : Abar = L^-T S L^-1 with S = 1/2 (Phi(P) + Phi(P)^T),
//                 P = L^T Lbar, Phi = lower triangle with halved diagonal.
//                 Three GEMMs (P, T = L^-T S, Abar = T L^-1) and two
//                 elementwise passes; L^-1 is recomputed unless given.
//
// The reference has no backward and reaches its block inverses through
// Triton leaves + cuBLAS bmm trees (_exp057_tri_inv_leaf32_kernel and
// _exp061_leaf_inverse, submission.py:3434-3466, :3562-3603); this is the
// same recursive-doubling identity written as gfx950 kernels.  The adjoint
// formula is the one torch's autograd uses for torch.linalg.cholesky.
#include "chol_common.h"

namespace chol {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GKC = 32;      // K chunk
constexpr int GKP = GKC + 4; // LDS row pitch
constexpr int kThreads = 256;

// Batched C = alpha * Aop Bop^T + beta * C over an (M x N) output, K deep.
//   Aop[m][k] = A[m*lda + k] (A_KC) or A[k*lda + m] (!A_KC)
//   Bop[n][k] = B[n*ldb + k] (B_KC) or B[k*ldb + n] (!B_KC)
// Batch index z = blockIdx.z splits as (z / nb2, z % nb2) with separate
// strides, so one launch covers matrices x sub-blocks.
struct Gemm {
  const float* A;
  const float* B;
  float* C;
  int64_t lda, ldb, ldc;
  int M, N, K;
  float alpha, beta;
  int64_t sA1, sA2, sB1, sB2, sC1, sC2;
  int nb2;
  // Triangular operands: the k range of an output tile (rows m0.., columns
  // n0..) is cut to k >= m0 (klo_m: Aop[m][k] = 0 for k < m), k >= n0
  // (klo_n: Bop[n][k] = 0 for k < n) and k < m0 + G (khi_m: Aop[m][k] = 0
  // for k > m); lower_tiles: tiles wholly above the diagonal are not formed.
  int klo_m, klo_n, khi_m, lower_tiles;
  // Set by launch_gemm: every operand address of the launch is 16-byte
  // aligned and a multiple of 4 floats apart (float4 global loads).
  int vec;
};

// Loads a G x GKC tile of op(X)[row][k] into registers, G * GKC / 256
// floats per thread.
template <int G, bool KC_CONTIG>
__device__ __forceinline__ void gload(float (&r)[G * GKC / kThreads], const float* base, int64_t ld, int row0,
                                      int rlim, int k0, int klim, int tid, bool vec) {
  constexpr int PER = G * GKC / kThreads;  // 8 (G = 64) or 16 (G = 128)
  if constexpr (KC_CONTIG) {
    // thread: row = tid / (GKC / PER), k = (tid % (GKC / PER)) * PER .. +PER-1
    constexpr int TPR = GKC / PER;
    const int gr = row0 + tid / TPR;
    const int kk = (tid % TPR) * PER;
    const float* src = base + (int64_t)gr * ld + k0 + kk;
#pragma unroll
    for (int u = 0; u < PER; u += 4) {
      const int gk = k0 + kk + u;
      if (vec && gr < rlim && gk + 3 < klim) {
        const float4 v = *reinterpret_cast<const float4*>(src + u);
        r[u] = v.x;
        r[u + 1] = v.y;
        r[u + 2] = v.z;
        r[u + 3] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) r[u + e] = (gr < rlim && gk + e < klim) ? src[u + e] : 0.f;
      }
    }
  } else {
    // thread: k = tid / (G / PER), rows (tid % (G / PER)) * PER .. +PER-1
    // (contiguous in memory)
    constexpr int TPK = G / PER;
    const int kk = tid / TPK;
    const int rr = (tid % TPK) * PER;
    const int gk = k0 + kk;
    const float* src = base + (int64_t)gk * ld + row0 + rr;
#pragma unroll
    for (int u = 0; u < PER; u += 4) {
      const int gr = row0 + rr + u;
      if (vec && gk < klim && gr + 3 < rlim) {
        const float4 v = *reinterpret_cast<const float4*>(src + u);
        r[u] = v.x;
        r[u + 1] = v.y;
        r[u + 2] = v.z;
        r[u + 3] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) r[u + e] = (gk < klim && gr + e < rlim) ? src[u + e] : 0.f;
      }
    }
  }
}

template <int G, bool KC_CONTIG>
__device__ __forceinline__ void gstore(float* lds, const float (&r)[G * GKC / kThreads], int tid) {
  constexpr int PER = G * GKC / kThreads;
  if constexpr (KC_CONTIG) {
    constexpr int TPR = GKC / PER;
    const int rr = tid / TPR;
    const int kk = (tid % TPR) * PER;
#pragma unroll
    for (int u = 0; u < PER; ++u) lds[rr * GKP + kk + u] = r[u];
  } else {
    constexpr int TPK = G / PER;
    const int kk = tid / TPK;
    const int rr = (tid % TPK) * PER;
#pragma unroll
    for (int u = 0; u < PER; ++u) lds[(rr + u) * GKP + kk] = r[u];
  }
}

// G x G output tile per workgroup of 4 wavefronts, each a (G/2) x (G/2)
// quarter of 16x16 MFMA blocks: G = 128 for large products (four times the
// operand reuse per LDS read of G = 64), G = 64 where 128-tiles would leave
// CUs idle.
template <int G, bool A_KC, bool B_KC>
__global__ __launch_bounds__(kThreads) void k_gemm(Gemm g) {
  constexpr int F = G / 32;  // 16x16 blocks per wave side
  constexpr int PER = G * GKC / kThreads;
  __shared__ __align__(16) float As[2][G * GKP];
  __shared__ __align__(16) float Bs[2][G * GKP];
  const int z = blockIdx.z;
  const int z1 = z / g.nb2, z2 = z - (z / g.nb2) * g.nb2;
  const float* A = g.A + z1 * g.sA1 + z2 * g.sA2;
  const float* B = g.B + z1 * g.sB1 + z2 * g.sB2;
  float* C = g.C + z1 * g.sC1 + z2 * g.sC2;
  // The slow grid dimension walks the tiles whose K range shrinks along it
  // (rows for klo_m, columns for klo_n), so the longest tiles of a matrix
  // are dispatched first and the short ones fill the tail.
  const int m0 = (g.klo_n ? blockIdx.x : blockIdx.y) * G;
  const int n0 = (g.klo_n ? blockIdx.y : blockIdx.x) * G;
  if (g.lower_tiles && n0 >= m0 + G) return;
  // m0 and n0 are multiples of G, so the range starts on a chunk boundary.
  int k_lo = 0, k_hi = g.K;
  if (g.klo_m) k_lo = m0 > k_lo ? m0 : k_lo;
  if (g.klo_n) k_lo = n0 > k_lo ? n0 : k_lo;
  if (g.khi_m) k_hi = m0 + G < k_hi ? m0 + G : k_hi;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * (G / 2);
  const int wn = (wave & 1) * (G / 2);
  const int li = lane & 15;
  const int lg = lane >> 4;

  floatx4 acc[F][F];
#pragma unroll
  for (int a = 0; a < F; ++a)
#pragma unroll
    for (int b = 0; b < F; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};

  float ra[PER], rb[PER];
  const int nch = k_hi > k_lo ? (k_hi - k_lo + GKC - 1) / GKC : 0;
  if (nch > 0) {
    gload<G, A_KC>(ra, A, g.lda, m0, g.M, k_lo, k_hi, tid, g.vec != 0);
    gload<G, B_KC>(rb, B, g.ldb, n0, g.N, k_lo, k_hi, tid, g.vec != 0);
    gstore<G, A_KC>(As[0], ra, tid);
    gstore<G, B_KC>(Bs[0], rb, tid);
    __syncthreads();
  }
  for (int ch = 0; ch < nch; ++ch) {
    const int cur = ch & 1;
    const bool more = ch + 1 < nch;
    if (more) {
      gload<G, A_KC>(ra, A, g.lda, m0, g.M, k_lo + (ch + 1) * GKC, k_hi, tid, g.vec != 0);
      gload<G, B_KC>(rb, B, g.ldb, n0, g.N, k_lo + (ch + 1) * GKC, k_hi, tid, g.vec != 0);
    }
#pragma unroll
    for (int ks = 0; ks < GKC; ks += 16) {
      float4 fa[F], fb[F];
#pragma unroll
      for (int a = 0; a < F; ++a)
        fa[a] = *reinterpret_cast<const float4*>(&As[cur][(wm + a * 16 + li) * GKP + ks + 4 * lg]);
#pragma unroll
      for (int b = 0; b < F; ++b)
        fb[b] = *reinterpret_cast<const float4*>(&Bs[cur][(wn + b * 16 + li) * GKP + ks + 4 * lg]);
      // k-step outermost: consecutive MFMAs write different accumulators.
#pragma unroll
      for (int a = 0; a < F; ++a)
#pragma unroll
        for (int b = 0; b < F; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].x, fb[b].x, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < F; ++a)
#pragma unroll
        for (int b = 0; b < F; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].y, fb[b].y, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < F; ++a)
#pragma unroll
        for (int b = 0; b < F; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].z, fb[b].z, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < F; ++a)
#pragma unroll
        for (int b = 0; b < F; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a].w, fb[b].w, acc[a][b], 0, 0, 0);
    }
    if (more) {
      gstore<G, A_KC>(As[cur ^ 1], ra, tid);
      gstore<G, B_KC>(Bs[cur ^ 1], rb, tid);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < F; ++a)
#pragma unroll
    for (int b = 0; b < F; ++b) {
      const int j = n0 + wn + b * 16 + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = m0 + wm + a * 16 + 4 * lg + v;
        if (i < g.M && j < g.N) {
          float* dst = C + (int64_t)i * g.ldc + j;
          const float r = g.alpha * acc[a][b][v];
          *dst = (g.beta == 0.f) ? r : fmaf(g.beta, *dst, r);
        }
      }
    }
}

template <int G>
void gemm_launch(const Gemm& g, bool a_kc, bool b_kc, int batch, hipStream_t s) {
  const unsigned tn = (unsigned)((g.N + G - 1) / G), tm = (unsigned)((g.M + G - 1) / G);
  dim3 grid(g.klo_n ? tm : tn, g.klo_n ? tn : tm, (unsigned)batch);
  if (a_kc && b_kc)
    hipLaunchKernelGGL((k_gemm<G, true, true>), grid, dim3(kThreads), 0, s, g);
  else if (a_kc)
    hipLaunchKernelGGL((k_gemm<G, true, false>), grid, dim3(kThreads), 0, s, g);
  else if (b_kc)
    hipLaunchKernelGGL((k_gemm<G, false, true>), grid, dim3(kThreads), 0, s, g);
  else
    hipLaunchKernelGGL((k_gemm<G, false, false>), grid, dim3(kThreads), 0, s, g);
}

// zcount products (matrices x nb2 sub-blocks) on grid.z, which holds at most
// 65535: launches of whole matrices each, the operand pointers advanced by
// the matrices before them.
int launch_gemm(const Gemm& g0, bool a_kc, bool b_kc, int64_t zcount, hipStream_t s) {
  if (g0.M <= 0 || g0.N <= 0 || zcount <= 0) return 0;
  Gemm g = g0;
  const uintptr_t addr = reinterpret_cast<uintptr_t>(g.A) | reinterpret_cast<uintptr_t>(g.B);
  const int64_t strides = g.lda | g.ldb | g.sA1 | g.sA2 | g.sB1 | g.sB2;
  g.vec = (addr & 15) == 0 && (strides & 3) == 0;
  const int64_t per = (int64_t)(65535 / g.nb2) * g.nb2;
  if (g.nb2 <= 0 || per <= 0 || zcount % g.nb2 != 0) {
    set_error("launch_gemm: %lld products do not split into matrices of %d", (long long)zcount, g.nb2);
    return CHOL_E_ARG;
  }
  // 128-tiles once they alone fill the GPU's 256 CUs twice over (decided on
  // the whole product count, so the tile shape does not depend on the split).
  const int64_t tiles128 = (int64_t)((g.N + 127) / 128) * ((g.M + 127) / 128) * zcount;
  for (int64_t z0 = 0; z0 < zcount; z0 += per) {
    Gemm c = g;
    const int64_t m0 = z0 / g.nb2;
    c.A += m0 * g.sA1;
    c.B += m0 * g.sB1;
    c.C += m0 * g.sC1;
    const int cnt = (int)((zcount - z0) < per ? (zcount - z0) : per);
    if (tiles128 >= 512)
      gemm_launch<128>(c, a_kc, b_kc, cnt, s);
    else
      gemm_launch<64>(c, a_kc, b_kc, cnt, s);
    const int rc = launch_status("k_gemm");
    if (rc) return rc;
  }
  return 0;
}

// Inverse of the 64x64 diagonal blocks of L (rows/cols past n: identity),
// one wavefront per (matrix, block); lane c forward-substitutes column c.
// Writes the blocks into X (zeros above the diagonal inside each block).
// Lane c keeps its column of the inverse in LDS (Ws[i][c], an address of its
// own, so no barrier is needed between its writes and reads): a register
// array indexed by the unrolled substitution kept too many values live and
// spilled.
__global__ __launch_bounds__(64) void k_leaf_inverse(const float* __restrict__ L, float* __restrict__ X, int n,
                                                     int nblk) {
  constexpr int NB = 64, P = NB + 4;
  __shared__ __align__(16) float Ls[NB * P];
  __shared__ float Ws[NB * NB];
  const int c = threadIdx.x;
  const int64_t b = blockIdx.x / nblk;
  const int k0 = (blockIdx.x % nblk) * NB;
  const int nb = min(NB, n - k0);
  const float* Lb = L + b * (int64_t)n * n;
  float* Xb = X + b * (int64_t)n * n;
  // stage the block (lane = column), identity past nb; 16 rows' loads in
  // flight at a time
#pragma unroll 16
  for (int r = 0; r < NB; ++r) {
    float v;
    if (r < nb)
      v = (c <= r && c < nb) ? Lb[(int64_t)(k0 + r) * n + k0 + c] : 0.f;
    else
      v = (c == r) ? 1.f : 0.f;
    Ls[r * P + c] = v;
  }
  __syncthreads();
  for (int i = 0; i < NB; ++i) {
    const float* Li = &Ls[i * P];  // row i, read by every lane (broadcast)
    float s0 = (i == c) ? 1.f : 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int k = 0;
    for (; k + 4 <= i; k += 4) {
      const float4 l = *reinterpret_cast<const float4*>(Li + k);
      s0 = fmaf(-l.x, Ws[k * NB + c], s0);
      s1 = fmaf(-l.y, Ws[(k + 1) * NB + c], s1);
      s2 = fmaf(-l.z, Ws[(k + 2) * NB + c], s2);
      s3 = fmaf(-l.w, Ws[(k + 3) * NB + c], s3);
    }
    for (; k < i; ++k) s0 = fmaf(-Li[k], Ws[k * NB + c], s0);
    Ws[i * NB + c] = ((s0 + s1) + (s2 + s3)) * __builtin_amdgcn_rcpf(Li[i]);
  }
  if (c < nb)
    for (int i = 0; i < nb; ++i) Xb[(int64_t)(k0 + i) * n + k0 + c] = Ws[i * NB + c];
}

__global__ void k_fill_zero(float* __restrict__ p, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

// S = 1/2 (Phi(P) + Phi(P)^T) from the lower triangle of P, in place
// (P becomes the full symmetric S).
// In place, over 32x32 tile pairs (I, J), I >= J, of matrix blockIdx.y,
// each pair owned by one workgroup and staged through LDS so that the
// transposed half is read and written a row at a time:
//   PHI : S = 1/2 (Phi(P) + Phi(P)^T) from P's lower triangle;
//   !PHI: Abar = 1/2 (Abar + Abar^T) (removes the rounding asymmetry of the
//         two products).
template <bool PHI>
__global__ __launch_bounds__(256) void k_sym_tiles(float* __restrict__ X, int n) {
  __shared__ float Ta[32][33];
  __shared__ float Tb[PHI ? 1 : 32][33];
  const int nt = (n + 31) / 32;
  const int I = blockIdx.x / nt, J = blockIdx.x - (blockIdx.x / nt) * nt;
  if (J > I) return;
  float* M = X + (int64_t)blockIdx.y * n * n;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int a = ty + 8 * r;
    const int i = I * 32 + a, j = J * 32 + tx;
    Ta[a][tx] = (i < n && j < n) ? M[(int64_t)i * n + j] : 0.f;
    if constexpr (!PHI) {
      const int i2 = J * 32 + a, j2 = I * 32 + tx;
      Tb[a][tx] = (i2 < n && j2 < n) ? M[(int64_t)i2 * n + j2] : 0.f;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int a = ty + 8 * r, c = tx;
    // tile (I, J): element (I*32 + a, J*32 + c)
    const int i = I * 32 + a, j = J * 32 + c;
    float v;
    if constexpr (PHI)
      v = 0.5f * (i >= j ? Ta[a][c] : Ta[c][a]);  // i < j only inside a diagonal tile
    else
      v = 0.5f * (Ta[a][c] + Tb[c][a]);
    if (i < n && j < n) M[(int64_t)i * n + j] = v;
    if (I != J) {
      // tile (J, I): element (J*32 + a, I*32 + c), the mirror of (I*32 + c, J*32 + a)
      const int i2 = J * 32 + a, j2 = I * 32 + c;
      const float w = PHI ? 0.5f * Ta[c][a] : 0.5f * (Tb[a][c] + Ta[c][a]);
      if (i2 < n && j2 < n) M[(int64_t)i2 * n + j2] = w;
    }
  }
}

// k_sym_tiles over a batch, at most 65535 matrices (grid.y) per launch.
template <bool PHI>
int sym_tiles(float* X, int64_t batch, int n, hipStream_t s) {
  const int nt = (n + 31) / 32;
  const int64_t nn = (int64_t)n * n;
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t cnt = (batch - b0) < 65535 ? (batch - b0) : 65535;
    hipLaunchKernelGGL((k_sym_tiles<PHI>), dim3((unsigned)(nt * nt), (unsigned)cnt), dim3(256), 0, s, X + b0 * nn, n);
    const int rc = launch_status("k_sym_tiles");
    if (rc) return rc;
  }
  return 0;
}

unsigned grid1d(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

// X = L^-1 for a contiguous batch; ws holds one n x n scratch per matrix.
int tri_inverse(const float* L, float* X, int64_t batch, int n, float* ws, hipStream_t s) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(k_fill_zero, dim3(grid1d(batch * nn)), dim3(256), 0, s, X, batch * nn);
  int rc = launch_status("k_fill_zero");
  if (rc) return rc;
  const int nblk = (n + 63) / 64;
  hipLaunchKernelGGL(k_leaf_inverse, dim3((unsigned)(batch * nblk)), dim3(64), 0, s, L, X, n, nblk);
  if ((rc = launch_status("k_leaf_inverse"))) return rc;
  for (int h = 64; h < n; h *= 2) {
    // pairs [2ih, 2ih+2h): A = [2ih, 2ih+h), C = [2ih+h, min(2ih+2h, n))
    const int pairs = (n + 2 * h - 1) / (2 * h);
    for (int pass = 0; pass < 2; ++pass) {
      // Full pairs in one batched launch; the last pair may be partial.
      const int full = (n / (2 * h));
      const int rest = pairs - full;
      for (int part = 0; part < 2; ++part) {
        const int cnt = part == 0 ? full : rest;
        if (cnt <= 0) continue;
        const int i0 = part == 0 ? 0 : full;
        const int a0 = 2 * i0 * h;          // first row/col of A
        const int c0 = a0 + h;              // first row/col of C
        if (c0 >= n) continue;
        const int hc = (n - c0) < h ? (n - c0) : h;  // rows of C
        Gemm g{};
        g.nb2 = cnt;
        if (pass == 0) {
          // ws[C rows, A cols] = L[C, A] * X[A, A]   (B^T A^-1 ... as NN)
          g.A = L + (int64_t)c0 * n + a0;  g.lda = n;  // Aop[m][k] = L[c0+m][a0+k]
          g.B = X + (int64_t)a0 * n + a0;  g.ldb = n;  // Bop[j][k] = X[a0+k][a0+j]
          g.C = ws + (int64_t)c0 * n + a0; g.ldc = n;
          g.M = hc; g.N = h; g.K = h;
          g.alpha = 1.f; g.beta = 0.f;
          g.klo_n = 1;  // X[A, A] lower triangular
        } else {
          // X[C, A] = -X[C, C] * ws[C, A]
          g.A = X + (int64_t)c0 * n + c0;  g.lda = n;  // Aop[m][k] = X[c0+m][c0+k]
          g.B = ws + (int64_t)c0 * n + a0; g.ldb = n;  // Bop[j][k] = ws[c0+k][a0+j]
          g.C = X + (int64_t)c0 * n + a0;  g.ldc = n;
          g.M = hc; g.N = h; g.K = hc;
          g.alpha = -1.f; g.beta = 0.f;
          g.khi_m = 1;  // X[C, C] lower triangular
        }
        g.sA1 = g.sB1 = g.sC1 = nn;
        g.sA2 = g.sB2 = g.sC2 = (int64_t)2 * h * n + 2 * h;  // next pair along the diagonal
        if ((rc = launch_gemm(g, true, false, batch * cnt, s))) return rc;
      }
    }
  }
  return 0;
}

// Abar = L^-T S L^-1, S from P = L^T Lbar.  ws: 2 n x n scratch per matrix
// (+1 more when Linv is recomputed).
int backward(const float* L, const float* Lbar, float* Abar, int64_t batch, int n, const float* Linv_in,
             float* ws, hipStream_t s) {
  const int64_t nn = (int64_t)n * n;
  float* P = ws;                 // batch x n x n
  float* T = ws + batch * nn;    // batch x n x n
  const float* Linv = Linv_in;
  int rc;
  if (!Linv) {
    float* Xi = ws + 2 * batch * nn;
    // the inverse's own scratch reuses T (free until step 3)
    if ((rc = tri_inverse(L, Xi, batch, n, T, s))) return rc;
    Linv = Xi;
  }
  // 1. P = L^T Lbar:  Aop[i][k] = L[k][i], Bop[j][k] = Lbar[k][j].  For
  //    i >= j only k >= i >= j contributes (L is lower), so the upper
  //    triangle of Lbar never reaches the lower triangle of P used below.
  Gemm g{};
  g.nb2 = 1;
  g.sA1 = g.sB1 = g.sC1 = nn;
  g.A = L; g.lda = n; g.B = Lbar; g.ldb = n; g.C = P; g.ldc = n;
  g.M = n; g.N = n; g.K = n; g.alpha = 1.f; g.beta = 0.f;
  g.klo_m = 1;        // L^T upper triangular: k >= i
  g.lower_tiles = 1;  // Phi reads the lower triangle of P only
  if ((rc = launch_gemm(g, false, false, batch, s))) return rc;
  // 2. S = 1/2 (Phi(P) + Phi(P)^T), in place
  if ((rc = sym_tiles<true>(P, batch, n, s))) return rc;
  // 3. T = L^-T S:  Aop[i][k] = Linv[k][i], Bop[j][k] = S[j][k]
  g.A = Linv; g.B = P; g.C = T;
  g.klo_m = 1;  // L^-T upper triangular: k >= i
  g.lower_tiles = 0;
  if ((rc = launch_gemm(g, false, true, batch, s))) return rc;
  // 4. Abar = T L^-1:  Aop[i][k] = T[i][k], Bop[j][k] = Linv[k][j]
  g.A = T; g.B = Linv; g.C = Abar;
  g.klo_m = 0;
  g.klo_n = 1;  // L^-1 lower triangular: Bop[j][k] = Linv[k][j] = 0 for k < j
  if ((rc = launch_gemm(g, true, false, batch, s))) return rc;
  return sym_tiles<false>(Abar, batch, n, s);
}

}  // namespace chol
