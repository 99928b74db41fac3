// chol_syrk256.hip — the large trailing update C -= P P^T in 256x256 tiles.
//
// The outer-block updates of big matrices (K = the outer block width, 256
// and more, thousands of tiles) are bound by how fast the CUs can pull
// operands, not by the matrix cores: a 128x128 tile does 32 useful flops per
// byte of FP32 operand it loads, which at the BF16x3 rate needs ~42 B per
// clock per CU.  A 256x256 tile halves that.  One workgroup of 8 wavefronts
// (2 per SIMD) per tile, each wavefront a 128x64 quarter-column of 16x16
// MFMA blocks (8 x 4, 128 accumulators).
//
// K is staged through LDS in chunks of 32, double buffered (2 x 64 KiB):
// the global loads of chunk c+1 are in flight in registers while chunk c is
// multiplied.  Rows are stored unpadded and XOR-swizzled by 16-byte slot so
// that both the row-wise staging writes and the MFMA fragment reads (lane i
// reads row i of a 16-row block) hit distinct banks:
//   FP32  : 32 floats = 8 slots per row, slot' = slot ^ (row & 7);
//   BF16x3: 32 bf16 = 4 slots per row (hi and lo arrays),
//           slot' = slot ^ ((row >> 1) & 3).
// BF16x3 splits x = hi + lo exactly as the 128-tile form does and
// accumulates hi*hi + hi*lo + lo*hi in FP32.
//
// Tile order: workgroups are dispatched round-robin over the 8 XCDs, each
// with its own L2; the block index is remapped so that every XCD takes a
// contiguous run of the tile list, and the list runs row by row through
// groups of 4 tile columns (decode_tile_grouped), so the ~32 tiles an XCD
// holds at once share 8 row blocks and 4 column blocks of the operand in
// its L2.
//
// Reference counterpart: the trailing SYRK-as-GEMM updates of the blocked
// drivers (submission.py:5163-5184, :2529-2534), run there on tensor cores.
#include "chol_common.h"
#include "chol_tiles.h"

namespace chol {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BT = 256;                  // tile edge
constexpr int NT = 512;                  // threads (8 wavefronts)
constexpr int KC = 32;                   // K chunk
constexpr int PER = BT * KC / 4 / NT;    // float4 per thread per operand and chunk (4)
constexpr int STAGE = 2 * BT * KC;       // floats per stage: A then B (64 KiB)
constexpr int WM = 128, WN = 64;         // per-wave tile
constexpr int FM = WM / 16, FN = WN / 16;

template <bool VEC>
__device__ __forceinline__ void load_chunk(float4 (&reg)[PER], const float* base, int64_t ld, int row0,
                                           int rlim, int k, int tid) {
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int c = tid + u * NT;
    const int r = c >> 3;
    const int q = (c & 7) * 4;
    const int gr = row0 + r;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (gr < rlim) {
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

// FP32 operand: row r at r*32 floats, 16-byte slot q at (q ^ (r & 7)) * 4.
__device__ __forceinline__ void put_f32(float* __restrict__ s, const float4 (&reg)[PER], int tid) {
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int c = tid + u * NT;
    const int r = c >> 3;
    const int q = c & 7;
    *reinterpret_cast<float4*>(s + r * KC + ((q ^ (r & 7)) << 2)) = reg[u];
  }
}

// BF16 operand (hi at s, lo at s + BT*KC): row r at r*32 elements, 16-byte
// slot g (k = 8g..8g+7) at (g ^ ((r >> 1) & 3)) * 8; a float4 (k = 4q..4q+3)
// fills half a slot.
__device__ __forceinline__ void put_split(__bf16* __restrict__ s, const float4 (&reg)[PER], int tid) {
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int c = tid + u * NT;
    const int r = c >> 3;
    const int q = c & 7;
    const float4 v = reg[u];
    uint2 hi, lo;
    split_bf16_pair(v.x, v.y, hi.x, lo.x);
    split_bf16_pair(v.z, v.w, hi.y, lo.y);
    const int off = r * KC + (((q >> 1) ^ ((r >> 1) & 3)) << 3) + ((q & 1) << 2);
    *reinterpret_cast<uint2*>(s + off) = hi;
    *reinterpret_cast<uint2*>(s + BT * KC + off) = lo;
  }
}

template <bool VEC, bool SPLIT>
__global__ __launch_bounds__(NT) void k_syrk256(Batch p, int col0, int col1, int k0, int kl) {
  __shared__ __align__(16) float lds[2 * STAGE];  // 128 KiB

  const int64_t b = blockIdx.y;
  const Mat M = mat_at(p, b);
  const int c1 = (col1 < 0 || col1 > M.n) ? M.n : col1;  // (ragged: clipped to this matrix)
  if (c1 <= col0) return;
  const int TI = (M.n - col0 + BT - 1) / BT;
  const int TJ = (c1 - col0 + BT - 1) / BT;
  const int T = (int)((int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2);
  // XCD-contiguous tile order (xcd_tile); grid.x covers the largest
  // matrix, so ids at or past this matrix's T exit.
  const int bid = blockIdx.x;
  if (bid >= T) return;
  int ti, tj;
  decode_tile_grouped(xcd_tile(bid, T), TI, TJ, 4, ti, tj);
  const int i0 = col0 + BT * ti;
  const int j0 = col0 + BT * tj;
  float* L = M.L;
  const float* Cs = p.first ? M.A : M.L;  // old values of C (first touch: A)
  const int64_t ld = M.ld;
  const int64_t ldcs = p.first ? M.lda : M.ld;  // row stride of Cs

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 2) * WM;
  const int wn = (wave & 3) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  floatx4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int c = 0; c < FN; ++c) acc[a][c] = floatx4{0.f, 0.f, 0.f, 0.f};

  float4 ra[PER], rb[PER];
  auto stage = [&](int buf) {
    float* s = lds + buf * STAGE;
    if constexpr (SPLIT) {
      __bf16* h = reinterpret_cast<__bf16*>(s);
      put_split(h, ra, tid);               // A hi, A lo
      put_split(h + 2 * BT * KC, rb, tid); // B hi, B lo
    } else {
      put_f32(s, ra, tid);
      put_f32(s + BT * KC, rb, tid);
    }
  };

  load_chunk<VEC>(ra, L, ld, i0, M.n, k0, tid);
  load_chunk<VEC>(rb, L, ld, j0, c1, k0, tid);
  stage(0);
  __syncthreads();

  const int nch = kl / KC;
  for (int ch = 0; ch < nch; ++ch) {
    const int cur = ch & 1;
    const bool more = ch + 1 < nch;
    if (more) {
      load_chunk<VEC>(ra, L, ld, i0, M.n, k0 + (ch + 1) * KC, tid);
      load_chunk<VEC>(rb, L, ld, j0, c1, k0 + (ch + 1) * KC, tid);
    }
    const float* s = lds + cur * STAGE;
    if constexpr (SPLIT) {
      const __bf16* Ah = reinterpret_cast<const __bf16*>(s);
      const __bf16* Al = Ah + BT * KC;
      const __bf16* Bh = Ah + 2 * BT * KC;
      const __bf16* Bl = Ah + 3 * BT * KC;
      const int sl = (lg ^ ((li >> 1) & 3)) << 3;  // swizzled slot of this lane
      bf16x8 bh[FN], bl[FN];
#pragma unroll
      for (int c = 0; c < FN; ++c) {
        const int o = (wn + c * 16 + li) * KC + sl;
        bh[c] = *reinterpret_cast<const bf16x8*>(Bh + o);
        bl[c] = *reinterpret_cast<const bf16x8*>(Bl + o);
      }
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        // The next chunk's split and LDS stores go in halfway through this
        // chunk's products, overlapping the matrix cores: its buffer is not
        // read in this chunk (every wave passed the last barrier after
        // reading it).  Staged after the products instead: 3 % slower
        // alone, 1-1.5 % on 1 x 16384 / 1 x 32768 calls.
        if (a == FM / 2 && more) stage(cur ^ 1);
        const int o = (wm + a * 16 + li) * KC + sl;
        const bf16x8 ah = *reinterpret_cast<const bf16x8*>(Ah + o);
        const bf16x8 al = *reinterpret_cast<const bf16x8*>(Al + o);
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[c], acc[a][c], 0, 0, 0);
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[c], acc[a][c], 0, 0, 0);
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[c], acc[a][c], 0, 0, 0);
        }
      }
    } else {
      const float* As = s;
      const float* Bs = s + BT * KC;
#pragma unroll
      for (int ks = 0; ks < KC; ks += 16) {
        const int sl = (((ks >> 2) + lg) ^ (li & 7)) << 2;  // swizzled slot of this lane
        float4 fb[FN];
#pragma unroll
        for (int c = 0; c < FN; ++c) fb[c] = *reinterpret_cast<const float4*>(Bs + (wn + c * 16 + li) * KC + sl);
#pragma unroll
        for (int a = 0; a < FM; ++a) {
          const float4 fa = *reinterpret_cast<const float4*>(As + (wm + a * 16 + li) * KC + sl);
#pragma unroll
          for (int c = 0; c < FN; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.x, fb[c].x, acc[a][c], 0, 0, 0);
#pragma unroll
          for (int c = 0; c < FN; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.y, fb[c].y, acc[a][c], 0, 0, 0);
#pragma unroll
          for (int c = 0; c < FN; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.z, fb[c].z, acc[a][c], 0, 0, 0);
#pragma unroll
          for (int c = 0; c < FN; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.w, fb[c].w, acc[a][c], 0, 0, 0);
        }
      }
    }
    if (!SPLIT && more) stage(cur ^ 1);  // FP32: after the products (halfway: 1-2 % slower)
    __syncthreads();
  }

  // C -= D on the lower part (j <= i), i < n, j < c1.  Lane holds
  // D[4*lg + v][li] of each 16x16 block.  Two rounds of 64 elements: every
  // load of a round is issued before its stores (a masked element reads
  // C at the tile's own corner, always valid, and is not stored); the
  // addresses are recomputed for the stores rather than held.
  const int jl = j0 + wn + li;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    float old[FM / 2][FN][4];
#pragma unroll
    for (int a = 0; a < FM / 2; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wm + (h * (FM / 2) + a) * 16 + 4 * lg + v;
        const float* row = Cs + (int64_t)(i < M.n ? i : i0) * ldcs;
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          const int j = jl + c * 16;
          const bool m = i < M.n && j < c1 && j <= i;
          old[a][c][v] = gld(m ? row + j : Cs + (int64_t)i0 * ldcs + j0);
        }
      }
#pragma unroll
    for (int a = 0; a < FM / 2; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wm + (h * (FM / 2) + a) * 16 + 4 * lg + v;
        float* row = L + (int64_t)i * ld;
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          const int j = jl + c * 16;
          if (i < M.n && j < c1 && j <= i) gst(row + j, old[a][c][v] - acc[h * (FM / 2) + a][c][v]);
        }
      }
  }
}

// E4M3 form (CHOL_MODE_FP8): the same tile, wavefront split and XCD order,
// with the operands quantised to OCP FP8 E4M3 as they are staged and one
// v_mfma_f32_16x16x32_fp8_fp8 per 16x16x32 block (three BF16 products in
// the BF16x3 form).  Scaling is per row, by a power of two taken from A's
// diagonal: the squares of row r of L sum to A[r][r], so |P[r][k]| <=
// sqrt(A[r][r]) <= 2^E for every k of every panel, and P[r][k] * 2^(8-E)
// lies within +-256 (E4M3's largest finite value is 448), with no pass
// over P to find its maximum.  E4M3 keeps 3 mantissa bits over its whole
// normal range (2^-6 .. 448), so the row's smaller entries lose no relative
// precision to the bound.  C -= s_i s_j (Q Q^T)_ij, exact in the scale
// (powers of two), accumulated in FP32.  Reference counterpart: the MXFP8
// trailing products of the large-n drivers (submission.py:1626-1675,
// :3952-3991), which scale per 32-element block.
constexpr int F8_STAGE = 2 * BT * KC;  // bytes per stage: A rows then B rows (16 KiB)

// 2^E with 2^E >= sqrt(d): d = f 2^e2, f in [0.5, 1), so sqrt(d) < 2^(e2/2)
// and E = ceil(e2 / 2).  Rows past the matrix (loaded as zeros) take 1.
__device__ __forceinline__ int row_exp(const Mat& M, int r, int rlim) {
  if (r >= rlim) return 0;
  int e2;
  (void)frexpf(gld(M.A + (int64_t)r * M.lda + r), &e2);
  return (e2 + 1) >> 1;
}

// Byte offset of 4-element group q (k = 4q..4q+3) of row r in a stage:
// 32 bytes per row, 8-byte slot g = q >> 1 at (g ^ ((r >> 3) & 3)), so the
// 16 rows a fragment read takes hit distinct banks.
__device__ __forceinline__ int f8_off(int r, int q) {
  return r * KC + ((((q >> 1) ^ (r >> 3)) & 3) << 3) + ((q & 1) << 2);
}

// inv: the rows' inverse scales 2^(8-E) in LDS (row r of the operand at inv[r]).
__device__ __forceinline__ void put_e4m3(unsigned char* __restrict__ s, const float4 (&reg)[PER],
                                         const float* inv, int tid) {
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int c = tid + u * NT;
    const float4 v = reg[u];
    const float f = inv[c >> 3];
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(v.x * f, v.y * f, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(v.z * f, v.w * f, w, true);
    *reinterpret_cast<int*>(s + f8_off(c >> 3, c & 7)) = w;
  }
}

template <bool VEC>
__global__ __launch_bounds__(NT) void k_syrk256_e4m3(Batch p, int col0, int col1, int k0, int kl) {
  __shared__ __align__(16) unsigned char f8[2 * F8_STAGE];  // 32 KiB
  __shared__ float sc[2][BT];                               // row scales 2^(E-8): A rows, B rows
  __shared__ float si8[2][BT];                              // their inverses 2^(8-E)

  const int64_t b = blockIdx.y;
  const Mat M = mat_at(p, b);
  const int c1 = (col1 < 0 || col1 > M.n) ? M.n : col1;
  if (c1 <= col0) return;
  const int TI = (M.n - col0 + BT - 1) / BT;
  const int TJ = (c1 - col0 + BT - 1) / BT;
  const int T = (int)((int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2);
  const int bid = blockIdx.x;
  if (bid >= T) return;
  int ti, tj;
  decode_tile_grouped(xcd_tile(bid, T), TI, TJ, 4, ti, tj);
  const int i0 = col0 + BT * ti;
  const int j0 = col0 + BT * tj;
  float* L = M.L;
  const float* Cs = p.first ? M.A : M.L;
  const int64_t ld = M.ld;
  const int64_t ldcs = p.first ? M.lda : M.ld;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 2) * WM;
  const int wn = (wave & 3) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  {
    // thread t < 256: row i0 + t of the A side; else row j0 + t - 256 of B
    const int side = tid >= BT, r = tid - side * BT;
    const int e = side ? row_exp(M, j0 + r, c1) : row_exp(M, i0 + r, M.n);
    sc[side][r] = ldexpf(1.f, e - 8);
    si8[side][r] = ldexpf(1.f, 8 - e);
  }
  __syncthreads();

  floatx4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int c = 0; c < FN; ++c) acc[a][c] = floatx4{0.f, 0.f, 0.f, 0.f};

  float4 ra[PER], rb[PER];
  auto stage = [&](int buf) {
    unsigned char* s = f8 + buf * F8_STAGE;
    put_e4m3(s, ra, si8[0], tid);
    put_e4m3(s + BT * KC, rb, si8[1], tid);
  };

  load_chunk<VEC>(ra, L, ld, i0, M.n, k0, tid);
  load_chunk<VEC>(rb, L, ld, j0, c1, k0, tid);
  stage(0);
  __syncthreads();

  const int nch = kl / KC;
  for (int ch = 0; ch < nch; ++ch) {
    const int cur = ch & 1;
    const bool more = ch + 1 < nch;
    if (more) {
      load_chunk<VEC>(ra, L, ld, i0, M.n, k0 + (ch + 1) * KC, tid);
      load_chunk<VEC>(rb, L, ld, j0, c1, k0 + (ch + 1) * KC, tid);
    }
    const unsigned char* s = f8 + cur * F8_STAGE;
    const unsigned char* Bq = s + BT * KC;
    long bq[FN];
#pragma unroll
    for (int c = 0; c < FN; ++c) {
      const int r = wn + c * 16 + li;
      bq[c] = *reinterpret_cast<const long*>(Bq + f8_off(r, 2 * lg));
    }
#pragma unroll
    for (int a = 0; a < FM; ++a) {
      // the next chunk's quantisation and LDS stores halfway through this
      // chunk's products (its buffer is not read in this chunk)
      if (a == FM / 2 && more) stage(cur ^ 1);
      const int r = wm + a * 16 + li;
      const long aq = *reinterpret_cast<const long*>(s + f8_off(r, 2 * lg));
#pragma unroll
      for (int c = 0; c < FN; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(aq, bq[c], acc[a][c], 0, 0, 0);
    }
    __syncthreads();
  }

  // C -= s_i s_j D on the lower part, as in k_syrk256.
  const int jl = j0 + wn + li;
  float sj[FN];
#pragma unroll
  for (int c = 0; c < FN; ++c) sj[c] = sc[1][wn + c * 16 + li];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    float old[FM / 2][FN][4];
#pragma unroll
    for (int a = 0; a < FM / 2; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wm + (h * (FM / 2) + a) * 16 + 4 * lg + v;
        const float* row = Cs + (int64_t)(i < M.n ? i : i0) * ldcs;
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          const int j = jl + c * 16;
          const bool m = i < M.n && j < c1 && j <= i;
          old[a][c][v] = gld(m ? row + j : Cs + (int64_t)i0 * ldcs + j0);
        }
      }
#pragma unroll
    for (int a = 0; a < FM / 2; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int il = wm + (h * (FM / 2) + a) * 16 + 4 * lg + v;
        const int i = i0 + il;
        const float si = sc[0][il];
        float* row = L + (int64_t)i * ld;
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          const int j = jl + c * 16;
          if (i < M.n && j < c1 && j <= i) gst(row + j, old[a][c][v] - acc[h * (FM / 2) + a][c][v] * (si * sj[c]));
        }
      }
  }
}

}  // namespace

int64_t syrk256_tiles(const Batch& p, int col0, int col1) {
  const int TI = (p.n - col0 + BT - 1) / BT;
  const int TJ = ((col1 < 0 ? p.n : col1) - col0 + BT - 1) / BT;
  return ((int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2) * p.batch;
}

int launch_syrk256(const Batch& p, int col0, int col1, int k0, int kl, bool split, bool vec, hipStream_t s) {
  const int TI = (p.n - col0 + BT - 1) / BT;
  const int TJ = ((col1 < 0 ? p.n : col1) - col0 + BT - 1) / BT;
  const int64_t tiles = (int64_t)TJ * TI - (int64_t)TJ * (TJ - 1) / 2;
  dim3 grid((unsigned)tiles, (unsigned)p.batch);
  if (split && p.f8) {
    if (vec)
      hipLaunchKernelGGL((k_syrk256_e4m3<true>), grid, dim3(NT), 0, s, p, col0, col1, k0, kl);
    else
      hipLaunchKernelGGL((k_syrk256_e4m3<false>), grid, dim3(NT), 0, s, p, col0, col1, k0, kl);
    return launch_status("k_syrk256_e4m3");
  }
  if (split) {
    if (vec)
      hipLaunchKernelGGL((k_syrk256<true, true>), grid, dim3(NT), 0, s, p, col0, col1, k0, kl);
    else
      hipLaunchKernelGGL((k_syrk256<false, true>), grid, dim3(NT), 0, s, p, col0, col1, k0, kl);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_syrk256<true, false>), grid, dim3(NT), 0, s, p, col0, col1, k0, kl);
    else
      hipLaunchKernelGGL((k_syrk256<false, false>), grid, dim3(NT), 0, s, p, col0, col1, k0, kl);
  }
  return launch_status("k_syrk256");
}

}  // namespace chol
