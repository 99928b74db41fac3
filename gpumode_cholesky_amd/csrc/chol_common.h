// chol_common.h — shared helpers of the gfx950 Cholesky kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <mutex>
#include <string>

#include "chol_amd.h"

namespace chol {

// Per-thread message for chol_last_error().
void set_error(const char* fmt, ...);
void clear_error();

// Successful calls per route, read by chol_route_counters (in its order).
enum RouteCounter { RC_SMALL, RC_LDS, RC_BLOCKED, RC_RAGGED, RC_BACKWARD, RC_TRI_INVERSE, RC_FUSED, RC_COUNT };
void note_call(RouteCounter r);

// Wave size of CDNA: every kernel here is written for 64-lane wavefronts.
constexpr int kWave = 64;

// Arguments of one batched factorisation (row-major, see chol_amd.h).
// Dense form: matrix b at A + b*bs, order n, leading dimension ld.
// Ragged form (Ap non-null): matrix b at Ap[b] / Lp[b], order ns[b], A
// contiguous (ld = ns[b]), L with leading dimension lds[b] (ns[b] when lds
// is null); its info slot is orig[b]; n is the largest order.
struct Batch {
  const float* A;
  float* L;
  int64_t batch;
  int n;
  int64_t bs;  // batch stride (elements)
  int64_t ld;  // leading dimension (elements)
  int32_t* info;
  const float* const* Ap = nullptr;
  float* const* Lp = nullptr;
  const int32_t* ns = nullptr;
  const int32_t* orig = nullptr;
  const int32_t* lds = nullptr;
  // First touch of a region: the kernel reads the old values of its output
  // region (C of a trailing update, A21 of a panel, the diagonal block)
  // from A (at A's row stride) instead of L, so no copy of tril(A) into L
  // precedes the factorisation.
  bool first = false;
  // Fused kernel of a ragged call: the first 128 columns of L already hold
  // tril(A) (k_copy_head); step 0's trailing update reads the old values of
  // the columns beyond them from A, and the kernel zeros the upper triangle.
  bool trailing_from_a = false;
  // CHOL_MODE_FP8 at n >= kF8MinN: the 256-tile bulk updates take E4M3
  // operands (k_syrk256_e4m3) where the mixed form takes BF16x3.
  bool f8 = false;
};

// One matrix of a Batch, dense or ragged.
struct Mat {
  const float* A;
  float* L;
  int n;
  int64_t ld;    // leading dimension of L (and of A in the dense form)
  int64_t slot;  // index into info
  int64_t lda;   // leading dimension of A
};

// Marks a pointer as global memory.  A pointer loaded from the ragged
// descriptor table is otherwise generic, and every access through it
// becomes a flat_load/flat_store, which the hardware counts against both
// the vector-memory and the LDS wait counters: an LDS wait then also waits
// for in-flight global prefetches.  The round trip through address space 1
// lets the compiler emit global_load/global_store.
template <typename T>
__device__ __forceinline__ T* as_global(T* q) {
  return (T*)((__attribute__((address_space(1))) T*)q);
}

// Global loads/stores through address-space-1 pointers: always
// global_load/global_store, whatever the compiler could infer about `q`.
typedef __attribute__((address_space(1))) float gfloat;
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) fvec4 gfvec4;
__device__ __forceinline__ float gld(const float* q) { return *(const gfloat*)q; }
__device__ __forceinline__ float4 gld4(const float* q) {
  const fvec4 v = *(const gfvec4*)q;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(float* q, float v) { *(gfloat*)q = v; }
__device__ __forceinline__ void gst4(float* q, float4 v) { *(gfvec4*)q = fvec4{v.x, v.y, v.z, v.w}; }
// Streaming (non-temporal) forms for data touched once by the kernel: the
// factor's final stores and the input's single read.  They keep such data
// from piling up as dirty lines in the XCD's L2, which the end-of-kernel
// release would otherwise write back after the last wave.
__device__ __forceinline__ float4 gld4_nt(const float* q) {
  const fvec4 v = __builtin_nontemporal_load((const gfvec4*)q);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst4_nt(float* q, float4 v) {
  __builtin_nontemporal_store(fvec4{v.x, v.y, v.z, v.w}, (gfvec4*)q);
}
__device__ __forceinline__ void gst_nt(float* q, float v) { __builtin_nontemporal_store(v, (gfloat*)q); }
__device__ __forceinline__ float gld_nt(const float* q) { return __builtin_nontemporal_load((const gfloat*)q); }

__device__ __forceinline__ Mat mat_at(const Batch& p, int64_t b) {
  if (p.Ap) {
    const int n = p.ns[b];
    return Mat{as_global(p.Ap[b]), as_global(p.Lp[b]), n, p.lds ? (int64_t)p.lds[b] : (int64_t)n,
               p.orig ? (int64_t)p.orig[b] : b, (int64_t)n};
  }
  return Mat{as_global(p.A + b * p.bs), as_global(p.L + b * p.bs), p.n, p.ld, b, p.ld};
}

// Compiler-level barrier between LDS traffic of lanes of one wavefront.
// LDS instructions of a wave execute in order, so lanes exchanging values
// through LDS inside one wave need no s_barrier, only this fence against
// compiler reordering.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The BF16x3 operand split of a pair of FP32 values, two to a dword:
// hi = x rounded to BF16 (nearest even; one v_cvt_pk_bf16_f32 for the
// pair), lo = the exact FP32 remainder x - hi rounded to BF16.  hi's FP32
// values come from the packed dword by a shift and a mask, so a pair costs
// six VALU operations (converting each hi on its own and packing it again
// cost eight).  Element 0 in the low half.
__device__ __forceinline__ void split_bf16_pair(float x0, float x1, uint32_t& hi, uint32_t& lo) {
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  const uint32_t u = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2){x0, x1}, bf2));
  const float h0 = __builtin_bit_cast(float, u << 16);
  const float h1 = __builtin_bit_cast(float, u & 0xffff0000u);
  hi = u;
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2){x0 - h0, x1 - h1}, bf2));
}

// Right-looking factorisation of the leading NP columns of a block held one
// row per lane: a[] holds columns 0..NP-1 of the lane's row r (lower part;
// the upper part is never read by other lanes), cb the per-block LDS column
// buffer (one float per row, 16-byte aligned), `writer` whether this lane
// owns row r (duplicate lanes compute but do not write).  Rows r >= NP
// (a wave of 64 rows over NP = 32 columns) come out as the panel below the
// factor, L[r][k] for k < NP.  Pivot step k: every row publishes its
// column-k entry, reads back the pivot c[k] and c[k+1..NP) with broadcast
// reads, and folds a[j] -= (a[k]/c[k]) c[j].
//
// Software pipelined: as soon as step k's multiplier is known, column k+1
// is brought up to date, published and its read-back issued; the bulk of
// step k's updates then runs while that LDS round trip is in flight.  The
// scheduling barriers keep the compiler from hoisting the next step's
// pivot arithmetic (and with it the LDS wait) above those updates.  The
// pivot reciprocal square root is v_rsq_f32 (~1 ulp); L[r][k] =
// a[k] * rsq(c[k]) for r > k, sqrt(c[k]) from c[k] * rsq(c[k]) plus one
// Newton correction for r == k (exact for perfect squares), 0 above the
// diagonal.  A non-positive pivot yields NaN from there on; the return
// value is the 1-based first failing pivot (0 if none).
template <int NP>
__device__ __forceinline__ int reg_factor(float (&a)[NP], int r, float* cb, bool writer) {
  int bad = 0;
  float c[NP];
  if (writer) cb[r] = a[0];
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < NP; q += 4) {
    const float4 v = *reinterpret_cast<const float4*>(cb + q);
    c[q] = v.x;
    c[q + 1] = v.y;
    c[q + 2] = v.z;
    c[q + 3] = v.w;
  }
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const float piv = c[k];
    if (bad == 0 && !(piv > 0.f)) bad = k + 1;
    const float is = __builtin_amdgcn_rsqf(piv);
    const float lk = a[k] * is;
    const float t = (r > k) ? lk * is : 0.f;
    float cn[NP];
    if (k + 1 < NP) {
      a[k + 1] = fmaf(-t, c[k + 1], a[k + 1]);
      wave_lds_sync();  // this step's reads of cb were issued before the write
      if (writer) cb[r] = a[k + 1];
      wave_lds_sync();
#pragma unroll
      for (int q = ((k + 1) / 4) * 4; q < NP; q += 4) {
        const float4 v = *reinterpret_cast<const float4*>(cb + q);
        cn[q] = v.x;
        cn[q + 1] = v.y;
        cn[q + 2] = v.z;
        cn[q + 3] = v.w;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = k + 2; j < NP; ++j) a[j] = fmaf(-t, c[j], a[j]);
    // diagonal: piv * rsq(piv), one Newton step on the residual piv - s^2
    const float s0 = piv * is;
    const float s = fmaf(0.5f * is, fmaf(-s0, s0, piv), s0);
    a[k] = (r > k) ? lk : ((r == k) ? s : 0.f);
    __builtin_amdgcn_sched_barrier(0);
    if (k + 1 < NP) {
#pragma unroll
      for (int q = ((k + 1) / 4) * 4; q < NP; ++q) c[q] = cn[q];
    }
  }
  return bad;
}

// Tiles of the lower region {(i, j): col0 <= j < col1, j <= i < n} in T x T
// tiles (ti >= tj), enumerated column by column: column tj holds TI - tj.
__device__ __forceinline__ void decode_tile(int idx, int TI, int& ti, int& tj) {
  const float twoTI1 = 2.f * TI + 1.f;
  const float disc = fmaxf(twoTI1 * twoTI1 - 8.f * (float)idx, 0.f);
  int t = (int)floorf((twoTI1 - sqrtf(disc)) * 0.5f);
  if (t < 0) t = 0;
  while (t > 0 && (int64_t)t * TI - (int64_t)t * (t - 1) / 2 > idx) --t;
  while ((int64_t)(t + 1) * TI - (int64_t)(t + 1) * t / 2 <= idx) ++t;
  tj = t;
  ti = t + (idx - (int)((int64_t)t * TI - (int64_t)t * (t - 1) / 2));
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remapping block id -> tile index so that every XCD takes a
// contiguous run of the column-by-column tile list keeps tiles that share
// operand blocks (the same column, neighbouring rows) in one L2.  A
// bijection of [0, T) for bid < T, whatever the dispatch order.
__device__ __forceinline__ int xcd_tile(int bid, int T) {
  const int xcd = bid & 7, q8 = T >> 3, r8 = T & 7;
  return xcd * q8 + (xcd < r8 ? xcd : r8) + (bid >> 3);
}

__device__ __forceinline__ float readlane_f(float x, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}

// The same right-looking step as reg_factor for a short panel, with the
// pivot column broadcast through scalar registers instead of LDS: the
// diagonal block's rows r0..r0+NP-1 are lanes r0..r0+NP-1 (lane = row,
// r0 wave-uniform), and c[j] = column k of row r0+j is v_readlane of lane
// r0+j's a[k].  Rows below the block (r >= r0+NP) come out as the panel
// L[r][k]; lanes r < r0 compute values that the caller discards.
//
// Critical path of a pivot step: the update of column k+1, its read-back,
// v_rcp_f32 of the pivot and one multiply, so the multiplier is
// t = a[k] * rcp(c[k]) (not (a[k] rsq) rsq) and is applied unmasked: the
// entries it touches in rows at or above the pivot lie above the diagonal
// and are zeroed when their own pivot comes.  L[r][k] = a[k] * rsq(c[k])
// for r >= k (the diagonal c[k] rsq(c[k]), within ~1 ulp of sqrt) is off
// that path.  A non-positive pivot yields NaN from there on; the return
// value is the 1-based first failing pivot (0 if none).
template <int NP>
__device__ __forceinline__ int reg_factor_rl(float (&a)[NP], int r, int r0) {
  int bad = 0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    // the whole pivot column first: the reads go to distinct scalar
    // registers while the reciprocal is in flight
    float c[NP];
#pragma unroll
    for (int j = k; j < NP; ++j) c[j] = readlane_f(a[k], r0 + j);
    const float piv = c[k];
    if (bad == 0 && !(piv > 0.f)) bad = k + 1;
    const float t = a[k] * __builtin_amdgcn_rcpf(piv);
#pragma unroll
    for (int j = k + 1; j < NP; ++j) a[j] = fmaf(-t, c[j], a[j]);
    const float lk = a[k] * __builtin_amdgcn_rsqf(piv);
    a[k] = (r >= r0 + k) ? lk : 0.f;
  }
  return bad;
}

inline int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

// Launchers implemented in the kernel translation units.
int launch_small(const Batch& p, hipStream_t s);                  // n <= 64
int launch_lds(const Batch& p, hipStream_t s);                    // n <= 256
int launch_copy_head(const Batch& p, hipStream_t s);              // the first 128 columns of L = tril(A)
int launch_zero_upper_rows(const Batch& p, hipStream_t s);        // L = 0 above the diagonal beyond them
int launch_zero_upper(const Batch& p, hipStream_t s);             // L = 0 right of the diagonal chunks
bool vec_ok(const Batch& p);  // float4 global access allowed for every matrix of p
// One-shot (K = kNB) SYRK tiles that a diagonal-block launch runs on the
// CUs its blocks leave idle: C -= P P^T for columns [c0, c1), P = the kNB
// columns from k0; first: old values of C from A; pre (split only, c0 =
// the launch's own block): the launch's block is one of those columns and
// its kernel applies its share of the update itself before factoring it.
struct DiagBulk {
  int c0, c1, k0;
  bool first, split;
  bool pre = false;
};
int64_t diag_bulk_tiles(const Batch& p, const DiagBulk& u);  // workgroups it adds (tiles x batch)
// A share of the zero fill of L's upper triangle carried by a
// diagonal-block launch: rows [r0, r1) of every matrix, by `wgs` extra
// workgroups.
struct DiagZero {
  int r0, r1, wgs;
};
int launch_diag_factor(const Batch& p, int k0, float* winv, hipStream_t s, const DiagBulk* bulk = nullptr,
                       const DiagZero* zero = nullptr);           // kNB x kNB block
// Whole factorisation, one workgroup per matrix (k_potrf_fused): winv holds
// kNB x kNB floats per matrix; p.first: A and L share ld and L is not yet
// written (else L holds tril(A)); split: BF16x3 panel and update products.
int launch_fused(const Batch& p, float* winv, bool split, hipStream_t s);
int launch_panel(const Batch& p, int k0, const float* winv, bool split,
                 hipStream_t s);  // L21 = A21 W^T (split: BF16x3 products)
// C -= P P^T over {col0 <= j < col1, j <= i < n}, K = [k0, k0+kl);
// split = BF16x3 products (mixed mode) instead of FP32.
int launch_syrk(const Batch& p, int col0, int col1, int k0, int kl, bool split,
                hipStream_t s);
int launch_zero_info(int32_t* info, int64_t batch, hipStream_t s);
// 256x256-tile form of the SYRK for large updates (K a multiple of 32);
// syrk256_tiles: its workgroup count (tiles x batch).
int64_t syrk256_tiles(const Batch& p, int col0, int col1);
int launch_syrk256(const Batch& p, int col0, int col1, int k0, int kl, bool split, bool vec, hipStream_t s);

// Contiguous batches (ld = n).  tri_inverse: X = L^-1, ws = batch*n*n floats.
// backward: Abar from L, Lbar; ws = 2 (Linv given) or 3 batch*n*n floats.
int tri_inverse(const float* L, float* X, int64_t batch, int n, float* ws, hipStream_t s);
int backward(const float* L, const float* Lbar, float* Abar, int64_t batch, int n, const float* Linv,
             float* ws, hipStream_t s);

// Diagonal block width of the blocked path.
constexpr int kNB = 128;
// Smallest order whose bulk updates run in E4M3 in CHOL_MODE_FP8.
constexpr int kF8MinN = 16384;

// The library's side stream of the current device for drivers other than
// the dense blocked path: side_begin makes it wait for s's work so far and
// returns it in *s2 (s itself while s is being captured into a graph),
// holding the device's enqueue lock in lk; side_end makes s wait for the
// work enqueued on *s2 since.
int side_begin(hipStream_t s, hipStream_t* s2, std::unique_lock<std::mutex>& lk);
int side_end(hipStream_t s, hipStream_t s2);

// Dense dispatch behind chol_potrf_f32 (argument checks, route, launches);
// side_ok = false keeps all of the blocked route on stream s.  `who` names
// the entry point in error messages.
int potrf_dispatch(const char* who, const float* A, float* L, int64_t batch, int64_t n, int64_t batch_stride,
                   int64_t ld, int32_t mode, int32_t* info, void* ws, size_t ws_bytes, hipStream_t s, bool side_ok);
// Device workspace bytes of the dense dispatch (chol_workspace_bytes).
size_t potrf_workspace(int64_t batch, int64_t n);
// Outer block width of the tuning table for (n, batch, mode): a multiple of
// kNB, -1 (the whole-matrix LDS route) or 0 when no row matches.
int64_t tuned_outer_width(int64_t n, int64_t batch, int32_t mode);

}  // namespace chol
