// chol_blocked.hip — copy and info kernels of the blocked
// Cholesky for n > 64 (the driver is blocked_potrf in chol_abi.cpp).
//
// The dense factor is built in place in the output L.  The matrix is walked
// in 128-wide block columns:
//   diag  : factor the 128x128 diagonal block in LDS and form its inverse
//           W = L11^-1 (kept in the workspace; chol_diag.hip);
//   panel : L21 = A21 W^T for every row below the block (a GEMM, K = 128);
//   syrk  : C -= P P^T on the lower trailing tiles.
// The two products run on the matrix cores (chol_gemm.hip, chol_syrk256.hip).
// Inner blocks are grouped into an outer block of `nbo` columns: inner
// steps update only the outer block's columns, and one SYRK with K = nbo then
// updates the rest of the trailing matrix, so the dominant update streams
// its C tile once per nbo columns instead of once per 128.
// The kernels that first touch a region of L read its old values from A, so
// a dense call moves no copy of tril(A) into L; the first diagonal-block
// launch zeros L's upper triangle with extra workgroups on the CUs that the
// batch leaves idle, or k_zero_upper does when it leaves none.  A ragged
// call copies only each matrix's first 128 columns (k_copy_head), reads the
// rest of A on first touch at A's own row stride, and zeros the rest of
// the upper triangle by rows (k_zero_upper_rows).
//
// Reference counterparts: the 128-wide diag+inverse kernel e62_diag128
// (submission.py:4694-4997), the panel-times-inverse bmm and trailing
// baddbmm_ of _exp062_factor (submission.py:5152-5196), the two-level
// micro/panel blocking of _split32_launch (submission.py:2375-2567) and the
// trsm-free inverse panels of the large-n drivers (submission.py:3606-3676).
#include "chol_common.h"

namespace chol {
namespace {

// ---------------------------------------------------------------- copy ----
// Ragged first touch, head part: L[i][j] = (j <= i ? A[i][j] : 0) for the
// first min(n, 128) columns of every row (the first column block, which the
// first diagonal block and panel read, so that they read L's aligned rows;
// A's rows, at their own stride n, are 16-byte aligned only for n % 4 ==
// 0).  Half a workgroup per row, workgroup x taking rows 2x, 2x + 1,
// 2x + 2 gridDim.x, ...
__global__ __launch_bounds__(256) void k_copy_head(Batch p) {
  const Mat M = mat_at(p, blockIdx.y);
  const int n = M.n;
  const int w = n < kNB ? n : kNB;
  const int t = threadIdx.x & (kNB - 1);
  for (int i = 2 * blockIdx.x + (threadIdx.x >> 7); i < n; i += 2 * gridDim.x)
    if (t < w) gst(M.L + (int64_t)i * M.ld + t, t <= i ? gld_nt(M.A + (int64_t)i * M.lda + t) : 0.f);
}

// Ragged first touch, the rest of the upper triangle: L[i][j] = 0 for
// max(128, the chunk after row i's diagonal chunk) <= j < n, by rows
// (workgroup x: rows x, x + gridDim.x, ...), float4 stores where L's row
// starts 16-byte aligned.  The head columns are k_copy_head's, the
// diagonal chunks the diagonal-block kernel's.
__global__ __launch_bounds__(256) void k_zero_upper_rows(Batch p) {
  const Mat M = mat_at(p, blockIdx.y);
  const int n = M.n;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    float* row = M.L + (int64_t)i * M.ld;
    const int j0 = max(kNB, ((i >> 2) + 1) * 4);
    if (j0 >= n) continue;
    if ((reinterpret_cast<uintptr_t>(row) & 15) == 0) {
      const int full = (n - j0) >> 2;
      for (int c = threadIdx.x; c < full; c += blockDim.x) gst4_nt(row + j0 + 4 * c, make_float4(0.f, 0.f, 0.f, 0.f));
      const int jt = j0 + 4 * full + threadIdx.x;
      if (jt < n) gst_nt(row + jt, 0.f);
    } else {
      for (int j = j0 + threadIdx.x; j < n; j += blockDim.x) gst_nt(row + j, 0.f);
    }
  }
}

// Zeros right of each row's diagonal chunk: L[i][c..c+3] = 0 for every
// 4-column chunk c > i (the diagonal-block kernel stores each row's
// diagonal chunk whole).  Used where the batch fills the GPU, so that the
// first diagonal-block launch has no idle CUs to do it.  Streaming stores:
// 2 n^2 bytes where a copy of tril(A) moves 8 n^2.
template <bool VEC>
__global__ __launch_bounds__(256) void k_zero_upper(Batch p) {
  const int n = p.n;
  const int n4 = (n + 3) / 4;
  const int64_t b = blockIdx.y;
  float* dst = p.L + b * p.bs;
  const int64_t total = (int64_t)n * n4;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n4);
    const int c = 4 * (int)(e - (int64_t)i * n4);
    if (c <= i) continue;
    float* q = as_global(dst + (int64_t)i * p.ld + c);
    if constexpr (VEC) {
      gst4_nt(q, make_float4(0.f, 0.f, 0.f, 0.f));
    } else {
      for (int j = 0; j < 4 && c + j < n; ++j) gst_nt(q + j, 0.f);
    }
  }
}

__global__ void k_zero_info(int32_t* info, int64_t batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch) info[i] = 0;
}

}  // namespace

int launch_zero_info(int32_t* info, int64_t batch, hipStream_t s) {
  if (!info || batch <= 0) return 0;
  hipLaunchKernelGGL(k_zero_info, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, info, batch);
  return launch_status("k_zero_info");
}

int launch_copy_head(const Batch& p, hipStream_t s) {
  if (p.batch <= 0) return 0;
  const unsigned gx = (unsigned)((p.n + 15) / 16);  // 16 rows of the largest matrix per workgroup
  hipLaunchKernelGGL(k_copy_head, dim3(gx, (unsigned)p.batch), dim3(256), 0, s, p);
  return launch_status("k_copy_head");
}

int launch_zero_upper_rows(const Batch& p, hipStream_t s) {
  if (p.batch <= 0 || p.n <= kNB) return 0;
  const unsigned gx = (unsigned)((p.n + 7) / 8);  // 8 rows of the largest matrix per workgroup
  hipLaunchKernelGGL(k_zero_upper_rows, dim3(gx, (unsigned)p.batch), dim3(256), 0, s, p);
  return launch_status("k_zero_upper_rows");
}

int launch_zero_upper(const Batch& p, hipStream_t s) {
  const bool vec = !p.Ap && (reinterpret_cast<uintptr_t>(p.L) & 15) == 0 && (p.ld % 4) == 0 &&
                   (p.bs % 4) == 0 && (p.n % 4) == 0;
  const int64_t elems = (int64_t)p.n * ((p.n + 3) / 4);
  int64_t gx = (elems + 255) / 256;
  if (gx > 4096) gx = 4096;
  if (vec)
    hipLaunchKernelGGL((k_zero_upper<true>), dim3((unsigned)gx, (unsigned)p.batch), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((k_zero_upper<false>), dim3((unsigned)gx, (unsigned)p.batch), dim3(256), 0, s, p);
  return launch_status("k_zero_upper");
}

}  // namespace chol
