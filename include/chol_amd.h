/*
 * chol_amd.h — C ABI of the MI355X (gfx950) batched Cholesky engine.
 *
 * Every entry point takes plain device pointers, sizes and strides, and the
 * HIP stream to launch on (passed as void*; a hipStream_t, 0 = null stream).
 * No torch or HIP types appear in the signatures, so ctypes, cgo or any other
 * FFI can bind them.  Nothing here synchronises the stream, allocates device
 * memory on the hot path, or frees caller memory.
 *
 * Layout: a batch of `batch` matrices, matrix b starting at
 * A + b*batch_stride, row-major with leading dimension `ld` (elements), so
 * A[b][i][j] = A[b*batch_stride + i*ld + j].  Only the lower triangle of A is
 * read.  L receives the lower-triangular factor with exact zeros above the
 * diagonal.  A is never written; A and L must not overlap.
 *
 * Return value: 0 on success, a negative CHOL_E* code on a rejected argument
 * (nothing is launched then), or a positive hipError_t from a launch.
 * chol_last_error() gives a message for the last non-zero return on the
 * calling thread.
 *
 * Non-SPD input is not an error: like LAPACK potrf with info, the factor then
 * holds NaN from the first non-positive pivot on, and info[b] (when info is
 * non-null) is set to the 1-based order of the first leading minor that is
 * not positive definite, 0 when the matrix factored.
 */
#ifndef CHOL_AMD_H
#define CHOL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHOL_ABI_VERSION 2

/* Error codes (negative; positive values are hipError_t). */
#define CHOL_E_ARG (-1)       /* bad size, stride or null pointer            */
#define CHOL_E_WORKSPACE (-2) /* workspace missing or smaller than required  */
#define CHOL_E_MODE (-3)      /* unknown precision mode                      */

/* Precision modes. */
#define CHOL_MODE_PLAIN 0 /* all-FP32 arithmetic                            */
#define CHOL_MODE_MIXED 1 /* BF16x3 split-product panels and trailing updates, FP32 acc */
/* CHOL_MODE_MIXED with the 256-tile bulk trailing updates of orders >= 16384
 * in OCP FP8 E4M3 (rows scaled by a power of two from A's diagonal), FP32
 * accumulation: the reference's MXFP8 large-n products (submission.py:
 * 1626-1675, :3952-4094); residual ~1e-3..1e-2, inside the reference's
 * gate, not the 1e-5 of the mixed form.  Below 16384 it is CHOL_MODE_MIXED. */
#define CHOL_MODE_FP8 2

/* Routes of chol_potrf_f32 (chol_route). */
#define CHOL_ROUTE_NONE 0    /* batch == 0 or n == 0: nothing launched        */
#define CHOL_ROUTE_SMALL 1   /* n <= 64: whole matrices in wavefront registers */
#define CHOL_ROUTE_LDS 2     /* 64 < n <= 256: whole matrix in one LDS         */
#define CHOL_ROUTE_BLOCKED 3 /* blocked path, outer blocks of nbo columns      */
#define CHOL_ROUTE_FUSED 4   /* n > 64: one workgroup factors a whole matrix  */

/* Version of this ABI (CHOL_ABI_VERSION). */
int chol_abi_version(void);

/* Message for the last non-zero return of any chol_* call on this thread. */
const char* chol_last_error(void);

/*
 * Device workspace (bytes) that chol_potrf_f32 needs for this problem.  The
 * caller allocates it (any 256-byte aligned device buffer) and passes it in.
 * Replaces: the module-global cached workspaces of the reference dispatcher
 * (submission.py:5134-5149 `_exp062_buffers`, :2569-2570 split32 buffers).
 */
size_t chol_workspace_bytes(int64_t batch, int64_t n, int32_t mode);

/*
 * Batched Cholesky factorisation A = L L^T, FP32 in and out.
 * Replaces: custom_kernel(data) -> L (submission.py:4101-4268) together with
 * the pybind launchers it reaches: chol32_launch (submission.py:168),
 * chol64_launch (:275), chol128_launch (:566), micro32_launch (:2412),
 * chol256_launch (:1091), e62_diag128_launch / e62_zero_upper_launch
 * (:5032-5101), and the torch.linalg.cholesky_ex fallback (:4268).
 *
 *   A, L          device pointers (see Layout above)
 *   batch, n      batch >= 0, n >= 0 (batch == 0 or n == 0 launches nothing)
 *   batch_stride  >= n*ld when batch > 1
 *   ld            >= max(n, 1)
 *   mode          CHOL_MODE_PLAIN, CHOL_MODE_MIXED or CHOL_MODE_FP8
 *   info          nullable device int32[batch]
 *   ws, ws_bytes  device workspace of at least chol_workspace_bytes()
 *   stream        hipStream_t to launch on
 */
int chol_potrf_f32(const float* A, float* L, int64_t batch, int64_t n,
                   int64_t batch_stride, int64_t ld, int32_t mode,
                   int32_t* info, void* ws, size_t ws_bytes, void* stream);

/*
 * chol_potrf_f32 without any allocation: no caller workspace, no hipMalloc,
 * no stream or event creation; every launch on `stream`.  The blocked
 * route's workspace (64 KiB per matrix of order > 64) comes from a device
 * array the library holds, CHOL_NOALLOC_SCRATCH_BYTES per device until the
 * process ends, so a call takes at most 1024 such matrices (more returns
 * CHOL_E_WORKSPACE, nothing launched); orders <= 64 have no limit.  Calls
 * that overlap on one device share that array and race: serialise them.
 * Same arguments and results as chol_potrf_f32 (bit for bit), without the
 * side-stream overlap.  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).
 */
#define CHOL_NOALLOC_SCRATCH_BYTES (64u << 20)
int chol_potrf_f32_noalloc(const float* A, float* L, int64_t batch, int64_t n,
                           int64_t batch_stride, int64_t ld, int32_t mode,
                           int32_t* info, void* stream);

/*
 * One outer block's bulk trailing update on its own, by the 256-tile kernel
 * chol_potrf_f32 runs for it: for one matrix (A and L at row stride ld),
 * with P = columns [k0, k0 + kl) of L,
 *     L[i][j] -= sum_k P[i][k] P[j][k]    for col0 <= j <= i < n,
 * L read and written at those elements only.  mode picks the products:
 * CHOL_MODE_PLAIN FP32, CHOL_MODE_MIXED BF16x3, CHOL_MODE_FP8 E4M3 with
 * row r scaled by 2^(8-E), 2^E >= sqrt(A[r][r]) (A supplies only its
 * diagonal; for any n).  kl a positive multiple of 32, k0 + kl <= col0.
 * Replaces the trailing SYRK-as-GEMM of the reference's blocked drivers
 * (submission.py:5163-5184, MXFP8 form :3952-3991), exposed so that the
 * product can be checked against the oracle on its own.
 */
int chol_trailing_update_f32(const float* A, float* L, int64_t n, int64_t ld,
                             int64_t col0, int64_t k0, int64_t kl, int32_t mode,
                             void* stream);

/*
 * The route chol_potrf_f32 takes for (batch, n, mode) under the current
 * tuning table (CHOL_ROUTE_*), and through nbo_out (nullable) the outer
 * block width of the blocked route (0 otherwise).  Host only, no HIP call.
 * Replaces: the (batch, n) tables of the reference's dispatcher, read
 * inline in custom_kernel (submission.py:4101-4268).
 */
int chol_route(int64_t batch, int64_t n, int32_t mode, int64_t* nbo_out);

/*
 * Dry run of the blocked route: the launches and stream-ordering steps that
 * chol_potrf_f32 would enqueue for one batch of at most 65535 matrices of
 * order n with leading dimension ld, on a device with `cus` compute units
 * (<= 0: 256), without any HIP call.  *count receives the number of
 * records; the first min(*count, cap) are written to rec, 10 int64 each:
 *   {kind, stream (0 = caller's, 1 = side), fields...}
 *   kind 1 diagonal block : k0, first-touch, carried update c0, c1, k0,
 *                           flags (first-touch | in-kernel << 1), zero-fill
 *                           rows r0, r1 (carried / zero fields -1 if none)
 *   kind 2 panel          : k0, first-touch
 *   kind 3 trailing update: col0, col1, k0, kl, first-touch
 *   kind 4 zero fill of L's upper triangle; kind 5 side stream waits for
 *   the caller; kind 6 mark on the side stream; kind 7 caller waits for
 *   the last mark.
 * Other routes give *count = 0.  Replaces nothing in the reference; it
 * exposes the schedule so it can be checked on a host without a GPU.
 */
int chol_plan_potrf(int64_t batch, int64_t n, int64_t ld, int32_t mode,
                    int32_t cus, int64_t* rec, int64_t cap, int64_t* count);

/*
 * Successful calls per route since the library was loaded, in the order
 * {small, LDS, blocked, ragged, backward, triangular inverse, fused}: the
 * first min(cap, 7) are written to out; returns 7.  Replaces the per-route
 * module counters of the reference (_CUDA32_HITS etc., submission.py:92,
 * :179-182, :194, :302, :620), which its runner collects
 * (scripts/_gpu_runner.py:2402-2414); there are no fallbacks here.
 */
int64_t chol_route_counters(int64_t* out, int64_t cap);

/*
 * Route / blocking table written by the autotuner: `rows` rows of 6 int64
 *   {n_lo, n_hi, batch_lo, batch_hi, mode, nbo}   (inclusive ranges)
 * mode = CHOL_MODE_PLAIN or CHOL_MODE_MIXED, or -1 for both forms;
 * nbo = outer block width of the blocked path (a multiple of 128, <= 4096),
 * -1 to keep the whole-matrix LDS kernel (only for n_hi <= 256), or -2 for
 * the fused one-workgroup-per-matrix kernel (only for n_lo > 64).  The
 * first matching row wins; unmatched shapes use built-in defaults; rows = 0
 * clears the table.  Replaces the reference's hand-written per-shape tables
 * (submission.py:2263-2281 _SPLIT32_SHAPES, :3262-3281 _LARGE_CFG,
 * :5364-5371 _EXP062_SHAPES).  Not thread-safe against concurrent
 * factorisations that read it (set it once, at start-up).
 */
int chol_set_tuning(const int64_t* table, int64_t rows);

/*
 * Ragged batch: `count` independent matrices of mixed orders n_host[i].
 * A_host[i] / L_host[i] are DEVICE pointers to matrix i; A is contiguous
 * (ld = n), L has leading dimension ld_host[i] (>= n; ld_host null: n).
 * The three or four arrays themselves are in HOST memory.  info: nullable
 * device int32[count], in the caller's order.  The library sorts by order
 * and runs each size class as one launch sequence: a persistent wavefront
 * kernel (n <= 64) fed by atomic work queues over the size-sorted list, and
 * for n > 64 a copy of tril(A) into L followed by the blocked path in lock
 * step over the matrices, with outer blocks of the tuning table's width
 * (rows of L padded to a multiple of 4, 16-byte aligned, give every kernel
 * after the copy float4 access).
 * Has no counterpart in the reference, which factors one fixed (batch, n)
 * per call; it generalises the per-shape dispatch of submission.py:4101-4268.
 */
size_t chol_ragged_workspace_bytes(const int64_t* n_host, int64_t count);
int chol_potrf_ragged_f32(const float* const* A_host, float* const* L_host,
                          const int64_t* n_host, const int64_t* ld_host,
                          int64_t count, int32_t mode, int32_t* info,
                          void* ws, size_t ws_bytes, void* stream);

/*
 * Adjoint of the factorisation: given L (from chol_potrf_f32) and the
 * upstream gradient Lbar (lower triangle read), writes the symmetric
 *   Abar = 1/2 * sym(L^-T Phi(L^T Lbar) L^-1),
 * Phi(X) = lower triangle of X with its diagonal halved, sym(X) = X + X^T.
 * This is the gradient torch's autograd gives for torch.linalg.cholesky
 * (full symmetric matrix, not the lower-triangle-only gradient).
 * Linv: nullable device buffer batch x n x n (ld = n).  When given with
 * have_linv = 1, it must hold L^-1 (e.g. from chol_tri_inverse_f32) and it is
 * used as is; when null, L^-1 is recomputed inside ws.
 * No counterpart in the reference (it has no backward); replaces the torch
 * autograd formula of torch.linalg.cholesky for callers that differentiate.
 */
size_t chol_backward_workspace_bytes(int64_t batch, int64_t n);
int chol_backward_f32(const float* L, const float* Lbar, float* Abar,
                      int64_t batch, int64_t n, const float* Linv,
                      int32_t have_linv, void* ws, size_t ws_bytes,
                      void* stream);

/* Lower-triangular inverse Linv = L^-1 (batch x n x n, ld = n, contiguous). */
size_t chol_tri_inverse_workspace_bytes(int64_t batch, int64_t n);
int chol_tri_inverse_f32(const float* L, float* Linv, int64_t batch,
                         int64_t n, void* ws, size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CHOL_AMD_H */
