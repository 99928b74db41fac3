// chol_ragged.cpp — ragged batches: many independent matrices of mixed
// orders factored together.
//
// The host sorts the matrices by order (largest first) and uploads one
// descriptor table (pointers, orders, leading dimensions of L, original
// positions) into the workspace.  Only each matrix's first 128 columns are
// copied into L (k_copy_head); the trailing updates that first touch the
// rest read its old values from A, at A's own row stride, and the upper
// triangle beyond is zeroed by rows.  The sorted list splits into three
// size classes, each one launch sequence over its contiguous slice:
//   n <= 64 : persistent wavefront kernel (k_potrf_small_queue), one
//             launch per padded order 64/32/16, fed by an atomic work queue
//             over the sorted slice;
//   64 < n <= 512 (or as the tuning table routes orders with nbo = -2):
//             one launch of the fused kernel (k_potrf_fused), a workgroup
//             per matrix that runs its whole factorisation, the largest
//             matrices dispatched first;
//   larger  : the zero fill of L's upper triangles, then the blocked path
//             in lock step: outer blocks of nbo columns (the tuning table's width
//             for the largest order, else 256, 512 above 1024); inside one,
//             128-column steps whose diagonal-block launch touches the
//             matrices with n > k0 (a prefix of the sorted list) and whose
//             panel and right-looking K = 128 update touch those with
//             n > k0 + 128, the update clipped to the outer block and to
//             each matrix's order; after it, one K = nbo update of every
//             matrix's columns beyond the block.  Each kernel reads its
//             matrix's order and leading dimension from the table (blocks
//             past a smaller matrix exit).  Rows of L padded to a multiple
//             of 4 (ld_host) give every kernel float4 access to L whatever
//             the orders.
// All but the lock step run on the library's side stream, beside it on the
// caller's stream.
// The reference factors one fixed (batch, n) per call (custom_kernel,
// submission.py:4101-4268); this generalises its per-shape dispatch.
#include <algorithm>
#include <numeric>
#include <vector>

#include "chol_common.h"

namespace chol {
namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Largest order that goes to the fused class when no tuning row matches.
constexpr int64_t kFusedMax = 512;

struct Layout {
  size_t queues, Ap, Lp, ns, orig, lds, winv, total;
  int64_t nbig;
};

Layout layout(const int64_t* n, int64_t count) {
  Layout l{};
  int64_t nbig = 0;
  for (int64_t i = 0; i < count; ++i) nbig += n[i] > 64;
  l.queues = 0;
  l.Ap = align256(l.queues + 8 * sizeof(int));
  l.Lp = align256(l.Ap + count * sizeof(void*));
  l.ns = align256(l.Lp + count * sizeof(void*));
  l.orig = align256(l.ns + count * sizeof(int32_t));
  l.lds = align256(l.orig + count * sizeof(int32_t));
  l.winv = align256(l.lds + count * sizeof(int32_t));
  l.total = align256(l.winv + (size_t)nbig * kNB * kNB * sizeof(float));
  l.nbig = nbig;
  return l;
}

}  // namespace

int launch_small_queue(const Batch& p, int64_t first, int64_t count, int np, int* queue, int grid, hipStream_t s);

}  // namespace chol

using namespace chol;

extern "C" {

size_t chol_ragged_workspace_bytes(const int64_t* n_host, int64_t count) {
  if (!n_host || count <= 0) return 0;
  return layout(n_host, count).total;
}

int chol_potrf_ragged_f32(const float* const* A_host, float* const* L_host, const int64_t* n_host,
                          const int64_t* ld_host, int64_t count, int32_t mode, int32_t* info, void* ws,
                          size_t ws_bytes, void* stream) {
  clear_error();
  if (count < 0 || (count > 0 && (!A_host || !L_host || !n_host))) {
    set_error("chol_potrf_ragged_f32: bad count or null host arrays");
    return CHOL_E_ARG;
  }
  if (mode != CHOL_MODE_PLAIN && mode != CHOL_MODE_MIXED) {
    set_error("chol_potrf_ragged_f32: unknown mode %d", (int)mode);
    return CHOL_E_MODE;
  }
  if (count == 0) return 0;
  if (count > (int64_t)1 << 30) {
    set_error("chol_potrf_ragged_f32: count %lld too large", (long long)count);
    return CHOL_E_ARG;
  }
  for (int64_t i = 0; i < count; ++i) {
    if (n_host[i] < 0 || n_host[i] > 32768 || (n_host[i] > 0 && (!A_host[i] || !L_host[i]))) {
      set_error("chol_potrf_ragged_f32: matrix %lld: n=%lld or null pointer", (long long)i, (long long)n_host[i]);
      return CHOL_E_ARG;
    }
    if (ld_host && (ld_host[i] < n_host[i] || ld_host[i] > (int64_t)1 << 30)) {
      set_error("chol_potrf_ragged_f32: matrix %lld: ld=%lld < n=%lld", (long long)i, (long long)ld_host[i],
                (long long)n_host[i]);
      return CHOL_E_ARG;
    }
  }
  const Layout lay = layout(n_host, count);
  if (!ws || ws_bytes < lay.total) {
    set_error("chol_potrf_ragged_f32: workspace of %zu bytes required, got %zu", lay.total, ws_bytes);
    return CHOL_E_WORKSPACE;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);

  // Sort by order, largest first (ties keep the caller's order).
  std::vector<int64_t> ord(count);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return n_host[a] > n_host[b]; });
  std::vector<char> desc(lay.winv - lay.Ap);
  auto* Ap = reinterpret_cast<const float**>(desc.data());
  auto* Lp = reinterpret_cast<float**>(desc.data() + (lay.Lp - lay.Ap));
  auto* ns = reinterpret_cast<int32_t*>(desc.data() + (lay.ns - lay.Ap));
  auto* orig = reinterpret_cast<int32_t*>(desc.data() + (lay.orig - lay.Ap));
  auto* lds = reinterpret_cast<int32_t*>(desc.data() + (lay.lds - lay.Ap));
  // float4 access to L by every kernel: each L row of a matrix with n > 64
  // starts 16-byte aligned (A, whose rows are n apart, is read element by
  // element: by the head copy and by the first-touch updates' old values)
  bool aligned = true;
  for (int64_t i = 0; i < count; ++i) {
    const int64_t j = ord[i];
    Ap[i] = A_host[j];
    Lp[i] = L_host[j];
    ns[i] = (int32_t)n_host[j];
    orig[i] = (int32_t)j;
    lds[i] = (int32_t)(ld_host ? ld_host[j] : n_host[j]);
    if (ns[i] > 64)
      aligned = aligned && (lds[i] % 4 == 0) && (reinterpret_cast<uintptr_t>(L_host[j]) & 15) == 0;
  }
  char* w = reinterpret_cast<char*>(ws);
  // Pageable source: the copy is staged before hipMemcpyAsync returns, so
  // the host vector may go out of scope afterwards.
  hipError_t e = hipMemcpyAsync(w + lay.Ap, desc.data(), desc.size(), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    set_error("chol_potrf_ragged_f32: descriptor upload failed: %s", hipGetErrorString(e));
    return (int)e;
  }
  int* queues = reinterpret_cast<int*>(w + lay.queues);
  int rc;
  if ((rc = launch_zero_info(queues, 8, s))) return rc;
  if ((rc = launch_zero_info(info, count, s))) return rc;

  Batch base{nullptr, nullptr, 0, 0, 0, aligned ? 4 : 1, info};
  base.Ap = reinterpret_cast<const float* const*>(w + lay.Ap);
  base.Lp = reinterpret_cast<float* const*>(w + lay.Lp);
  base.ns = reinterpret_cast<const int32_t*>(w + lay.ns);
  base.orig = reinterpret_cast<const int32_t*>(w + lay.orig);
  base.lds = reinterpret_cast<const int32_t*>(w + lay.lds);

  // Class boundaries in the sorted list (orders decrease along it).
  auto first_at_most = [&](int64_t lim) {  // first sorted index with n <= lim
    int64_t i = 0;
    while (i < count && ns[i] > lim) ++i;
    return i;
  };
  const int64_t c64 = first_at_most(64), c32 = first_at_most(32), c16 = first_at_most(16), cz = first_at_most(0);
  // The fused class [cF, c64): the smallest orders above 64 that the
  // tuning table routes to the fused kernel for a batch of `count`
  // (nbo = -2), or, with no matching row, those up to kFusedMax.
  auto fused_order = [&](int64_t n) {
    const int64_t t = tuned_outer_width(n, count, mode);
    return t == 0 ? n <= kFusedMax : t == -2;
  };
  int64_t cF = c64;
  while (cF > 0 && fused_order(ns[cF - 1])) --cF;
  auto slice = [&](int64_t c0, int64_t cnt) {
    Batch q = base;
    q.Ap = base.Ap + c0;
    q.Lp = base.Lp + c0;
    q.ns = base.ns + c0;
    q.orig = base.orig + c0;
    q.lds = base.lds + c0;
    q.batch = cnt;
    q.n = ns[c0];
    return q;
  };
  float* const winv0 = reinterpret_cast<float*>(w + lay.winv);
  // No copy of tril(A) into L beyond each matrix's first 128 columns
  // (k_copy_head), which the first diagonal block and panel read from L's
  // aligned rows (A's rows, at stride n, are unaligned for odd n); every
  // later region is first touched by a trailing update, which reads its old
  // values from A at A's own row stride, and the rest of the upper triangle
  // is zeroed by rows.  (The matrix on grid.y: chunks of at most 65535.)
  for (int64_t c0 = 0; c0 < c64; c0 += 65535)
    if ((rc = launch_copy_head(slice(c0, std::min<int64_t>(c64 - c0, 65535)), s))) return rc;

  // Side stream (when the lock-step class below has work for the caller's
  // stream, so the classes overlap): the zero fill of the lock-step
  // matrices (no kernel of theirs reads there), the fused launch of the mid
  // orders [cF, c64) (a workgroup per matrix, largest first), then the
  // n <= 64 queues.
  const bool split = mode == CHOL_MODE_MIXED;
  hipStream_t s2 = s;
  std::unique_lock<std::mutex> side_lock;
  if (cF > 0 && cz > cF && (rc = side_begin(s, &s2, side_lock))) return rc;
  for (int64_t c0 = 0; c0 < cF; c0 += 65535)
    if ((rc = launch_zero_upper_rows(slice(c0, std::min<int64_t>(cF - c0, 65535)), s2))) return rc;
  if (c64 > cF) {
    Batch q = slice(cF, c64 - cF);
    q.trailing_from_a = true;
    if ((rc = launch_fused(q, winv0 + cF * kNB * kNB, split, s2))) return rc;
  }
  // n <= 64: persistent wavefronts (4 per workgroup, 64/np matrices each).
  auto waves = [](int64_t cnt, int np) { return (cnt * np + 63) / 64; };
  if ((rc = launch_small_queue(base, c64, c32 - c64, 64, queues + 2,
                               (int)std::min<int64_t>((waves(c32 - c64, 64) + 3) / 4, 512), s2)))
    return rc;
  if ((rc = launch_small_queue(base, c32, c16 - c32, 32, queues + 3,
                               (int)std::min<int64_t>((waves(c16 - c32, 32) + 3) / 4, 1024), s2)))
    return rc;
  if ((rc = launch_small_queue(base, c16, cz - c16, 16, queues + 4,
                               (int)std::min<int64_t>((waves(cz - c16, 16) + 3) / 4, 1024), s2)))
    return rc;

  // The larger orders: lock-step blocked path over the prefix [0, cF), in chunks
  // of at most 65535 matrices (the kernels put the matrix on grid.y).
  for (int64_t c0 = 0; c0 < cF; c0 += 65535) {
    const int64_t cnt = std::min<int64_t>(cF - c0, 65535);
    const int32_t* cns = ns + c0;  // host copy of the chunk's orders (sorted, largest first)
    Batch q = slice(c0, cnt);
    Batch qf = q;  // first-touch updates (outer block 0: old values from A)
    qf.first = true;
    float* winv = winv0 + c0 * kNB * kNB;
    auto active = [&](int64_t above) {  // matrices of the chunk with n > above
      int64_t a = 0;
      while (a < cnt && cns[a] > above) ++a;
      return a;
    };
    int64_t nbo = tuned_outer_width(cns[0], cnt, mode);
    if (nbo < kNB || nbo % kNB != 0) nbo = cns[0] > 1024 ? 512 : 256;
    for (int K0 = 0; K0 < cns[0]; K0 += (int)nbo) {
      const int E = K0 + (int)nbo;  // end of the outer block
      // Inside the block, in the mixed form, left-looking when the strip's
      // tiles fill the GPU (each column block takes the block's earlier
      // columns at once, K = k0 - K0, just before its diagonal step: the
      // strip is read and written once per column block), since the BF16x3
      // updates are bound by that traffic; else right-looking K = 128
      // updates of the block's later columns.  512-matrix workload, outer
      // blocks of 512: mixed 2924 -> 2837 us; the plain form, whose FP32
      // updates are bound by the matrix cores, 4226 -> 4565 us, so it
      // stays right-looking.
      const bool left = split && (int64_t)((cns[0] - K0) / kNB) * active(K0) >= 256;
      // First touch in the first outer block (columns from 128 on): each
      // column block's left-looking update, the step-0 right-looking update
      // of the block's later columns, and the update of every column beyond
      // the block.
      Batch& q0 = K0 == 0 ? qf : q;
      for (int k0 = K0; k0 < E && k0 < cns[0]; k0 += kNB) {
        Batch& qk = k0 == 0 ? qf : q;
        q.batch = qf.batch = active(k0);
        if (left && k0 > K0 && (rc = launch_syrk(q0, k0, k0 + kNB, K0, k0 - K0, split, s))) return rc;
        if ((rc = launch_diag_factor(q, k0, winv, s))) return rc;
        const int64_t act2 = active(k0 + kNB);
        if (act2 == 0) break;
        q.batch = qf.batch = act2;
        if ((rc = launch_panel(q, k0, winv, split, s))) return rc;
        if (!left && k0 + kNB < E && (rc = launch_syrk(qk, k0 + kNB, E, k0, kNB, split, s))) return rc;
      }
      const int64_t actE = active(E);
      if (actE == 0) break;
      q.batch = qf.batch = actE;
      if ((rc = launch_syrk(q0, E, -1, K0, (int)nbo, split, s))) return rc;
    }
  }
  if ((rc = side_end(s, s2))) return rc;
  note_call(RC_RAGGED);
  return 0;
}

}  // extern "C"
