// chol_abi.cpp — the extern "C" boundary (include/chol_amd.h): argument
// validation, workspace sizing and dispatch to the gfx950 kernels.
//
// Dispatch by order n and the tuning table (the reference dispatches by
// (batch, n) tables in custom_kernel, submission.py:4101-4268; here every
// (batch, n) takes a HIP route and there is no library fallback):
//   n <= 64        : chol_small.hip, whole matrices in wavefront registers /
//                    LDS;
//   64 < n <= 256  : the blocked path below a batch of 8192, else
//                    chol_lds.hip, whole matrix in one workgroup's LDS (the
//                    tuning table can pick either);
//   n >  256       : the blocked path: 128-wide diagonal blocks with explicit
//                    inverses (chol_diag.hip), MFMA panels and SYRK
//                    (chol_gemm.hip, chol_syrk256.hip), driven by
//                    blocked_potrf below;
//   any n > 64     : where the table says so, the fused kernel
//                    (k_potrf_fused, chol_diag.hip): one workgroup factors
//                    a whole matrix, for batches that fill the GPU.
#include <array>
#include <atomic>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "chol_common.h"

namespace chol {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Largest order of the whole-matrix-in-LDS kernel.
constexpr int64_t kLdsMaxN = 256;
// Table values of nbo that name a route instead of an outer width.
constexpr int64_t kRouteLds = -1, kRouteFused = -2;

// Tuning table (chol_set_tuning): rows {n_lo, n_hi, batch_lo, batch_hi, mode,
// nbo}; mode -1 matches both forms.  nbo > 0: blocked path with that outer
// block width (a multiple of kNB = 128); nbo = -1: the whole-matrix LDS
// kernel (n <= 256 only); nbo = -2: the fused kernel (n > 64 only).  First
// match wins.
struct TuneRow {
  int64_t n_lo, n_hi, b_lo, b_hi, mode, nbo;
};
static std::mutex g_tune_mu;
static std::vector<TuneRow> g_tune;

// Default when no row matches: the fused kernel for 128 < n <= 512 when
// the batch fills a 256-CU part one to a few times over (against the
// blocked path, dense inputs, both forms, one MI355X: 256..512 x 160..256
// 1.02-1.16x; 256..4096 x 300 1.16-1.55x; 256..2048 x 384 1.03-1.29x;
// 256..1024 x 448 1.04-1.29x; 256..512 x 512 0.98-1.12x; fewer matrices
// leave CUs idle while each runs its chain, more favour the blocked path's
// wide launches: 4096 x 256 0.86-0.88x); else the blocked path with one
// outer block up to n = 256 (measured 1.2-2.2x faster than the
// whole-matrix LDS kernel from 16x100 to 4096x128 and 512x256), the LDS
// kernel only for batches of 8192 and more (8192x96: 663 us against
// 721 us), then outer blocks of 256 columns (512 from n = 12288, where the
// trailing matrix outgrows the 256 MiB Infinity Cache and the SYRK's C
// traffic dominates).
static int64_t pick_nbo(int64_t n, int64_t batch, int32_t mode) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const TuneRow& r : g_tune)
      if (n >= r.n_lo && n <= r.n_hi && batch >= r.b_lo && batch <= r.b_hi && (r.mode < 0 || r.mode == mode))
        return r.nbo;
  }
  if (n > kNB && n <= 512 && batch >= 256) {
    const int64_t cap = n <= 256 ? 512 : n <= 384 ? 4096 : n <= 448 ? 1024 : 512;
    if (batch <= cap) return kRouteFused;
  }
  if (n <= kLdsMaxN) return batch >= 8192 ? -1 : (n <= kNB ? kNB : 2 * kNB);
  return n >= 12288 ? 512 : 256;
}

static size_t potrf_ws(int64_t batch, int64_t n) {
  if (n <= 64 || batch <= 0) return 0;
  return align256((size_t)batch * kNB * kNB * sizeof(float));
}

size_t potrf_workspace(int64_t batch, int64_t n) { return potrf_ws(batch, n); }

int64_t tuned_outer_width(int64_t n, int64_t batch, int32_t mode) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const TuneRow& r : g_tune)
    if (n >= r.n_lo && n <= r.n_hi && batch >= r.b_lo && batch <= r.b_hi && (r.mode < 0 || r.mode == mode))
      return r.nbo;
  return 0;
}

// Per-device side stream (default priority) and the two events that order
// it against the caller's stream, created on first use and kept for the
// life of the process.
struct SideStream {
  hipStream_t s2 = nullptr;
  hipEvent_t to_side = nullptr;   // recorded on the caller's stream
  hipEvent_t from_side = nullptr; // recorded on the side stream
  std::mutex enqueue;  // one call's event record / wait pairs at a time
};
static std::mutex g_side_mu;
static SideStream g_side[64];

static int side_stream(SideStream*& out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= 64) {
    set_error("side stream: hipGetDevice failed (%d)", (int)e);
    return e != hipSuccess ? (int)e : CHOL_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_side_mu);
  SideStream& r = g_side[dev];
  if (!r.s2) {
    if ((e = hipStreamCreateWithFlags(&r.s2, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&r.to_side, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&r.from_side, hipEventDisableTiming)) != hipSuccess) {
      set_error("side stream: creation failed: %s", hipGetErrorString(e));
      return (int)e;
    }
  }
  out = &r;
  return 0;
}

int side_begin(hipStream_t s, hipStream_t* s2, std::unique_lock<std::mutex>& lk) {
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (s && hipStreamIsCapturing(s, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone) {
    *s2 = s;  // under capture: one stream (see chol_potrf_f32)
    return 0;
  }
  SideStream* side = nullptr;
  int rc = side_stream(side);
  if (rc) return rc;
  lk = std::unique_lock<std::mutex>(side->enqueue);
  hipError_t e;
  if ((e = hipEventRecord(side->to_side, s)) != hipSuccess ||
      (e = hipStreamWaitEvent(side->s2, side->to_side, 0)) != hipSuccess) {
    set_error("side stream: event ordering failed: %s", hipGetErrorString(e));
    return (int)e;
  }
  *s2 = side->s2;
  return 0;
}

int side_end(hipStream_t s, hipStream_t s2) {
  if (s2 == s) return 0;
  SideStream* side = nullptr;
  int rc = side_stream(side);
  if (rc) return rc;
  hipError_t e;
  if ((e = hipEventRecord(side->from_side, s2)) != hipSuccess || (e = hipStreamWaitEvent(s, side->from_side, 0)) != hipSuccess) {
    set_error("side stream: event ordering failed: %s", hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

// Compute units of the current device (hipDeviceAttributeMultiprocessorCount),
// read once per device; the blocked driver sizes its idle-CU work by it.
static int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v > 0) return v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
  cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

// Dry-run record of a blocked factorisation (chol_plan_potrf): the launch
// sequence blocked_potrf would enqueue, with no HIP call made.  One record
// = 10 int64: {kind, stream (0 caller, 1 side), then per kind:
//   DIAG : k0, first, carry c0, carry c1, carry k0, carry flags (first | pre << 1), zero r0, zero r1
//          (carry / zero fields -1 when absent)
//   PANEL: k0, first
//   SYRK : col0, col1, k0, kl, first
//   ZERO (k_zero_upper), FORK (side waits for the caller), MARK (event on
//   the side stream), JOIN (caller waits for the last MARK): no fields}.
enum : int64_t { PK_DIAG = 1, PK_PANEL = 2, PK_SYRK = 3, PK_ZERO = 4, PK_FORK = 5, PK_MARK = 6, PK_JOIN = 7 };
struct Plan {
  std::vector<std::array<int64_t, 10>> rec;
  void add(std::array<int64_t, 10> r) { rec.push_back(r); }
};

// Calls per route since load (chol_route_counters).
static std::atomic<int64_t> g_route_calls[RC_COUNT];
void note_call(RouteCounter r) { g_route_calls[r].fetch_add(1, std::memory_order_relaxed); }

// Blocked factorisation with one outer block of look-ahead.  Within outer block
// [K0, K0+W) the 128-wide steps run on the caller's stream s.  Across outer
// blocks the update is right-looking: the block's trailing update (K = W)
// is split: the next outer block's columns (narrow, needed next) stay on s;
// the columns beyond (bulk) go to the side stream, where they overlap the
// next block's chain of small launches.  A narrow update first waits for the
// previous bulk update, which wrote the same columns; the caller's stream
// joins the side streams at the end.  The event pairs of one call are
// enqueued under the device's lock, so calls from several host threads do
// not interleave them (their work on the shared side streams runs in order).
// side_ok = false: everything on the caller's stream (no side stream, no
// events: the same launches in the same order, without the overlap).
static int blocked_potrf(const Batch& p, float* winv, bool split, int nbo, hipStream_t s, int cus, bool side_ok,
                         Plan* plan = nullptr) {
  SideStream* side = nullptr;
  int rc;
  std::unique_lock<std::mutex> lk;
  if (!plan && side_ok) {
    if ((rc = side_stream(side))) return rc;
    lk = std::unique_lock<std::mutex>(side->enqueue);
  }
  hipStream_t s2 = side ? side->s2 : s;
  // Every launch and stream-ordering step goes through these, which either
  // enqueue it or (plan) record it.
  auto fork = [&]() -> int {  // the side stream waits for the caller's work so far
    if (plan) return plan->add({PK_FORK, 1}), 0;
    if (!side) return 0;
    hipError_t e;
    if ((e = hipEventRecord(side->to_side, s)) != hipSuccess ||
        (e = hipStreamWaitEvent(side->s2, side->to_side, 0)) != hipSuccess) {
      set_error("look-ahead: event ordering failed: %s", hipGetErrorString(e));
      return (int)e;
    }
    return 0;
  };
  auto mark = [&]() -> int {  // the side stream's work so far, for a later join
    if (plan) return plan->add({PK_MARK, 1}), 0;
    if (!side) return 0;
    return hipEventRecord(side->from_side, side->s2) != hipSuccess ? launch_status("hipEventRecord") : 0;
  };
  auto join = [&]() -> int {  // the caller waits for the side stream up to the last mark
    if (plan) return plan->add({PK_JOIN, 0}), 0;
    if (!side) return 0;
    return hipStreamWaitEvent(s, side->from_side, 0) != hipSuccess ? launch_status("hipStreamWaitEvent") : 0;
  };
  auto zero_side = [&]() -> int {
    if (plan) return plan->add({PK_ZERO, 1}), 0;
    return launch_zero_upper(p, s2);
  };
  auto diag = [&](const Batch& q, int k0, const DiagBulk* c, const DiagZero* z) -> int {
    if (plan)
      return plan->add({PK_DIAG, 0, k0, q.first, c ? c->c0 : -1, c ? c->c1 : -1, c ? c->k0 : -1,
                        c ? (int64_t)(c->first | (c->pre << 1)) : -1, z ? z->r0 : -1, z ? z->r1 : -1}),
             0;
    return launch_diag_factor(q, k0, winv, s, c, z);
  };
  auto panel = [&](const Batch& q, int k0) -> int {
    if (plan) return plan->add({PK_PANEL, 0, k0, q.first}), 0;
    return launch_panel(q, k0, winv, split, s);
  };
  auto syrk = [&](const Batch& q, int c0, int c1, int k0, int kl, int on_side) -> int {
    if (plan) return plan->add({PK_SYRK, on_side, c0, c1, k0, kl, q.first}), 0;
    return launch_syrk(q, c0, c1, k0, kl, split, on_side ? s2 : s);
  };
  const int n = p.n;
  // No copy of tril(A) into L: the kernels that first touch a region read
  // it from A (pf).  L's upper triangle is zeroed by extra workgroups of the
  // diagonal-block launches, one per CU that the last round of the batch
  // leaves idle (every workgroup of that kernel takes a CU's LDS), each
  // launch an equal share of the area in consecutive rows, while a
  // workgroup's share stays under ~160 KiB (one workgroup per CU streams
  // far slower than the separate fill kernel); else by k_zero_upper.  Each
  // diagonal-block kernel stores its block whole.
  Batch pf = p;
  pf.first = true;
  bool bulk_pending = false;  // an outer-block bulk update on the side stream
  bool side_pending = false;  // the zero fill on the side stream
  const int nblk = (n + kNB - 1) / kNB;
  const int idle = (int)((cus - p.batch % cus) % cus);  // CUs the last round of diagonal blocks leaves idle
  const double share = 2.0 * (double)n * (double)n * (double)p.batch / nblk;  // bytes per launch
  const bool zero_in_diag = n > kNB && idle >= 32 && share <= 160.0 * 1024.0 * idle;
  std::vector<int> zrow;  // launch j zeros rows [zrow[j], zrow[j + 1])
  if (zero_in_diag) {
    // area above the diagonal of rows < R: sum (n - 1 - r) = R (n - 1) - R (R - 1) / 2
    auto area = [n](int64_t R) { return R * (n - 1) - R * (R - 1) / 2; };
    const int64_t total = area(n);
    zrow.resize(nblk + 1);
    zrow[0] = 0;
    for (int j = 1; j < nblk; ++j) {
      const int64_t target = total * j / nblk;
      int lo = zrow[j - 1], hi = n;  // first R with area(R) >= target
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (area(mid) >= target) hi = mid; else lo = mid + 1;
      }
      zrow[j] = lo;
    }
    zrow[nblk] = n;
  } else if (n > kNB) {
    // on the side stream, overlapping the first diagonal blocks: it only
    // writes zeros where no kernel of the factorisation reads, and the
    // diagonal-block kernels store the same zeros in their own blocks
    if ((rc = fork()) || (rc = zero_side()) || (rc = mark())) return rc;
    side_pending = true;
  }
  auto zero_of = [&](int k0, DiagZero& z) -> const DiagZero* {
    if (!zero_in_diag) return nullptr;
    const int j = k0 / kNB;
    z = DiagZero{zrow[j], zrow[j + 1], idle};
    return &z;
  };
  DiagZero zcur{};
  for (int K0 = 0; K0 < n; K0 += nbo) {
    const int W = (n - K0) < nbo ? (n - K0) : nbo;
    const int E = K0 + W;  // end of this outer block
    // Inside the block: left-looking when the strip is tall enough that its
    // updates are throughput work (each column block is brought up to date
    // with the block's earlier columns, K = k0 - K0, just before its
    // diagonal step: the strip is read and written once per column block);
    // right-looking (K = 128 updates of all later columns of the block, one
    // short launch each) when they are latency work.  These steps stay on
    // one stream: a cross-stream event pair costs ~10-20 us of idle time on
    // this stack, more than the microsecond-scale launches it could overlap.
    // From n = 24576 on every outer block is left-looking: there the
    // right-looking K = 128 updates of the block's later columns are launches
    // of thousands of short tiles that the side stream's bulk update
    // starves of CUs (1 x 32768: -5 % mixed, -2 % plain; 8192 and 16384
    // lose with it).
    const bool left = (int64_t)((n - K0) / kNB) * p.batch >= cus || (n >= 24576 && (n - K0) > kNB);
    // first touch in the first outer block: its diagonal block and panel at
    // k0 = 0, the first update of each of its column blocks (right-looking:
    // the step-0 update of all of them; left-looking: each one's update),
    // and its narrow and bulk updates
    const Batch& p0 = K0 == 0 ? pf : p;
    DiagBulk bulk{};
    const DiagBulk* carry = nullptr;
    for (int k0 = K0; k0 < E; k0 += kNB) {
      const Batch& pk = k0 == 0 ? pf : p;
      if (left && k0 > K0 && (rc = syrk(p0, k0, k0 + kNB, K0, k0 - K0, 0))) return rc;
      if ((rc = diag(pk, k0, carry, zero_of(k0, zcur)))) return rc;
      carry = nullptr;
      if (k0 + kNB >= n) break;
      if ((rc = panel(pk, k0))) return rc;
      if (!left && k0 + kNB < E) {
        // Right-looking update by this panel: the next column block now (the
        // next diagonal step needs it); the columns beyond it with the next
        // diagonal-block launch, on the CUs its blocks leave idle, when they
        // fit there in a few rounds (fewer while the previous outer block's
        // bulk update runs on the side stream and takes CUs too).
        const int cn = k0 + 2 * kNB;
        const int rounds = bulk_pending ? 3 : 8;
        const bool hand_in = !p.Ap && 4 * p.batch <= 3 * (int64_t)cus && vec_ok(pk);
        if (split && hand_in) {
          // mixed form: the whole update, the next block's own part included,
          // rides on the next diagonal-block launch
          bulk = DiagBulk{k0 + kNB, E, k0, pk.first, split, true};
          if (diag_bulk_tiles(p, bulk) <= (int64_t)rounds * (cus - p.batch)) {
            carry = &bulk;
            continue;
          }
        }
        bulk = DiagBulk{cn, E, k0, pk.first, split};
        if (cn < E && hand_in && diag_bulk_tiles(p, bulk) <= (int64_t)rounds * (cus - p.batch)) {
          if ((rc = syrk(pk, k0 + kNB, cn, k0, kNB, 0))) return rc;
          carry = &bulk;
        } else if ((rc = syrk(pk, k0 + kNB, E, k0, kNB, 0))) {
          return rc;
        }
      }
    }
    const int c0 = E;  // first column after this outer block
    if (c0 >= n) break;
    const int c1 = (n - c0) < nbo ? n : c0 + nbo;  // end of the next outer block
    if (bulk_pending) {
      if ((rc = join())) return rc;
      bulk_pending = false;
    }
    if ((rc = syrk(p0, c0, c1, K0, W, 0))) return rc;  // narrow
    if (c1 < n) {
      if ((rc = fork()) || (rc = syrk(p0, c1, n, K0, W, 1)) || (rc = mark())) return rc;  // bulk
      bulk_pending = true;
    }
  }
  if ((bulk_pending || side_pending) && (rc = join())) return rc;
  return 0;
}

// Argument checks and routing of one dense batched factorisation (the body
// of chol_potrf_f32 and of its allocation-free variant).  side_ok: the
// blocked route may use the library's side stream.
int potrf_dispatch(const char* who, const float* A, float* L, int64_t batch, int64_t n, int64_t batch_stride,
                   int64_t ld, int32_t mode, int32_t* info, void* ws, size_t ws_bytes, hipStream_t s, bool side_ok) {
  if (batch < 0 || n < 0) {
    set_error("%s: batch (%lld) and n (%lld) must be >= 0", who, (long long)batch, (long long)n);
    return CHOL_E_ARG;
  }
  if (mode != CHOL_MODE_PLAIN && mode != CHOL_MODE_MIXED && mode != CHOL_MODE_FP8) {
    set_error("%s: unknown mode %d", who, (int)mode);
    return CHOL_E_MODE;
  }
  if (batch == 0 || n == 0) return 0;
  // CHOL_MODE_FP8 is the mixed form (BF16x3 panels and narrow updates, its
  // tuning rows) with E4M3 bulk updates from kF8MinN on.
  const bool split = mode != CHOL_MODE_PLAIN;
  const bool fp8 = mode == CHOL_MODE_FP8;
  if (fp8) mode = CHOL_MODE_MIXED;
  if (n > 65536 || batch > (int64_t)1 << 31) {
    set_error("%s: n=%lld or batch=%lld out of range", who, (long long)n, (long long)batch);
    return CHOL_E_ARG;
  }
  if (!A || !L) {
    set_error("%s: null A or L", who);
    return CHOL_E_ARG;
  }
  if (ld < n || (batch > 1 && batch_stride < n * ld)) {
    set_error("%s: ld=%lld must be >= n=%lld and batch_stride=%lld >= n*ld", who, (long long)ld, (long long)n,
              (long long)batch_stride);
    return CHOL_E_ARG;
  }
  const size_t need = potrf_ws(batch, n);
  if (need > 0 && (ws == nullptr || ws_bytes < need)) {
    set_error("%s: workspace of %zu bytes required, got %zu", who, need, ws_bytes);
    return CHOL_E_WORKSPACE;
  }
  Batch p{A, L, batch, (int)n, batch > 1 ? batch_stride : n * ld, ld, info};
  p.f8 = fp8 && n >= kF8MinN;
  if (n <= 64) {
    int rc = launch_small(p, s);
    if (!rc) note_call(RC_SMALL);
    return rc;
  }
  int64_t nbo = pick_nbo(n, batch, mode);
  if (nbo == kRouteLds && n <= kLdsMaxN) {
    int rc = launch_lds(p, s);
    if (!rc) note_call(RC_LDS);
    return rc;
  }
  if (nbo == kRouteFused) {
    int rc = launch_zero_info(info, batch, s);
    if (rc) return rc;
    Batch q = p;
    q.first = true;  // A and L share ld: step 0 reads A where it first touches
    if ((rc = launch_fused(q, reinterpret_cast<float*>(ws), split, s))) return rc;
    note_call(RC_FUSED);
    return 0;
  }
  const int cus = device_cus();
  if (nbo < kNB || nbo % kNB != 0) nbo = 256;
  int rc = launch_zero_info(info, batch, s);
  if (rc) return rc;
  // The blocked kernels put the batch on grid.y (at most 65535 blocks).
  constexpr int64_t kMaxY = 65535;
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxY) {
    Batch q = p;
    q.batch = (batch - b0) < kMaxY ? (batch - b0) : kMaxY;
    q.A = A + b0 * p.bs;
    q.L = L + b0 * p.bs;
    q.info = info ? info + b0 : nullptr;
    if ((rc = blocked_potrf(q, reinterpret_cast<float*>(ws), split, (int)nbo, s, cus, side_ok)))
      return rc;
  }
  note_call(RC_BLOCKED);
  return 0;
}


}  // namespace chol

using namespace chol;

extern "C" {

int chol_abi_version(void) { return CHOL_ABI_VERSION; }

const char* chol_last_error(void) { return g_err; }

size_t chol_workspace_bytes(int64_t batch, int64_t n, int32_t mode) {
  (void)mode;
  return potrf_ws(batch, n);
}

int chol_potrf_f32(const float* A, float* L, int64_t batch, int64_t n, int64_t batch_stride, int64_t ld,
                   int32_t mode, int32_t* info, void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // Under stream capture the side stream stays out of the graph: a stream
  // forked into a capture belongs to it until the capture ends, and the
  // library's side stream is shared by every caller on the device.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  const bool capturing = s && hipStreamIsCapturing(s, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone;
  return potrf_dispatch("chol_potrf_f32", A, L, batch, n, batch_stride, ld, mode, info, ws, ws_bytes, s,
                        !capturing);
}

int chol_trailing_update_f32(const float* A, float* L, int64_t n, int64_t ld, int64_t col0, int64_t k0, int64_t kl,
                             int32_t mode, void* stream) {
  clear_error();
  if (mode != CHOL_MODE_PLAIN && mode != CHOL_MODE_MIXED && mode != CHOL_MODE_FP8) {
    set_error("chol_trailing_update_f32: unknown mode %d", (int)mode);
    return CHOL_E_MODE;
  }
  if (!A || !L || n <= 0 || n > 65536 || ld < n || col0 < 0 || col0 >= n || k0 < 0 || kl <= 0 || kl % 32 != 0 ||
      k0 + kl > col0) {
    set_error("chol_trailing_update_f32: bad arguments (n=%lld ld=%lld col0=%lld k0=%lld kl=%lld)", (long long)n,
              (long long)ld, (long long)col0, (long long)k0, (long long)kl);
    return CHOL_E_ARG;
  }
  Batch p{A, L, 1, (int)n, n * ld, ld, nullptr};
  p.f8 = mode == CHOL_MODE_FP8;
  return launch_syrk256(p, (int)col0, -1, (int)k0, (int)kl, mode != CHOL_MODE_PLAIN, vec_ok(p),
                        reinterpret_cast<hipStream_t>(stream));
}

int chol_route(int64_t batch, int64_t n, int32_t mode, int64_t* nbo_out) {
  int64_t nbo = 0;
  if (mode == CHOL_MODE_FP8) mode = CHOL_MODE_MIXED;  // the mixed form's routes and rows
  int route = CHOL_ROUTE_NONE;
  if (batch > 0 && n > 0) {
    if (n <= 64) {
      route = CHOL_ROUTE_SMALL;
    } else {
      nbo = pick_nbo(n, batch, mode);
      if (nbo == kRouteLds && n <= kLdsMaxN) {
        route = CHOL_ROUTE_LDS;
      } else if (nbo == kRouteFused) {
        route = CHOL_ROUTE_FUSED;
      } else {
        if (nbo < kNB || nbo % kNB != 0) nbo = 256;
        route = CHOL_ROUTE_BLOCKED;
      }
    }
  }
  if (nbo_out) *nbo_out = route == CHOL_ROUTE_BLOCKED ? nbo : 0;
  return route;
}

int chol_plan_potrf(int64_t batch, int64_t n, int64_t ld, int32_t mode, int32_t cus, int64_t* rec, int64_t cap,
                    int64_t* count) {
  clear_error();
  if (batch <= 0 || n <= 0 || ld < n || !count || (cap > 0 && !rec)) {
    set_error("chol_plan_potrf: bad arguments");
    return CHOL_E_ARG;
  }
  int64_t nbo = 0;
  if (chol_route(batch, n, mode, &nbo) != CHOL_ROUTE_BLOCKED) {
    *count = 0;
    return 0;
  }
  // Host-side stand-ins for 256-byte aligned device buffers (never
  // dereferenced): only alignment and sizes steer the schedule.
  const float* fa = reinterpret_cast<const float*>(uintptr_t(1) << 20);
  float* fl = reinterpret_cast<float*>(uintptr_t(1) << 40);
  const int64_t b = batch < 65535 ? batch : 65535;
  Batch p{fa, fl, b, (int)n, n * ld, ld, nullptr};
  Plan plan;
  int rc = blocked_potrf(p, nullptr, mode != CHOL_MODE_PLAIN, (int)nbo, nullptr, cus > 0 ? cus : 256, true, &plan);
  if (rc) return rc;
  *count = (int64_t)plan.rec.size();
  for (int64_t i = 0; i < *count && i < cap; ++i)
    for (int j = 0; j < 10; ++j) rec[10 * i + j] = plan.rec[i][j];
  return 0;
}

int64_t chol_route_counters(int64_t* out, int64_t cap) {
  for (int64_t i = 0; i < cap && i < RC_COUNT; ++i) out[i] = g_route_calls[i].load(std::memory_order_relaxed);
  return RC_COUNT;
}

int chol_set_tuning(const int64_t* table, int64_t rows) {
  clear_error();
  if (rows < 0 || (rows > 0 && !table)) {
    set_error("chol_set_tuning: bad table");
    return CHOL_E_ARG;
  }
  std::vector<TuneRow> t;
  for (int64_t i = 0; i < rows; ++i) {
    const int64_t* r = table + 6 * i;
    const int64_t mode = r[4], nbo = r[5];
    if (!(nbo == kRouteLds || nbo == kRouteFused || (nbo >= kNB && nbo % kNB == 0 && nbo <= 4096)) ||
        (nbo == kRouteLds && r[1] > kLdsMaxN) || (nbo == kRouteFused && r[0] <= 64)) {
      set_error("chol_set_tuning: row %lld: nbo=%lld must be -1 (n <= %lld), -2 (n > 64) or a multiple of %d",
                (long long)i, (long long)nbo, (long long)kLdsMaxN, kNB);
      return CHOL_E_ARG;
    }
    if (mode != -1 && mode != CHOL_MODE_PLAIN && mode != CHOL_MODE_MIXED) {
      set_error("chol_set_tuning: row %lld: mode=%lld must be -1, %d or %d", (long long)i, (long long)mode,
                CHOL_MODE_PLAIN, CHOL_MODE_MIXED);
      return CHOL_E_MODE;
    }
    t.push_back(TuneRow{r[0], r[1], r[2], r[3], mode, nbo});
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.swap(t);
  return 0;
}

size_t chol_backward_workspace_bytes(int64_t batch, int64_t n) {
  if (batch <= 0 || n <= 0) return 0;
  // P, T and (when L^-1 is not supplied) the recomputed inverse.
  return align256((size_t)3 * batch * n * n * sizeof(float));
}

static int check_square_batch(const char* who, const void* a, const void* b, int64_t batch, int64_t n) {
  if (batch < 0 || n < 0 || n > 32768) {
    set_error("%s: batch=%lld n=%lld out of range", who, (long long)batch, (long long)n);
    return CHOL_E_ARG;
  }
  if (batch > 0 && n > 0 && (!a || !b)) {
    set_error("%s: null pointer", who);
    return CHOL_E_ARG;
  }
  return 0;
}

int chol_backward_f32(const float* Lm, const float* Lbar, float* Abar, int64_t batch, int64_t n, const float* Linv,
                      int32_t have_linv, void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  int rc = check_square_batch("chol_backward_f32", Lm, Abar, batch, n);
  if (rc) return rc;
  if (batch == 0 || n == 0) return 0;
  if (!Lbar || (have_linv && !Linv)) {
    set_error("chol_backward_f32: null Lbar, or have_linv set without Linv");
    return CHOL_E_ARG;
  }
  const size_t need = (size_t)(have_linv ? 2 : 3) * batch * n * n * sizeof(float);
  if (!ws || ws_bytes < need) {
    set_error("chol_backward_f32: workspace of %zu bytes required, got %zu", need, ws_bytes);
    return CHOL_E_WORKSPACE;
  }
  rc = backward(Lm, Lbar, Abar, batch, (int)n, have_linv ? Linv : nullptr, reinterpret_cast<float*>(ws),
                reinterpret_cast<hipStream_t>(stream));
  if (!rc) note_call(RC_BACKWARD);
  return rc;
}

size_t chol_tri_inverse_workspace_bytes(int64_t batch, int64_t n) {
  if (batch <= 0 || n <= 0) return 0;
  return align256((size_t)batch * n * n * sizeof(float));
}

int chol_tri_inverse_f32(const float* Lm, float* Linv, int64_t batch, int64_t n, void* ws, size_t ws_bytes,
                         void* stream) {
  clear_error();
  int rc = check_square_batch("chol_tri_inverse_f32", Lm, Linv, batch, n);
  if (rc) return rc;
  if (batch == 0 || n == 0) return 0;
  const size_t need = (size_t)batch * n * n * sizeof(float);
  if (!ws || ws_bytes < need) {
    set_error("chol_tri_inverse_f32: workspace of %zu bytes required, got %zu", need, ws_bytes);
    return CHOL_E_WORKSPACE;
  }
  rc = tri_inverse(Lm, Linv, batch, (int)n, reinterpret_cast<float*>(ws), reinterpret_cast<hipStream_t>(stream));
  if (!rc) note_call(RC_TRI_INVERSE);
  return rc;
}

}  // extern "C"
