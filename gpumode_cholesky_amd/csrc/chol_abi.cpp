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
//                    blocked_potrf below.
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

// Tuning table (chol_set_tuning): rows {n_lo, n_hi, batch_lo, batch_hi, nbo}.
// nbo > 0: blocked path with that outer block width (a multiple of kNB = 128);
// nbo = -1: the whole-matrix LDS kernel (n <= 256 only).  First match wins.
struct TuneRow {
  int64_t n_lo, n_hi, b_lo, b_hi, nbo;
};
static std::mutex g_tune_mu;
static std::vector<TuneRow> g_tune;

// Default when no row matches: the blocked path with one outer block up to
// n = 256 (measured 1.2-2.2x faster than the whole-matrix LDS kernel from
// 16x100 to 4096x128 and 512x256), the LDS kernel only for batches of 8192
// and more (8192x96: 663 us against 721 us), then outer blocks of 256
// columns (512 from n = 12288, where the trailing matrix outgrows the
// 256 MiB Infinity Cache and the SYRK's C traffic dominates).
static int64_t pick_nbo(int64_t n, int64_t batch) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const TuneRow& r : g_tune)
      if (n >= r.n_lo && n <= r.n_hi && batch >= r.b_lo && batch <= r.b_hi) return r.nbo;
  }
  if (n <= kLdsMaxN) return batch >= 8192 ? -1 : (n <= kNB ? kNB : 2 * kNB);
  return n >= 12288 ? 512 : 256;
}

static size_t potrf_ws(int64_t batch, int64_t n) {
  if (n <= 64 || batch <= 0) return 0;
  return align256((size_t)batch * kNB * kNB * sizeof(float));
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
static int blocked_potrf(const Batch& p, float* winv, bool split, int nbo, hipStream_t s) {
  SideStream* side = nullptr;
  int rc = side_stream(side);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(side->enqueue);
  auto order = [](hipEvent_t ev, hipStream_t from, hipStream_t to) -> int {
    hipError_t e;
    if ((e = hipEventRecord(ev, from)) != hipSuccess || (e = hipStreamWaitEvent(to, ev, 0)) != hipSuccess) {
      set_error("look-ahead: event ordering failed: %s", hipGetErrorString(e));
      return (int)e;
    }
    return 0;
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
  const int idle = (int)((256 - p.batch % 256) % 256);
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
    if ((rc = order(side->to_side, s, side->s2))) return rc;
    if ((rc = launch_zero_upper(p, side->s2))) return rc;
    if (hipEventRecord(side->from_side, side->s2) != hipSuccess) return launch_status("hipEventRecord");
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
    const bool left = (int64_t)((n - K0) / kNB) * p.batch >= 256;
    // first touch in the first outer block: its diagonal block and panel at
    // k0 = 0, the first update of each of its column blocks (right-looking:
    // the step-0 update of all of them; left-looking: each one's update),
    // and its narrow and bulk updates
    const Batch& p0 = K0 == 0 ? pf : p;
    DiagBulk bulk{};
    const DiagBulk* carry = nullptr;
    for (int k0 = K0; k0 < E; k0 += kNB) {
      const Batch& pk = k0 == 0 ? pf : p;
      if (left && k0 > K0 && (rc = launch_syrk(p0, k0, k0 + kNB, K0, k0 - K0, split, s))) return rc;
      if ((rc = launch_diag_factor(pk, k0, winv, s, carry, zero_of(k0, zcur)))) return rc;
      carry = nullptr;
      if (k0 + kNB >= n) break;
      if ((rc = launch_panel(pk, k0, winv, split, s))) return rc;
      if (!left && k0 + kNB < E) {
        // Right-looking update by this panel: the next column block now (the
        // next diagonal step needs it); the columns beyond it with the next
        // diagonal-block launch, on the CUs its blocks leave idle, when they
        // fit there in a few rounds (fewer while the previous outer block's
        // bulk update runs on the side stream and takes CUs too).
        const int cn = k0 + 2 * kNB;
        const int rounds = bulk_pending ? 3 : 8;
        const bool hand_in = !p.Ap && p.batch <= 192 && vec_ok(pk);
        if (split && hand_in) {
          // mixed form: the whole update, the next block's own part included,
          // rides on the next diagonal-block launch
          bulk = DiagBulk{k0 + kNB, E, k0, pk.first, split, true};
          if (diag_bulk_tiles(p, bulk) <= (int64_t)rounds * (256 - p.batch)) {
            carry = &bulk;
            continue;
          }
        }
        bulk = DiagBulk{cn, E, k0, pk.first, split};
        if (cn < E && hand_in && diag_bulk_tiles(p, bulk) <= (int64_t)rounds * (256 - p.batch)) {
          if ((rc = launch_syrk(pk, k0 + kNB, cn, k0, kNB, split, s))) return rc;
          carry = &bulk;
        } else if ((rc = launch_syrk(pk, k0 + kNB, E, k0, kNB, split, s))) {
          return rc;
        }
      }
    }
    const int c0 = E;  // first column after this outer block
    if (c0 >= n) break;
    const int c1 = (n - c0) < nbo ? n : c0 + nbo;  // end of the next outer block
    if (bulk_pending) {
      if (hipStreamWaitEvent(s, side->from_side, 0) != hipSuccess) return launch_status("hipStreamWaitEvent");
      bulk_pending = false;
    }
    if ((rc = launch_syrk(p0, c0, c1, K0, W, split, s))) return rc;  // narrow
    if (c1 < n) {
      if ((rc = order(side->to_side, s, side->s2))) return rc;
      if ((rc = launch_syrk(p0, c1, n, K0, W, split, side->s2))) return rc;  // bulk
      if (hipEventRecord(side->from_side, side->s2) != hipSuccess) return launch_status("hipEventRecord");
      bulk_pending = true;
    }
  }
  if ((bulk_pending || side_pending) && hipStreamWaitEvent(s, side->from_side, 0) != hipSuccess)
    return launch_status("hipStreamWaitEvent");
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
  if (batch < 0 || n < 0) {
    set_error("chol_potrf_f32: batch (%lld) and n (%lld) must be >= 0", (long long)batch, (long long)n);
    return CHOL_E_ARG;
  }
  if (mode != CHOL_MODE_PLAIN && mode != CHOL_MODE_MIXED) {
    set_error("chol_potrf_f32: unknown mode %d", (int)mode);
    return CHOL_E_MODE;
  }
  if (batch == 0 || n == 0) return 0;
  if (n > 65536 || batch > (int64_t)1 << 31) {
    set_error("chol_potrf_f32: n=%lld or batch=%lld out of range", (long long)n, (long long)batch);
    return CHOL_E_ARG;
  }
  if (!A || !L) {
    set_error("chol_potrf_f32: null A or L");
    return CHOL_E_ARG;
  }
  if (ld < n || (batch > 1 && batch_stride < n * ld)) {
    set_error("chol_potrf_f32: ld=%lld must be >= n=%lld and batch_stride=%lld >= n*ld", (long long)ld,
              (long long)n, (long long)batch_stride);
    return CHOL_E_ARG;
  }
  const size_t need = potrf_ws(batch, n);
  if (need > 0 && (ws == nullptr || ws_bytes < need)) {
    set_error("chol_potrf_f32: workspace of %zu bytes required, got %zu", need, ws_bytes);
    return CHOL_E_WORKSPACE;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Batch p{A, L, batch, (int)n, batch > 1 ? batch_stride : n * ld, ld, info};
  if (n <= 64) return launch_small(p, s);
  int64_t nbo = pick_nbo(n, batch);
  if (nbo < 0 && n <= kLdsMaxN) return launch_lds(p, s);
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
    if ((rc = blocked_potrf(q, reinterpret_cast<float*>(ws), mode == CHOL_MODE_MIXED, (int)nbo, s))) return rc;
  }
  return 0;
}

int chol_set_tuning(const int64_t* table, int64_t rows) {
  clear_error();
  if (rows < 0 || (rows > 0 && !table)) {
    set_error("chol_set_tuning: bad table");
    return CHOL_E_ARG;
  }
  std::vector<TuneRow> t;
  for (int64_t i = 0; i < rows; ++i) {
    const int64_t* r = table + 5 * i;
    const int64_t nbo = r[4];
    if (!(nbo == -1 || (nbo >= kNB && nbo % kNB == 0 && nbo <= 4096)) || (nbo == -1 && r[1] > kLdsMaxN)) {
      set_error("chol_set_tuning: row %lld: nbo=%lld must be -1 (n <= %lld) or a multiple of %d", (long long)i,
                (long long)nbo, (long long)kLdsMaxN, kNB);
      return CHOL_E_ARG;
    }
    t.push_back(TuneRow{r[0], r[1], r[2], r[3], nbo});
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
  return backward(Lm, Lbar, Abar, batch, (int)n, have_linv ? Linv : nullptr, reinterpret_cast<float*>(ws),
                  reinterpret_cast<hipStream_t>(stream));
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
  return tri_inverse(Lm, Linv, batch, (int)n, reinterpret_cast<float*>(ws), reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
