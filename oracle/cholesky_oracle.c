/*
 * cholesky_oracle.c — FP64 CPU oracle for the batched Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and bench.py's
 * CPU-baseline leg may load this library, and only as the checker (or the
 * timed CPU baseline) — never as the thing measured or shipped.
 *
 * What it restates: the reference's ground truth for this path is
 * `ref_kernel(data) = torch.linalg.cholesky_ex(data, check_errors=False).L`
 * (reference/reference.py:129-130), i.e. LAPACK potrf semantics: read the
 * lower triangle of each n x n matrix, return the unique lower-triangular L
 * with positive diagonal and A = L L^T, and report info = order of the first
 * leading minor that is not positive definite.  The arithmetic here is the
 * textbook Cholesky-Banachiewicz recurrence (row by row, dot-product form)
 *     L[j][j] = sqrt(A[j][j] - sum_{k<j} L[j][k]^2)
 *     L[i][j] = (A[i][j] - sum_{k<j} L[i][k] L[j][k]) / L[j][j],  i > j
 * carried in double precision on FP32 inputs.  Parity of this restatement is
 * pinned in tests/test_oracle.py against LAPACK (numpy.linalg.cholesky),
 * closed forms (diagonal, identity, n = 1, tridiagonal recurrence, A built as
 * L L^T from a chosen L) and planted mistakes.
 *
 * Also here: the FP64 reconstruction residual the reference's checker uses
 * (reference/reference.py:170-185, L1 matrix norm of L L^T - A) together with
 * the relative Frobenius residual ||A - L L^T||_F / ||A||_F, and a triangular
 * inverse.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Factor one matrix.  A: n x n FP32 row-major (lower triangle read).
 * L: n x n double, written in full (zeros above the diagonal).
 * Returns 0 or the 1-based order of the first non-positive pivot; on failure
 * the remaining columns are NaN (same convention as the GPU path). */
static int32_t potrf_one(const float* A, double* L, int64_t n) {
  int32_t info = 0;
  memset(L, 0, sizeof(double) * (size_t)(n * n));
  for (int64_t i = 0; i < n; ++i) {
    double* Li = L + i * n;
    for (int64_t j = 0; j <= i; ++j) {
      const double* Lj = L + j * n;
      double s = (double)A[i * n + j];
      for (int64_t k = 0; k < j; ++k) s -= Li[k] * Lj[k];
      if (i == j) {
        if (!(s > 0.0)) {
          if (info == 0) info = (int32_t)(j + 1);
          Li[j] = NAN;
        } else {
          Li[j] = sqrt(s);
        }
      } else {
        Li[j] = s / Lj[j];
      }
    }
  }
  if (info) {
    /* Make the failure unambiguous: everything from the failing column on. */
    for (int64_t i = info - 1; i < n; ++i)
      for (int64_t j = info - 1; j <= i; ++j) L[i * n + j] = NAN;
  }
  return info;
}

/* Batched FP64 factorisation of contiguous FP32 matrices. */
int oracle_potrf_f64(const float* A, double* L, int64_t batch, int64_t n,
                     int32_t* info) {
  if (batch < 0 || n < 0) return -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < batch; ++b) {
    int32_t r = potrf_one(A + b * n * n, L + b * n * n, n);
    if (info) info[b] = r;
  }
  return 0;
}

/* Same recurrence with every operation rounded to FP32 (the plain-FP32
 * arithmetic class of the GPU path); used to bound what FP32 can reach. */
int oracle_potrf_f32(const float* A, float* L, int64_t batch, int64_t n,
                     int32_t* info) {
  if (batch < 0 || n < 0) return -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < batch; ++b) {
    const float* a = A + b * n * n;
    float* l = L + b * n * n;
    int32_t inf = 0;
    memset(l, 0, sizeof(float) * (size_t)(n * n));
    for (int64_t i = 0; i < n; ++i) {
      for (int64_t j = 0; j <= i; ++j) {
        float s = a[i * n + j];
        for (int64_t k = 0; k < j; ++k) s = fmaf(-l[i * n + k], l[j * n + k], s);
        if (i == j) {
          if (!(s > 0.0f)) {
            if (!inf) inf = (int32_t)(j + 1);
            l[i * n + j] = NAN;
          } else {
            l[i * n + j] = sqrtf(s);
          }
        } else {
          l[i * n + j] = s / l[j * n + j];
        }
      }
    }
    if (info) info[b] = inf;
  }
  return 0;
}

/* Per-matrix FP64 residuals of an FP32 factor L against FP32 A:
 *   fro[b]   = ||L L^T - A||_F / ||A||_F
 *   l1res[b] = ||L L^T - A||_1     (max column sum, reference.py:133-134)
 *   l1a[b]   = ||A||_1
 * L is used as given (its upper triangle included), like the reference's
 * checker which forms output @ output.T (reference/reference.py:176). */
int oracle_residual(const float* A, const float* L, int64_t batch, int64_t n,
                    double* fro, double* l1res, double* l1a) {
  if (batch < 0 || n < 0) return -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < batch; ++b) {
    const float* a = A + b * n * n;
    const float* l = L + b * n * n;
    double* colres = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* cola = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double fr = 0.0, fa = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      for (int64_t j = 0; j < n; ++j) {
        double s = 0.0;
        const float* li = l + i * n;
        const float* lj = l + j * n;
        for (int64_t k = 0; k < n; ++k) s += (double)li[k] * (double)lj[k];
        double av = (double)a[i * n + j];
        double r = s - av;
        fr += r * r;
        fa += av * av;
        colres[j] += fabs(r);
        cola[j] += fabs(av);
      }
    }
    double mr = 0.0, ma = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (colres[j] > mr) mr = colres[j];
      if (cola[j] > ma) ma = cola[j];
    }
    fro[b] = fa > 0.0 ? sqrt(fr) / sqrt(fa) : sqrt(fr);
    l1res[b] = mr;
    l1a[b] = ma;
    free(colres);
    free(cola);
  }
  return 0;
}

/* FP64 inverse of a lower-triangular FP64 factor (forward substitution on
 * each column of the identity). */
int oracle_tri_inverse_f64(const double* L, double* X, int64_t batch,
                           int64_t n) {
  if (batch < 0 || n < 0) return -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < batch; ++b) {
    const double* l = L + b * n * n;
    double* x = X + b * n * n;
    memset(x, 0, sizeof(double) * (size_t)(n * n));
    for (int64_t c = 0; c < n; ++c) {
      for (int64_t i = c; i < n; ++i) {
        double s = (i == c) ? 1.0 : 0.0;
        for (int64_t k = c; k < i; ++k) s -= l[i * n + k] * x[k * n + c];
        x[i * n + c] = s / l[i * n + i];
      }
    }
  }
  return 0;
}
