/*
 * CPU oracle (test infrastructure only): FP64 block Clenshaw filter over a CSR
 * matrix, restating apply_poly_block! (polfed.jl:464-501,
 * chebyshev_subspace_iteration.jl:554-591) on column-major n x b blocks (Julia
 * layout).  Like the reference's threaded block operator (polfed.jl:209-219,
 * Threads.@threads over columns) it parallelises over block columns with
 * OpenMP; each column's SpMV and the three-term update run in one pass.
 *
 *   B_{K+1} = B_{K+2} = 0
 *   for k = K+1 .. 2:  TMP = c[k-1] Q + α A B1 + β B1 - B2;  (B1,B2,TMP) <- (TMP,B1,B2)
 *   Y = c0 Q + invd A B1 - c invd B1 - B2
 *   c = (smin+smax)/2, invd = 2/(smax-smin), α = 2 invd, β = -2 c invd
 */
#include <omp.h>
#include <stdint.h>
#include <string.h>

static void step(int64_t n, const int64_t* ptr, const int64_t* idx, const double* val,
                 const double* q, double ck, const double* b1, double al, double be,
                 const double* b2, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double ax = 0.0;
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) ax += val[p] * b1[idx[p]];
    out[i] = ck * q[i] + al * ax + be * b1[i] - b2[i];
  }
}

int oracle_clenshaw(int64_t n, int64_t b, const int64_t* ptr, const int64_t* idx,
                    const double* val, const double* Q, const double* coeffs, int K, double smin,
                    double smax, double* Y, double* work, int nthreads) {
  const double c = 0.5 * (smin + smax);
  const double invd = 2.0 / (smax - smin);
  const double al = 2.0 * invd, be = -2.0 * c * invd;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < b; ++j) {
    double* B1 = work + (size_t)(3 * j + 0) * n;
    double* B2 = work + (size_t)(3 * j + 1) * n;
    double* T = work + (size_t)(3 * j + 2) * n;
    memset(B1, 0, sizeof(double) * n);
    memset(B2, 0, sizeof(double) * n);
    const double* q = Q + (size_t)j * n;
    for (int k = K + 1; k >= 2; --k) {
      step(n, ptr, idx, val, q, coeffs[k - 1], B1, al, be, B2, T);
      double* t = B2;
      B2 = B1;
      B1 = T;
      T = t;
    }
    step(n, ptr, idx, val, q, coeffs[0], B1, invd, -c * invd, B2, Y + (size_t)j * n);
  }
  return 0;
}

/*
 * The same filter on row-major n x b blocks (element (i,c) at i*b + c), with
 * OpenMP over rows: each step reads the matrix once for all b columns, so a
 * wide block costs one matrix pass per step instead of one per column.
 */
static void step_rm(int64_t n, int64_t b, const int64_t* ptr, const int64_t* idx,
                    const double* val, const double* q, double ck, const double* b1, double al,
                    double be, const double* b2, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double* o = out + i * b;
    const double* qi = q + i * b;
    const double* x = b1 + i * b;
    const double* y = b2 ? b2 + i * b : 0;
    for (int64_t c = 0; c < b; ++c) o[c] = ck * qi[c] + be * x[c] - (y ? y[c] : 0.0);
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      const double w = al * val[p];
      const double* xr = b1 + idx[p] * b;
      for (int64_t c = 0; c < b; ++c) o[c] += w * xr[c];
    }
  }
}

int oracle_clenshaw_rm(int64_t n, int64_t b, const int64_t* ptr, const int64_t* idx,
                       const double* val, const double* Q, const double* coeffs, int K,
                       double smin, double smax, double* Y, double* work, int nthreads) {
  const double c = 0.5 * (smin + smax);
  const double invd = 2.0 / (smax - smin);
  const double al = 2.0 * invd, be = -2.0 * c * invd;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double* B1 = work;
  double* B2 = work + (size_t)n * b;
  double* T = work + (size_t)2 * n * b;
  memset(B1, 0, sizeof(double) * n * b);
  memset(B2, 0, sizeof(double) * n * b);
  for (int k = K + 1; k >= 2; --k) {
    step_rm(n, b, ptr, idx, val, Q, coeffs[k - 1], B1, al, be, B2, T);
    double* t = B2;
    B2 = B1;
    B1 = T;
    T = t;
  }
  step_rm(n, b, ptr, idx, val, Q, coeffs[0], B1, invd, -c * invd, B2, Y);
  return 0;
}
