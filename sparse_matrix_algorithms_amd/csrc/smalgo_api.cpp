// C ABI (include/smalgo.h): argument checking, error capture, ownership.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/smalgo.h"
#include "smalgo_coeffs.h"
#include "smalgo_mlp.h"
#include "smalgo_solver.h"

using namespace smalgo;

struct smalgo_ctx : Ctx {};
struct smalgo_mat : Mat {};
// smalgo_result stays an incomplete type: handles are Result objects.
static Result* res_of(smalgo_result* r) { return reinterpret_cast<Result*>(r); }
static const Result* res_of(const smalgo_result* r) { return reinterpret_cast<const Result*>(r); }
static smalgo_result* handle_of(Result* r) { return reinterpret_cast<smalgo_result*>(r); }

namespace {
std::atomic<long long> g_mat_ids{1};

template <typename F>
int guarded(smalgo_ctx* ctx, F&& f) {
  try {
    if (ctx) SMALGO_HIP(hipSetDevice(ctx->device));
    f();
    return SMALGO_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return SMALGO_ERR_INTERNAL;
  }
}
thread_local std::string g_io_err;
template <typename F>
int io_guarded(F&& f) {
  try {
    f();
    g_io_err.clear();
    return SMALGO_OK;
  } catch (const Error& e) {
    g_io_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_io_err = e.what();
    return SMALGO_ERR_INTERNAL;
  }
}
}  // namespace

extern "C" {

const char* smalgo_version(void) { return "smalgo 0.1 gfx950"; }

int smalgo_julia_float_string(double v, char* buf, int64_t len) {
  if (!buf || len <= 0) return SMALGO_ERR_ARG;
  try {
    const std::string s = format_julia_float(v);
    if ((int64_t)s.size() >= len) return SMALGO_ERR_ARG;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return SMALGO_OK;
  } catch (const std::exception&) {
    return SMALGO_ERR_INTERNAL;
  }
}

int smalgo_ctx_create(int device, smalgo_ctx** out) {
  if (!out) return SMALGO_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SMALGO_ERR_HIP;
  if (device < 0 || device >= ndev) return SMALGO_ERR_ARG;
  auto* c = new smalgo_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SMALGO_ERR_HIP;
  }
  if (rocblas_create_handle(&c->blas) != rocblas_status_success ||
      rocblas_set_stream(c->blas, c->stream) != rocblas_status_success) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return SMALGO_ERR_INTERNAL;
  }
  // device address of the no-allocation entry point's scratch, once per
  // device (a failure is reported by that entry point, not here)
  noalloc_scratch_init(device);
  *out = c;
  return SMALGO_OK;
}

void smalgo_ctx_destroy(smalgo_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  if (ctx->blas) rocblas_destroy_handle(ctx->blas);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* smalgo_last_error(const smalgo_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int smalgo_csr_upload(smalgo_ctx* ctx, int64_t n, const int64_t* ptr, const int64_t* idx,
                      const double* val, int index_base, smalgo_mat** out) {
  if (!ctx || !out) return SMALGO_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    if (n < 0 || n >= (int64_t)std::numeric_limits<int>::max()) throw Error(-1, "n out of range");
    if (!ptr) throw Error(-1, "ptr is null");
    if (index_base != 0 && index_base != 1) throw Error(-1, "index_base must be 0 or 1");
    const int64_t nnz = ptr[n] - ptr[0];
    if (nnz < 0 || nnz >= (int64_t)std::numeric_limits<int>::max())
      throw Error(-1, "nnz out of range");
    if (nnz > 0 && (!idx || !val)) throw Error(-1, "idx/val are null");
    // the whole pointer array first, in int64: every offset in [0, nnz] and
    // nondecreasing, before any idx/val element is touched
    for (int64_t i = 0; i < n; ++i)
      if (ptr[i + 1] < ptr[i]) throw Error(-1, "ptr must be nondecreasing");
    std::vector<int> rp(n + 1), ci(nnz);
    std::vector<double> v64(val, val + nnz);
    std::vector<float> v32(nnz);
    std::vector<double> dg(n, 0.0), rs(n, 0.0);
    for (int64_t i = 0; i <= n; ++i) rp[i] = (int)(ptr[i] - ptr[0]);
    for (int64_t i = 0; i < n; ++i) {
      for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const int64_t j = idx[p] - index_base;
        if (j < 0 || j >= n) throw Error(-1, "column index out of range");
        ci[p] = (int)j;
        v32[p] = (float)v64[p];
        if (j == i)
          dg[i] += v64[p];
        else
          rs[i] += std::fabs(v64[p]);
      }
    }
    auto* m = new smalgo_mat();
    m->ctx = ctx;
    m->id = g_mat_ids++;
    {
      // FNV-1a over n, the row pointers and the column indices
      uint64_t h = 1469598103934665603ull;
      auto mix = [&h](uint64_t v) {
        for (int k = 0; k < 8; ++k) {
          h ^= (v >> (8 * k)) & 0xff;
          h *= 1099511628211ull;
        }
      };
      mix((uint64_t)n);
      for (int64_t i = 0; i <= n; ++i) mix((uint64_t)(uint32_t)rp[i]);
      for (int64_t q = 0; q < nnz; ++q) mix((uint64_t)(uint32_t)ci[q]);
      m->structure = (long long)h;
    }
    m->A.n = (int)n;
    m->A.nnz = (int)nnz;
    // Gershgorin bounds on the host copy (polfed.jl:254-266)
    double lo = std::numeric_limits<double>::infinity(), hi = -lo;
    for (int64_t i = 0; i < n; ++i) {
      lo = std::min(lo, dg[i] - rs[i]);
      hi = std::max(hi, dg[i] + rs[i]);
    }
    m->A.gersh_min = lo;
    m->A.gersh_max = hi;
    // exact symmetry of pattern and values (issymmetric): compare with the transpose
    {
      std::vector<int> cnt(n + 1, 0), tp;
      for (int64_t p = 0; p < nnz; ++p) cnt[ci[p] + 1]++;
      for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
      std::vector<int> tcol(nnz), fill(cnt.begin(), cnt.end() - 1);
      std::vector<double> tval(nnz);
      for (int64_t i = 0; i < n; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
          const int d = fill[ci[p]]++;
          tcol[d] = (int)i;
          tval[d] = v64[p];
        }
      bool sym = true;
      for (int64_t i = 0; i < n && sym; ++i) {
        // compare row i of A with row i of Aᵀ as (col,val) multisets, both sorted by col
        std::vector<std::pair<int, double>> a, b;
        for (int p = rp[i]; p < rp[i + 1]; ++p) a.push_back({ci[p], v64[p]});
        for (int p = cnt[i]; p < cnt[i + 1]; ++p) b.push_back({tcol[p], tval[p]});
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        // merge duplicates
        auto merge = [](std::vector<std::pair<int, double>>& x) {
          std::vector<std::pair<int, double>> y;
          for (auto& e : x) {
            if (!y.empty() && y.back().first == e.first)
              y.back().second += e.second;
            else
              y.push_back(e);
          }
          std::vector<std::pair<int, double>> z;
          for (auto& e : y)
            if (e.second != 0.0) z.push_back(e);
          x.swap(z);
        };
        merge(a);
        merge(b);
        sym = (a == b);
      }
      m->symmetric = sym ? 1 : 0;
    }
    try {
      SMALGO_HIP(hipMalloc((void**)&m->A.rowptr, sizeof(int) * (n + 1)));
      SMALGO_HIP(hipMalloc((void**)&m->A.col, sizeof(int) * std::max<int64_t>(nnz, 1)));
      SMALGO_HIP(hipMalloc((void**)&m->A.val64, sizeof(double) * std::max<int64_t>(nnz, 1)));
      SMALGO_HIP(hipMalloc((void**)&m->A.val32, sizeof(float) * std::max<int64_t>(nnz, 1)));
      SMALGO_HIP(hipMemcpy(m->A.rowptr, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
      if (nnz) {
        SMALGO_HIP(hipMemcpy(m->A.col, ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
        SMALGO_HIP(hipMemcpy(m->A.val64, v64.data(), sizeof(double) * nnz, hipMemcpyHostToDevice));
        SMALGO_HIP(hipMemcpy(m->A.val32, v32.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
      }
    } catch (...) {
      smalgo_mat_free(m);
      throw;
    }
    *out = m;
  });
}

void smalgo_mat_free(smalgo_mat* m) {
  if (!m) return;
  if (m->ctx) (void)hipSetDevice(m->ctx->device);
  if (m->A.rowptr) (void)hipFree(m->A.rowptr);
  if (m->A.col) (void)hipFree(m->A.col);
  if (m->A.val64) (void)hipFree(m->A.val64);
  if (m->A.val32) (void)hipFree(m->A.val32);
  delete m;
}

int64_t smalgo_mat_n(const smalgo_mat* m) { return m ? m->A.n : -1; }
int64_t smalgo_mat_nnz(const smalgo_mat* m) { return m ? m->A.nnz : -1; }

int smalgo_mat_is_symmetric(const smalgo_mat* m, int* s) {
  if (!m || !s) return SMALGO_ERR_ARG;
  *s = m->symmetric;
  return SMALGO_OK;
}

int smalgo_gershgorin_bounds(const smalgo_mat* m, double* smin, double* smax) {
  if (!m || !smin || !smax) return SMALGO_ERR_ARG;
  *smin = m->A.gersh_min;
  *smax = m->A.gersh_max;
  return SMALGO_OK;
}

// ---- coefficients --------------------------------------------------------
// The coefficient calls have no context to hold a message: an exception
// (allocation failure for a huge K or nsamp) becomes a status code here and
// never crosses the C ABI.
extern "C++" {
template <typename F>
static int put(double* out, F&& make) {
  if (!out) return SMALGO_ERR_ARG;
  try {
    const std::vector<double> v = make();
    std::copy(v.begin(), v.end(), out);
    return SMALGO_OK;
  } catch (const Error& e) {
    return e.code;
  } catch (...) {
    return SMALGO_ERR_INTERNAL;
  }
}
}
int smalgo_jackson_kernel(int K, double* g) {
  return put(g, [&] { return jackson_kernel(K); });
}
int smalgo_cheb_delta_coeffs(double sh, int K, int jackson, double* c) {
  if (K < 0) return SMALGO_ERR_ARG;
  return put(c, [&] { return cheb_delta_coeffs(sh, K, jackson != 0); });
}
int smalgo_cheb_window_coeffs(double a, double b, int K, int jackson, double* c) {
  if (K < 0) return SMALGO_ERR_ARG;
  return put(c, [&] { return cheb_window_coeffs(a, b, K, jackson != 0); });
}
int smalgo_cheb_smooth_window_coeffs(double a, double b, int K, int jackson, int nsamp,
                                     double delta, double* c) {
  if (K < 0) return SMALGO_ERR_ARG;
  return put(c, [&] { return cheb_smooth_window_coeffs(a, b, K, jackson != 0, nsamp, delta); });
}
int smalgo_build_coeffs(int degree, double smin, double smax, double sigma, double emin,
                        double emax, int window_type, int jackson, double* c) {
  if (degree < 0 || window_type < 0 || window_type > 2) return SMALGO_ERR_ARG;
  return put(c, [&] {
    return build_coeffs(degree, smin, smax, sigma, emin, emax, window_type, jackson != 0);
  });
}

// ---- filter ---------------------------------------------------------------
static int apply_poly_common(smalgo_ctx* ctx, const Op& op, int64_t b, const double* Q, int64_t ldq,
                             const double* coeffs, int K, double smin, double smax, int precision,
                             double* Y, int64_t ldy) {
  return guarded(ctx, [&] {
    if (b < 0 || K < 0 || !coeffs || (b > 0 && (!Q || !Y))) throw Error(-1, "bad arguments");
    if (ldq < op.n || ldy < op.n) throw Error(-1, "leading dimension smaller than n");
    if (!(smax > smin)) throw Error(-1, "need smin < smax");
    const int n = op.n, bb = (int)b;
    const int ld = (bb + 3) / 4 * 4;
    std::vector<double> co(coeffs, coeffs + K + 1);
    DBuf<double> cm((size_t)n * std::max(bb, 1));
    for (int j = 0; j < bb; ++j)
      SMALGO_HIP(hipMemcpyAsync(cm.p + (size_t)j * n, Q + (size_t)j * ldq, sizeof(double) * n,
                                hipMemcpyHostToDevice, ctx->stream));
    if (precision == SMALGO_PREC_FP32_REFINE64) {
      if (op.is_callback()) throw Error(-1, "matrix-free filter runs in FP64 only");
      DBuf<float> q((size_t)n * ld), y((size_t)n * ld), b0((size_t)n * ld), b1((size_t)n * ld);
      DBuf<double> y64((size_t)n * ld);
      cm_to_rm<double, float>(ctx->stream, n, bb, cm.p, n, q.p, ld);
      apply_filter<float>(*ctx, op, bb, q.p, ld, co, smin, smax, y.p, ld, b0.p, b1.p, ld, nullptr);
      copy_block<float, double>(ctx->stream, n, bb, y.p, ld, y64.p, ld);
      rm_to_cm<double>(ctx->stream, n, bb, y64.p, ld, nullptr, cm.p, n);
    } else {
      DBuf<double> q((size_t)n * ld), y((size_t)n * ld), b0((size_t)n * ld), b1((size_t)n * ld);
      DBuf<double> pb;
      if (op.is_callback()) pb.alloc((size_t)n * ld);
      cm_to_rm<double, double>(ctx->stream, n, bb, cm.p, n, q.p, ld);
      apply_filter<double>(*ctx, op, bb, q.p, ld, co, smin, smax, y.p, ld, b0.p, b1.p, ld, pb.p);
      rm_to_cm<double>(ctx->stream, n, bb, y.p, ld, nullptr, cm.p, n);
    }
    for (int j = 0; j < bb; ++j)
      SMALGO_HIP(hipMemcpyAsync(Y + (size_t)j * ldy, cm.p + (size_t)j * n, sizeof(double) * n,
                                hipMemcpyDeviceToHost, ctx->stream));
    SMALGO_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int smalgo_apply_poly_block_noalloc(smalgo_ctx* ctx, const smalgo_mat* A, int64_t b,
                                    const double* dQ, int64_t ldq, const double* coeffs, int K,
                                    double smin, double smax, double* dY, int64_t ldy) {
  if (!ctx) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    if (!A) throw Error(-1, "A is null");
    if (b < 0 || K < 0 || !coeffs || (b > 0 && (!dQ || !dY))) throw Error(-1, "bad arguments");
    if (ldq < b || ldy < b) throw Error(-1, "leading dimension smaller than b (row-major blocks)");
    if (!(smax > smin)) throw Error(-1, "need smin < smax");
    if (b == 0 || A->A.n == 0) return;
    if (b > std::numeric_limits<int>::max() || ldq > std::numeric_limits<int>::max() ||
        ldy > std::numeric_limits<int>::max())
      throw Error(-1, "block width or leading dimension out of range");
    const int64_t n = A->A.n;
    size_t cap = 0;
    double* scr = noalloc_scratch(ctx->device, &cap);
    const double need = (double)n * (double)ldy * sizeof(double);
    if (need > (double)cap)
      throw Error(-1, "the no-allocation filter needs n*ldy*8 = " + std::to_string((long long)need) +
                          " bytes of scratch; its device scratch holds " + std::to_string(cap) +
                          " bytes (use smalgo_apply_poly_block)");
    const char* q0 = reinterpret_cast<const char*>(dQ);
    const char* y0 = reinterpret_cast<const char*>(dY);
    if (q0 < y0 + (size_t)n * ldy * sizeof(double) && y0 < q0 + (size_t)n * ldq * sizeof(double))
      throw Error(-1, "Q and Y overlap");
    Op op;
    op.A = &A->A;
    op.n = (int)n;
    op.mat_id = A->structure;
    apply_filter_noalloc(*ctx, op, (int)b, dQ, (int)ldq, coeffs, K, smin, smax, dY, (int)ldy, scr);
    SMALGO_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int smalgo_apply_poly_block(smalgo_ctx* ctx, const smalgo_mat* A, int64_t b, const double* Q,
                            int64_t ldq, const double* coeffs, int K, double smin, double smax,
                            int precision, double* Y, int64_t ldy) {
  if (!ctx || !A) return SMALGO_ERR_ARG;
  Op op;
  op.A = &A->A;
  op.n = A->A.n;
  op.mat_id = A->structure;
  return apply_poly_common(ctx, op, b, Q, ldq, coeffs, K, smin, smax, precision, Y, ldy);
}

int smalgo_apply_poly_block_mf(smalgo_ctx* ctx, int64_t n, smalgo_matblock_fn fn, void* user,
                               int64_t b, const double* Q, int64_t ldq, const double* coeffs,
                               int K, double smin, double smax, double* Y, int64_t ldy) {
  if (!ctx || !fn || n < 0) return SMALGO_ERR_ARG;
  Op op;
  op.fn = fn;
  op.user = user;
  op.n = (int)n;
  return apply_poly_common(ctx, op, b, Q, ldq, coeffs, K, smin, smax, SMALGO_PREC_FP64, Y, ldy);
}

// ---- block kernels -------------------------------------------------------
int smalgo_orth_against(smalgo_ctx* ctx, int64_t n, int64_t b, double* W, int64_t ldw, int64_t m,
                        const double* Q, int64_t ldq, int precision) {
  if (!ctx) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    if (n < 0 || b < 0 || m < 0 || ldw < n || ldq < n || (b && !W) || (m && !Q))
      throw Error(-1, "bad arguments");
    orth_against_host(*ctx, (int)n, (int)b, W, (int)ldw, (int)m, Q, (int)ldq, precision);
  });
}

int smalgo_block_qr(smalgo_ctx* ctx, int64_t n, int64_t b, const double* W, int64_t ldw,
                    double rtol, int precision, double* Q, int64_t ldq, double* R, int64_t ldr,
                    int* rank) {
  if (!ctx || !rank) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    if (n < 1 || b < 1 || b > n || ldw < n || ldq < n || !W || !Q || (R && ldr < b))
      throw Error(-1, "bad arguments");
    block_qr_host(*ctx, (int)n, (int)b, W, (int)ldw, rtol, precision, Q, (int)ldq, R, (int)ldr,
                  rank);
  });
}

int smalgo_build_X(smalgo_ctx* ctx, int64_t n, int64_t M, int64_t k, const double* V, int64_t ldv,
                   const double* U, int64_t ldu, double* X, int64_t ldx) {
  if (!ctx) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    if (n < 0 || M < 0 || k < 0 || ldv < n || ldu < M || ldx < n) throw Error(-1, "bad arguments");
    build_X_host(*ctx, (int)n, (int)M, (int)k, V, (int)ldv, U, (int)ldu, X, (int)ldx);
  });
}

static bool make_op(const smalgo_mat* A, smalgo_matblock_fn fn, void* user, int64_t n, Op& op);

int smalgo_rayleigh(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                    int64_t n, int64_t k, const double* X, int64_t ldx, double* lam, double* res) {
  if (!ctx) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, fn, user, n, op)) throw Error(-1, "exactly one of A or fn must be given (with matching n)");
    if (k < 0 || ldx < op.n || (k && (!X || !lam || !res))) throw Error(-1, "bad arguments");
    rayleigh_host(*ctx, op, (int)k, X, (int)ldx, lam, res);
  });
}

int smalgo_residuals(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                     int64_t n, int64_t k, const double* X, int64_t ldx, const double* lam,
                     double* res) {
  if (!ctx) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, fn, user, n, op)) throw Error(-1, "exactly one of A or fn must be given (with matching n)");
    if (k < 0 || ldx < op.n || (k && (!X || !lam || !res))) throw Error(-1, "bad arguments");
    residuals_host(*ctx, op, (int)k, X, (int)ldx, lam, res);
  });
}

int smalgo_refine_original_rr(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn,
                              void* user, int64_t n, int64_t k, const double* X, int64_t ldx,
                              smalgo_result** out) {
  if (!ctx || !out) return SMALGO_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, fn, user, n, op)) throw Error(-1, "exactly one of A or fn must be given (with matching n)");
    if (k < 1 || ldx < op.n || !X) throw Error(-1, "bad arguments");
    *out = handle_of(refine_rr_host(*ctx, op, (int)k, X, (int)ldx));
  });
}

int smalgo_subspace_rr(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                       int64_t n, int64_t k, const double* Y, int64_t ldy, double* Q,
                       int64_t ldq, smalgo_result** out) {
  if (!ctx || !out) return SMALGO_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, fn, user, n, op)) throw Error(-1, "exactly one of A or fn must be given (with matching n)");
    if (k < 1 || k > op.n || ldy < op.n || !Y || (Q && ldq < op.n)) throw Error(-1, "bad arguments");
    *out = handle_of(subspace_rr_host(*ctx, op, (int)k, Y, (int)ldy, Q, (int)ldq));
  });
}

int smalgo_kpm_moments(smalgo_ctx* ctx, const smalgo_mat* A, int nmom, int R, uint64_t seed,
                       double smin, double smax, double* mu) {
  if (!ctx || !A || !mu) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    Op op;
    op.A = &A->A;
    op.n = A->A.n;
    op.mat_id = A->structure;
    kpm_moments(*ctx, op, nmom, R, seed, smin, smax, mu);
  });
}

int smalgo_copy_bench(smalgo_ctx* ctx, int64_t bytes, int reps, double* seconds) {
  if (!ctx || !seconds || bytes < 16 || reps <= 0) return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    const size_t nb = (size_t)bytes / 16 * 16;
    DBuf<float> a(nb / 4), b(nb / 4);
    fill_block<float>(ctx->stream, 1, (int)(nb / 4), a.p, (int)(nb / 4), 1.0f);
    stream_copy(ctx->stream, a.p, b.p, nb);
    Event e0, e1;
    SMALGO_HIP(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps; ++r) stream_copy(ctx->stream, a.p, b.p, nb);
    SMALGO_HIP(hipEventRecord(e1, ctx->stream));
    SMALGO_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SMALGO_HIP(hipEventElapsedTime(&ms, e0, e1));
    *seconds = ms * 1e-3 / reps;
  });
}

// ---- MLP -------------------------------------------------------------------
struct smalgo_mlp {
  smalgo_ctx* ctx;
  Mlp* m;
};

int smalgo_mlp_create(smalgo_ctx* ctx, int nin, int nh, int nout, uint64_t seed, smalgo_mlp** out) {
  if (!ctx || !out || nin < 1 || nh < 1 || nout < 1) return SMALGO_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    auto* h = new smalgo_mlp{ctx, nullptr};
    try {
      h->m = new Mlp(*ctx, nin, nh, nout, seed);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
void smalgo_mlp_free(smalgo_mlp* h) {
  if (!h) return;
  (void)hipSetDevice(h->ctx->device);
  delete h->m;
  delete h;
}
int64_t smalgo_mlp_num_params(const smalgo_mlp* h) { return h ? (int64_t)h->m->np : -1; }
int smalgo_mlp_get_params(const smalgo_mlp* h, float* p) {
  if (!h || !p) return SMALGO_ERR_ARG;
  return guarded(h->ctx, [&] { h->m->get_params(p); });
}
int smalgo_mlp_set_params(smalgo_mlp* h, const float* p) {
  if (!h || !p) return SMALGO_ERR_ARG;
  return guarded(h->ctx, [&] { h->m->set_params(p); });
}
int smalgo_mlp_predict(smalgo_mlp* h, const float* X, int B, float* Y) {
  if (!h || !X || !Y || B < 1) return SMALGO_ERR_ARG;
  return guarded(h->ctx, [&] { h->m->predict(X, B, Y); });
}
int smalgo_mlp_gradient(smalgo_mlp* h, const float* X, const float* T, int B, float* grad,
                        float* loss) {
  if (!h || !X || !T || B < 1) return SMALGO_ERR_ARG;
  return guarded(h->ctx, [&] {
    const float l = h->m->gradient(X, T, B, grad);
    if (loss) *loss = l;
  });
}
int smalgo_mlp_train(smalgo_mlp* h, const float* X, const float* T, int B, int epochs, float lr,
                     float* loss_hist, float* loss) {
  if (!h || !X || !T || B < 1 || epochs < 0) return SMALGO_ERR_ARG;
  return guarded(h->ctx, [&] {
    const float l = h->m->train(X, T, B, epochs, lr, loss_hist);
    if (loss) *loss = l;
  });
}

// ---- drivers ---------------------------------------------------------------
void smalgo_polfed_default_params(smalgo_polfed_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->s = 8;
  p->has_bounds = 0;
  p->sigma = 0.0;
  p->emin = -std::numeric_limits<double>::infinity();
  p->emax = std::numeric_limits<double>::infinity();
  p->tol = 1e-10;
  p->overextract = 2;
  p->refine = 1;
  p->window_type = SMALGO_WINDOW_STEP;
  p->jackson = 1;
  p->mmax = 0;
  p->rtol = 1e-12;
  p->reorth_passes = 1;
  p->extra_keep = 8;
  p->seed = 1;
  p->chunk = 64;
  p->show_progress = 0;
  p->precision = SMALGO_PREC_FP64;
  p->refine_guard = -1;
  p->refine_degree = -1;
  p->refine_maxiter = 3;
  p->refine_tol = -1.0;
  p->refine_guard_degree = -1;
}

void smalgo_cheb_default_params(smalgo_cheb_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->has_bounds = 0;
  p->sigma = 0.0;
  p->m0 = 200;
  p->degree = 80;
  p->maxiter = 8;
  p->tol = 1e-10;
  p->res_gate = 1e-6;
  p->jackson = 1;
  p->debug = 0;
  p->two_hit = 1;
  p->seed = 1;
  p->window_type = SMALGO_WINDOW_STEP;
}

static bool make_op(const smalgo_mat* A, smalgo_matblock_fn fn, void* user, int64_t n, Op& op) {
  if ((A == nullptr) == (fn == nullptr)) return false;
  if (A) {
    op.A = &A->A;
    op.n = A->A.n;
    op.mat_id = A->structure;
    if (n >= 0 && n != A->A.n) return false;
  } else {
    if (n <= 0 || n >= (int64_t)std::numeric_limits<int>::max()) return false;
    op.fn = fn;
    op.user = user;
    op.n = (int)n;
  }
  return true;
}

int smalgo_polfed(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                  int64_t n, int nev, int degree, const smalgo_polfed_params* p, const double* Q1,
                  smalgo_result** out) {
  if (!ctx || !out) return SMALGO_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, fn, user, n, op)) throw Error(-1, "exactly one of A or fn must be given (with matching n)");
    smalgo_polfed_params prm;
    if (p)
      prm = *p;
    else
      smalgo_polfed_default_params(&prm);
    Result* r = polfed_solve(*ctx, op, nev, degree, prm, Q1);
    *out = handle_of(r);
  });
}

int smalgo_polfed_slices(smalgo_ctx* ctx, const smalgo_mat* A, int nslices,
                         const smalgo_slice* slices, const smalgo_polfed_params* p, int use_graph,
                         smalgo_result** out) {
  if (!ctx || !A || !out || !slices || nslices < 1) return SMALGO_ERR_ARG;
  for (int i = 0; i < nslices; ++i) out[i] = nullptr;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, nullptr, nullptr, A->A.n, op)) throw Error(-1, "bad matrix");
    smalgo_polfed_params prm;
    if (p)
      prm = *p;
    else
      smalgo_polfed_default_params(&prm);
    std::vector<SliceSpec> sl(nslices);
    for (int i = 0; i < nslices; ++i)
      sl[i] = SliceSpec{slices[i].emin, slices[i].emax, slices[i].sigma, slices[i].nev,
                        slices[i].s,    slices[i].degree, slices[i].mmax};
    std::vector<Result*> rs = polfed_slices_solve(*ctx, op, sl, prm, use_graph != 0);
    for (int i = 0; i < nslices; ++i) out[i] = handle_of(rs[i]);
  });
}

int smalgo_apply_poly_slices(smalgo_ctx* ctx, const smalgo_mat* A, int nslices, const int* widths,
                             const int* degrees, const double* coeffs, double smin, double smax,
                             int precision, int panel_width, int use_graph, int replays,
                             const double* Q, int64_t ldq, double* Y, int64_t ldy, double* stats) {
  if (!ctx || !A || nslices < 1 || !widths || !degrees || !coeffs || !Q || !Y || replays < 1)
    return SMALGO_ERR_ARG;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, nullptr, nullptr, A->A.n, op)) throw Error(-1, "bad matrix");
    if (ldq < op.n || ldy < op.n) throw Error(-1, "leading dimension smaller than n");
    if (!(smax > smin)) throw Error(-1, "need smin < smax");
    std::vector<int> w(widths, widths + nslices);
    std::vector<std::vector<double>> co(nslices);
    size_t off = 0;
    for (int i = 0; i < nslices; ++i) {
      if (degrees[i] < 0) throw Error(-1, "degrees must be nonnegative");
      co[i].assign(coeffs + off, coeffs + off + degrees[i] + 1);
      off += (size_t)degrees[i] + 1;
    }
    apply_poly_slices_host(*ctx, op, w, co, smin, smax, precision, panel_width, use_graph != 0,
                           replays, Q, (int)ldq, Y, (int)ldy, stats);
  });
}

int smalgo_chebyshev(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                     int64_t n, double emin, double emax, const smalgo_cheb_params* p,
                     const double* Q0, smalgo_result** out) {
  if (!ctx || !out) return SMALGO_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    Op op;
    if (!make_op(A, fn, user, n, op)) throw Error(-1, "exactly one of A or fn must be given (with matching n)");
    smalgo_cheb_params prm;
    if (p)
      prm = *p;
    else
      smalgo_cheb_default_params(&prm);
    if (A && A->symmetric == 0)
      throw Error(-1, "Chebyshev filtered solver requires a real symmetric matrix");
    Result* r = chebyshev_solve(*ctx, op, emin, emax, prm, Q0, fn != nullptr);
    *out = handle_of(r);
  });
}

// ---- on-disk fixtures ------------------------------------------------------
const char* smalgo_io_last_error(void) { return g_io_err.c_str(); }

int smalgo_csr_save(const char* path, int64_t n, const int64_t* rowptr, const int64_t* colidx,
                    const double* val, int index_base, int symmetric) {
  if (!path) return SMALGO_ERR_ARG;
  return io_guarded([&] { csr_save(path, n, rowptr, colidx, val, index_base, symmetric); });
}

int smalgo_csr_info(const char* path, int64_t* n, int64_t* nnz, int* symmetric) {
  if (!path) return SMALGO_ERR_ARG;
  return io_guarded([&] { csr_info(path, n, nnz, symmetric); });
}

int smalgo_csr_load(const char* path, int64_t* rowptr, int64_t* colidx, double* val) {
  if (!path) return SMALGO_ERR_ARG;
  return io_guarded([&] { csr_load(path, rowptr, colidx, val); });
}

int smalgo_csr_upload_file(smalgo_ctx* ctx, const char* path, smalgo_mat** out) {
  if (!ctx || !path || !out) return SMALGO_ERR_ARG;
  *out = nullptr;
  int64_t n = 0, nnz = 0;
  int sym = 0;
  int rc = smalgo_csr_info(path, &n, &nnz, &sym);
  if (rc != SMALGO_OK) {
    ctx->err = g_io_err;
    return rc;
  }
  // the header was checked against the file size; allocation failures still
  // return an error code instead of unwinding through extern "C"
  try {
    std::vector<int64_t> rp(n + 1), ci(nnz);
    std::vector<double> v(nnz);
    rc = smalgo_csr_load(path, rp.data(), ci.data(), v.data());
    if (rc != SMALGO_OK) {
      ctx->err = g_io_err;
      return rc;
    }
    return smalgo_csr_upload(ctx, n, rp.data(), ci.data(), v.data(), 0, out);
  } catch (const std::exception& e) {
    ctx->err = std::string("loading ") + path + ": " + e.what();
    return SMALGO_ERR_INTERNAL;
  }
}

int smalgo_eig_save(const char* path, int64_t count, const double* lam, const double* res) {
  if (!path) return SMALGO_ERR_ARG;
  return io_guarded([&] { eig_save(path, count, lam, res); });
}

int smalgo_eig_load(const char* path, int64_t* count, int* has_res, double* lam, double* res) {
  if (!path) return SMALGO_ERR_ARG;
  return io_guarded([&] { eig_load(path, count, has_res, lam, res); });
}

// ---- results --------------------------------------------------------------
int64_t smalgo_result_count(const smalgo_result* h) {
  const Result* r = res_of(h);
  return r ? (int64_t)r->lam.size() : -1;
}
int64_t smalgo_result_n(const smalgo_result* h) {
  const Result* r = res_of(h);
  return r ? r->n : -1;
}

}  // extern "C"

namespace {
// Pinned staging for large result downloads, one per device, kept for the life
// of the process: two 64 MiB blocks and a copy stream.  Chunk c+1 crosses PCIe
// into one block while host threads copy chunk c out of the other into the
// caller's (pageable) array.
struct Download {
  hipStream_t st = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double* buf[2] = {nullptr, nullptr};
  size_t cap = 0;  // doubles per block
};
std::mutex g_dl_mu;
std::map<int, Download> g_dl;
constexpr size_t kDownloadChunkBytes = size_t(64) << 20;

hipError_t download_setup(Download& d, size_t doubles) {
  hipError_t e = hipSuccess;
  if (!d.st && (e = hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking)) != hipSuccess) return e;
  for (auto& ev : d.ev)
    if (!ev && (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
  if (d.cap >= doubles) return hipSuccess;
  for (auto& b : d.buf) {
    if (b) (void)hipHostFree(b);
    b = nullptr;
  }
  d.cap = 0;
  for (auto& b : d.buf)
    if ((e = hipHostMalloc((void**)&b, doubles * sizeof(double), hipHostMallocDefault)) !=
        hipSuccess)
      return e;
  d.cap = doubles;
  return hipSuccess;
}

// columns [c0, c0+nc) of a column-major n-row block from pinned src into X
void host_copy_cols(const double* src, size_t n, size_t nc, double* X, size_t ldx, size_t c0) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nthr = std::min<size_t>(std::min<size_t>(8, hw), std::max<size_t>(1, nc * n >> 20));
  auto part = [&](size_t t) {
    if (ldx == n) {  // one contiguous range: split by bytes
      const size_t tot = n * nc, lo = tot * t / nthr, hi = tot * (t + 1) / nthr;
      std::memcpy(X + c0 * n + lo, src + lo, (hi - lo) * sizeof(double));
    } else {
      for (size_t c = nc * t / nthr; c < nc * (t + 1) / nthr; ++c)
        std::memcpy(X + (c0 + c) * ldx, src + c * n, n * sizeof(double));
    }
  };
  std::vector<std::thread> th;
  size_t t = 1;
  try {
    for (; t < nthr; ++t) th.emplace_back(part, t);
  } catch (const std::system_error&) {
    // no more threads available: this thread copies the remaining parts
  }
  for (size_t u = t; u < nthr; ++u) part(u);
  part(0);
  for (auto& x : th) x.join();
}

int download_pinned(const Result* r, double* X, int64_t ldx, size_t k) {
  const size_t n = (size_t)r->n;
  const size_t cpc = std::max<size_t>(1, kDownloadChunkBytes / (n * sizeof(double)));
  std::lock_guard<std::mutex> lock(g_dl_mu);
  Download& d = g_dl[r->device];
  if (download_setup(d, cpc * n) != hipSuccess) return SMALGO_ERR_HIP;
  const size_t nch = (k + cpc - 1) / cpc;
  auto issue = [&](size_t c) {
    const size_t c0 = c * cpc, nc = std::min(cpc, k - c0);
    hipError_t e = hipMemcpyAsync(d.buf[c & 1], r->Xdev + c0 * n, nc * n * sizeof(double),
                                  hipMemcpyDeviceToHost, d.st);
    if (e == hipSuccess) e = hipEventRecord(d.ev[c & 1], d.st);
    return e;
  };
  if (issue(0) != hipSuccess) return SMALGO_ERR_HIP;
  for (size_t c = 0; c < nch; ++c) {
    // block (c+1)&1 last held chunk c-1, whose host copy finished in the previous pass
    if (c + 1 < nch && issue(c + 1) != hipSuccess) return SMALGO_ERR_HIP;
    if (hipEventSynchronize(d.ev[c & 1]) != hipSuccess) return SMALGO_ERR_HIP;
    const size_t c0 = c * cpc, nc = std::min(cpc, k - c0);
    host_copy_cols(d.buf[c & 1], n, nc, X, (size_t)ldx, c0);
  }
  return SMALGO_OK;
}
}  // namespace

extern "C" {

int smalgo_result_copy(const smalgo_result* h, double* lam, double* X, int64_t ldx, double* res) {
  const Result* r = res_of(h);
  if (!r) return SMALGO_ERR_ARG;
  const size_t k = r->lam.size();
  if (lam) std::copy(r->lam.begin(), r->lam.end(), lam);
  if (res) std::copy(r->res.begin(), r->res.end(), res);
  if (X && k > 0) {
    if (ldx < r->n) return SMALGO_ERR_ARG;
    if (hipSetDevice(r->device) != hipSuccess) return SMALGO_ERR_HIP;
    const size_t bytes = sizeof(double) * (size_t)r->n * k;
    if (bytes >= 2 * kDownloadChunkBytes) {
      try {
        return download_pinned(r, X, ldx, k);
      } catch (const std::exception&) {
        return SMALGO_ERR_INTERNAL;
      }
    }
    hipError_t e;
    if (ldx == r->n)
      e = hipMemcpy(X, r->Xdev, bytes, hipMemcpyDeviceToHost);
    else
      e = hipMemcpy2D(X, sizeof(double) * ldx, r->Xdev, sizeof(double) * r->n,
                      sizeof(double) * r->n, k, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return SMALGO_ERR_HIP;
  }
  return SMALGO_OK;
}

int smalgo_result_stat(const smalgo_result* h, const char* name, double* value) {
  const Result* r = res_of(h);
  if (!r || !name || !value) return SMALGO_ERR_ARG;
  auto it = r->stats.find(name);
  if (it == r->stats.end()) return SMALGO_ERR_ARG;
  *value = it->second;
  return SMALGO_OK;
}

const char* smalgo_result_progress_line(const smalgo_result* h) {
  const Result* r = res_of(h);
  return r ? r->line.c_str() : "";
}

void smalgo_result_free(smalgo_result* h) {
  Result* r = res_of(h);
  if (!r) return;
  (void)hipSetDevice(r->device);
  delete r;
}

}  // extern "C"
