// Host drivers for the solve-free interior eigensolvers on one MI355X.
//
// polfed           polfed.jl:1128-1158 (driver), :1022-1099 (basis), :892-984 (extraction)
// chebyshev        chebyshev_subspace_iteration.jl:667-757 (sparse), :771-848 (matvec!)
//
// Everything between the caller's arrays and the results stays in HBM: the
// CSR matrix, the block basis V = [Q_1 … Q_m] (row-major n x (mmax·s), block j
// at columns [j·s, j·s + r_j)), the Clenshaw ping-pong buffers, the lifted Ritz
// vectors.  Small dense FP64 steps (Cholesky, triangular inverse, symmetric
// eigensolve) run through rocSOLVER on the same stream; the block-tridiagonal
// T and the rank decisions are assembled on the host from r x r blocks.
#include "smalgo_solver.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <map>
#include <memory>
#include <numeric>

#include "smalgo_coeffs.h"

namespace smalgo {

static inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// Dense (rocSOLVER, column-major FP64)
// ---------------------------------------------------------------------------
static void check_rb(rocblas_status s, const char* what) {
  if (s != rocblas_status_success)
    throw Error(-4, std::string(what) + " failed with rocblas status " + std::to_string((int)s));
}

int Dense::potrf_upper(int b, double* A, int lda) {
  check_rb(rocsolver_dpotrf(cx.blas, rocblas_fill_upper, b, A, lda, info.p), "rocsolver_dpotrf");
  int h = 0;
  SMALGO_HIP(hipMemcpyAsync(&h, info.p, sizeof(int), hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  return h;
}

void Dense::potrf_upper_async(int b, double* A, int lda, int* dinfo) {
  check_rb(rocsolver_dpotrf(cx.blas, rocblas_fill_upper, b, A, lda, dinfo), "rocsolver_dpotrf");
}

void Dense::trtri_upper(int b, double* A, int lda) {
  check_rb(rocsolver_dtrtri(cx.blas, rocblas_fill_upper, rocblas_diagonal_non_unit, b, A, lda,
                            info.p),
           "rocsolver_dtrtri");
}

void Dense::syevd(int n, double* A, int lda, double* w, double* work_e, bool upper) {
  check_rb(rocsolver_dsyevd(cx.blas, rocblas_evect_original,
                            upper ? rocblas_fill_upper : rocblas_fill_lower, n, A, lda, w, work_e,
                            info.p),
           "rocsolver_dsyevd");
  int h = 0;
  SMALGO_HIP(hipMemcpyAsync(&h, info.p, sizeof(int), hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  if (h != 0) throw Error(-4, "rocsolver_dsyevd did not converge (info=" + std::to_string(h) + ")");
}

// ---------------------------------------------------------------------------
// device memory cache (per thread and device)
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kPoolCacheBytes = size_t(24) << 30;
struct PoolCache {
  std::multimap<size_t, void*> free_blocks[64];
  size_t cached[64] = {};
  ~PoolCache() {
    for (int d = 0; d < 64; ++d)
      for (auto& e : free_blocks[d]) {
        if (hipSetDevice(d) == hipSuccess) (void)hipFree(e.second);
      }
  }
};
thread_local PoolCache g_pool;
// requests round up to 2 MiB above 2 MiB (256 B below), so blocks of nearby
// sizes are interchangeable
size_t pool_round(size_t b) {
  const size_t g = b > (size_t(2) << 20) ? (size_t(2) << 20) : 256;
  return (b + g - 1) / g * g;
}
}  // namespace

void* pool_alloc(size_t bytes, int* device) {
  int d = 0;
  SMALGO_HIP(hipGetDevice(&d));
  *device = d;
  const size_t b = pool_round(bytes);
  if (d >= 0 && d < 64) {
    auto& fl = g_pool.free_blocks[d];
    auto it = fl.lower_bound(b);
    if (it != fl.end() && it->first <= b + b / 4) {
      void* p = it->second;
      g_pool.cached[d] -= it->first;
      fl.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess && d >= 0 && d < 64 && !g_pool.free_blocks[d].empty()) {
    (void)hipGetLastError();
    for (auto& f : g_pool.free_blocks[d]) (void)hipFree(f.second);  // give the cache back
    g_pool.free_blocks[d].clear();
    g_pool.cached[d] = 0;
    e = hipMalloc(&p, b);
  }
  SMALGO_HIP(e);
  return p;
}

void pool_free(void* p, size_t bytes, int device) {
  if (!p) return;
  if (device < 0 || device >= 64) {
    (void)hipFree(p);
    return;
  }
  const size_t b = pool_round(bytes);
  auto& fl = g_pool.free_blocks[device];
  fl.emplace(b, p);
  g_pool.cached[device] += b;
  while (g_pool.cached[device] > kPoolCacheBytes && !fl.empty()) {
    auto last = std::prev(fl.end());  // the largest cached block goes back
    g_pool.cached[device] -= last->first;
    (void)hipFree(last->second);
    fl.erase(last);
  }
}

// ---------------------------------------------------------------------------
// phase timer
// ---------------------------------------------------------------------------
size_t PhaseTimer::begin(const std::string& name) {
  hipEvent_t a = nullptr, b = nullptr;
  SMALGO_HIP(hipEventCreate(&a));
  ev.push_back({name, {a, nullptr}});  // owned by the timer from here on
  SMALGO_HIP(hipEventCreate(&b));
  ev.back().second.second = b;
  SMALGO_HIP(hipEventRecord(a, st));
  return ev.size() - 1;
}
void PhaseTimer::end(size_t h) { SMALGO_HIP(hipEventRecord(ev[h].second.second, st)); }
std::map<std::string, double> PhaseTimer::resolve() {
  std::map<std::string, double> out;
  SMALGO_HIP(hipStreamSynchronize(st));
  for (auto& e : ev) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess)
      out[e.first] += ms * 1e-3;
  }
  return out;
}
PhaseTimer::~PhaseTimer() {
  for (auto& e : ev) {
    if (e.second.first) (void)hipEventDestroy(e.second.first);
    if (e.second.second) (void)hipEventDestroy(e.second.second);
  }
}

// ---------------------------------------------------------------------------
// filter tile choice: widest aligned vector, then enough lanes to cover b
// ---------------------------------------------------------------------------
FilterTile choose_tile(size_t elem, int b, std::initializer_list<const void*> ptrs,
                       std::initializer_list<int> lds) {
  int vec = elem == 4 ? 4 : 2;
  auto ok = [&](int v) {
    for (const void* p : ptrs)
      if (p && reinterpret_cast<uintptr_t>(p) % (elem * v)) return false;
    for (int ld : lds)
      if (ld % v) return false;
    return true;
  };
  while (vec > 1 && (!ok(vec) || b < vec)) vec /= 2;
  // one column chunk of at most 512 bytes per row: keeps the rows that a
  // row's CSR neighbours gather inside the per-XCD L2 for wide blocks
  const int chunk_elems = std::max(vec, (int)(512 / elem));
  int need = (std::min(b, chunk_elems) + vec - 1) / vec;
  int lpr = 4;
  while (lpr < need && lpr < 64) lpr *= 2;
  return FilterTile{vec, lpr, 1};
}

// ---------------------------------------------------------------------------
// operator application
// ---------------------------------------------------------------------------
void Staging::ensure(size_t count, int chunks) {
  if (!up) SMALGO_HIP(hipStreamCreateWithFlags(&up, hipStreamNonBlocking));
  if (!done) SMALGO_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
  while ((int)ev.size() < chunks) {
    hipEvent_t e = nullptr;
    SMALGO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev.push_back(e);
  }
  if (count <= cap) return;
  if (x) (void)hipHostFree(x);
  if (y) (void)hipHostFree(y);
  x = y = nullptr;
  cap = 0;
  SMALGO_HIP(hipHostMalloc((void**)&x, sizeof(double) * count, hipHostMallocDefault));
  SMALGO_HIP(hipHostMalloc((void**)&y, sizeof(double) * count, hipHostMallocDefault));
  cap = count;
}

Staging::~Staging() {
  if (x) (void)hipHostFree(x);
  if (y) (void)hipHostFree(y);
  for (auto e : ev) (void)hipEventDestroy(e);
  if (done) (void)hipEventDestroy(done);
  if (up) (void)hipStreamDestroy(up);
}

// Y = A X through the host callback.  The device block goes column-major,
// then to pinned host memory in column chunks; the callback runs on chunk c
// while later chunks are still in flight device->host, and each result chunk
// goes back host->device on a second stream while the callback works on the
// next one.  matblock!(Y, X) acts column by column, so calling it per chunk
// is the same operator application.
static void callback_apply(Ctx& cx, const Op& op, int b, const double* X, int ldx, double* Y,
                           int ldy) {
  const int n = op.n;
  if (b == 0 || n == 0) return;
  // about four chunks, none under 4 MB
  const int min_cols = std::max(1, (int)((4u << 20) / (sizeof(double) * (size_t)n)));
  const int cw = std::max(min_cols, (b + 3) / 4);
  const int nch = (b + cw - 1) / cw;
  Staging& sg = cx.staging;
  sg.ensure((size_t)n * b, nch);
  DBuf<double> dx((size_t)n * b), dy((size_t)n * b);
  rm_to_cm<double>(cx.stream, n, b, X, ldx, nullptr, dx.p, n);
  for (int c = 0; c < nch; ++c) {
    const size_t off = (size_t)c * cw * n, cnt = (size_t)std::min(cw, b - c * cw) * n;
    SMALGO_HIP(hipMemcpyAsync(sg.x + off, dx.p + off, sizeof(double) * cnt,
                              hipMemcpyDeviceToHost, cx.stream));
    SMALGO_HIP(hipEventRecord(sg.ev[c], cx.stream));
  }
  for (int c = 0; c < nch; ++c) {
    const int cb = std::min(cw, b - c * cw);
    const size_t off = (size_t)c * cw * n, cnt = (size_t)cb * n;
    SMALGO_HIP(hipEventSynchronize(sg.ev[c]));
    const int rc = op.fn(op.user, n, cb, sg.x + off, n, sg.y + off, n);
    if (rc != 0) {
      (void)hipStreamSynchronize(sg.up);
      throw Error(-5, "matblock callback returned " + std::to_string(rc));
    }
    SMALGO_HIP(hipMemcpyAsync(dy.p + off, sg.y + off, sizeof(double) * cnt,
                              hipMemcpyHostToDevice, sg.up));
  }
  SMALGO_HIP(hipEventRecord(sg.done, sg.up));
  SMALGO_HIP(hipStreamWaitEvent(cx.stream, sg.done, 0));
  cm_to_rm<double, double>(cx.stream, n, b, dy.p, n, Y, ldy);
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}

template <typename T>
void apply_A(Ctx& cx, const Op& op, int b, const T* X, int ldx, T* Y, int ldy) {
  if (op.is_callback()) {
    if constexpr (std::is_same<T, double>::value) {
      callback_apply(cx, op, b, X, ldx, Y, ldy);
      return;
    } else {
      throw Error(-3, "matrix-free operators run in FP64 only");
    }
  }
  const FilterTile t = choose_tile(sizeof(T), b, {X, Y}, {ldx, ldy});
  launch_clenshaw_step<T>(cx.stream, *op.A, b, nullptr, 0, T(0), X, ldx, T(1), T(0), nullptr, 0,
                          T(0), Y, ldy, t);
}
template void apply_A<double>(Ctx&, const Op&, int, const double*, int, double*, int);
template void apply_A<float>(Ctx&, const Op&, int, const float*, int, float*, int);

// ---------------------------------------------------------------------------
// Split filter steps (FilterTile::split = 2): a step of width b runs as two
// launches over the column halves [0, h) and [h, b), the first on the
// context's stream and the second on its side stream.  The halves share no
// column, so their recurrences are independent; run concurrently, each chain
// fills the other's launch ramp and drain.  h is b/2 rounded to the nearest
// whole 128-byte line, and the split is made only when every block pointer and
// leading dimension keeps the halves on separate cache lines.
// ---------------------------------------------------------------------------
static int split_half(int b, size_t elem) {
  const int g = (int)(128 / elem);
  const int h = (b + g) / (2 * g) * g;  // the whole-line boundary nearest b/2
  return (h >= g && b - h >= g) ? h : 0;
}

static bool split_ok(int b, size_t elem, std::initializer_list<const void*> ptrs,
                     std::initializer_list<int> lds) {
  if (split_half(b, elem) == 0) return false;
  for (const void* q : ptrs)
    if (q && reinterpret_cast<uintptr_t>(q) % 128) return false;
  for (int ld : lds)
    if (ld && ((size_t)ld * elem) % 128) return false;
  return true;
}

static void side_stream(Ctx& cx) {
  if (cx.side) return;
  SMALGO_HIP(hipEventCreateWithFlags(&cx.fork_ev, hipEventDisableTiming));
  SMALGO_HIP(hipEventCreateWithFlags(&cx.join_ev, hipEventDisableTiming));
  SMALGO_HIP(hipStreamCreateWithFlags(&cx.side, hipStreamNonBlocking));
}

// the side stream waits for all work issued so far on the main stream
static void split_fork(Ctx& cx) {
  side_stream(cx);
  SMALGO_HIP(hipEventRecord(cx.fork_ev, cx.stream));
  SMALGO_HIP(hipStreamWaitEvent(cx.side, cx.fork_ev, 0));
}

// the main stream waits for all work issued so far on the side stream
static void split_join(Ctx& cx) {
  SMALGO_HIP(hipEventRecord(cx.join_ev, cx.side));
  SMALGO_HIP(hipStreamWaitEvent(cx.stream, cx.join_ev, 0));
}

// Fork on entry when `on`; join on exit, also when a launch throws.  Split
// launches happen only inside such a scope (an inner scope adds nothing).
struct SplitScope {
  Ctx& cx;
  bool on;
  SplitScope(Ctx& c, bool o) : cx(c), on(o && !c.split_active) {
    if (on) {
      split_fork(cx);
      cx.split_active = true;
      cx.split_h = -1;
    }
  }
  ~SplitScope() {
    if (on) {
      cx.split_active = false;
      (void)hipEventRecord(cx.join_ev, cx.side);
      (void)hipStreamWaitEvent(cx.stream, cx.join_ev, 0);
    }
  }
};

// Column split of the next launch (0: one launch on the main stream).  When it
// differs from the last launch's, the streams meet first (join, then fork):
// the buffers the recurrence reuses (ping-pong blocks, panels) change hands
// between the streams at the new column boundary.
static int split_begin(Ctx& cx, int h) {
  if (!cx.split_active) return 0;
  if (cx.split_h >= 0 && h != cx.split_h) {
    split_join(cx);
    split_fork(cx);
  }
  cx.split_h = h;
  return h;
}

// One fused step with the tile's launch shape: one launch, or (split = 2)
// one per column half on the two streams, inside a SplitScope of the caller.
// A split tile outside a scope, or on operands that do not allow the split,
// runs as one launch.
template <typename T>
static void launch_step(Ctx& cx, const DevCSR& A, int b, const T* Q, int ldq, T cq, const T* X1,
                        int ld1, T a1, T b1, const T* X2, int ld2, T s2, T* out, int ldo,
                        FilterTile t) {
  const int h = split_begin(
      cx, t.split == 2 && split_ok(b, sizeof(T), {Q, X1, X2, out}, {ldq, ld1, ld2, ldo})
              ? split_half(b, sizeof(T))
              : 0);
  if (h == 0) {
    launch_clenshaw_step<T>(cx.stream, A, b, Q, ldq, cq, X1, ld1, a1, b1, X2, ld2, s2, out, ldo, t);
    return;
  }
  auto off = [h](const T* q) { return q ? q + h : nullptr; };
  ++cx.split_steps;
  launch_clenshaw_step<T>(cx.stream, A, h, Q, ldq, cq, X1, ld1, a1, b1, X2, ld2, s2, out, ldo, t);
  launch_clenshaw_step<T>(cx.side, A, b - h, off(Q), ldq, cq, off(X1), ld1, a1, b1, off(X2), ld2,
                          s2, out + h, ldo, t);
}

// the same for per-column Q coefficients (row cq of b values)
template <typename T>
static void launch_step_cols(Ctx& cx, const DevCSR& A, int b, const T* Q, int ldq, const T* cq,
                             const T* X1, int ld1, T a1, T b1, const T* X2, int ld2, T s2, T* out,
                             int ldo, FilterTile t) {
  const int h = split_begin(
      cx, t.split == 2 && split_ok(b, sizeof(T), {Q, X1, X2, out}, {ldq, ld1, ld2, ldo})
              ? split_half(b, sizeof(T))
              : 0);
  if (h == 0) {
    launch_clenshaw_step_cols<T>(cx.stream, A, b, Q, ldq, cq, X1, ld1, a1, b1, X2, ld2, s2, out,
                                 ldo, t);
    return;
  }
  auto off = [h](const T* q) { return q ? q + h : nullptr; };
  ++cx.split_steps;
  launch_clenshaw_step_cols<T>(cx.stream, A, h, Q, ldq, cq, X1, ld1, a1, b1, X2, ld2, s2, out,
                               ldo, t);
  launch_clenshaw_step_cols<T>(cx.side, A, b - h, off(Q), ldq, cq + h, off(X1), ld1, a1, b1,
                               off(X2), ld2, s2, out + h, ldo, t);
}

// One Clenshaw step out = cq*Q + a1*(A X1) + b1*X1 - s2*X2 with either
// operator kind.  X1/X2 are given as (pointer, implicit multiple of Q).
template <typename T>
struct FBlk {
  const T* p;
  double qs;
};

template <typename T>
static void filter_step(Ctx& cx, const Op& op, int b, const T* Q, int ldq, double cq, double a1,
                        double b1, FBlk<T> X1, FBlk<T> X2, int ldb, T* out, int ldo, double* pbuf,
                        const FilterTile* tile) {
  const T* x1p = X1.p;
  int ld1 = ldb;
  double a1e = a1, b1e = b1;
  if (!x1p) {
    if (X1.qs == 0.0) {
      a1e = b1e = 0.0;
    } else {
      x1p = Q;
      ld1 = ldq;
      a1e = a1 * X1.qs;
      b1e = b1 * X1.qs;
    }
  }
  const T* x2p = X2.p;
  double cqe = cq, s2e = 1.0;
  if (!x2p) {
    cqe = cq - X2.qs;
    s2e = 0.0;
  }
  if (op.is_callback()) {
    if constexpr (std::is_same<T, double>::value) {
      const double* P = nullptr;
      if (x1p) {
        callback_apply(cx, op, b, x1p, ld1, pbuf, ldb);
        P = pbuf;
      } else {
        a1e = 0.0;
      }
      launch_clenshaw_step_pre(cx.stream, op.n, b, Q, ldq, cqe, P ? P : nullptr, ldb, x1p, ld1,
                               P ? a1e : 0.0, b1e, x2p, ldb, s2e, out, ldo);
      return;
    } else {
      throw Error(-3, "matrix-free operators run in FP64 only");
    }
  }
  FilterTile t = tile ? *tile : choose_tile(sizeof(T), b, {Q, x1p, x2p, out}, {ldq, ld1, ldb, ldo});
  launch_step<T>(cx, *op.A, b, Q, ldq, (T)cqe, x1p, ld1, (T)a1e, (T)b1e, x2p, ldb, (T)s2e, out, ldo,
                 t);
}

// Autotuner: on the first filter of a (precision, width, matrix, alignment)
// combination, time a few fused steps of every admissible (vec, lanes per
// row) tile on the real operands and keep the fastest.  Rows per wavefront is
// 64/lpr; the column chunk per row is lpr*vec elements.
template <typename T>
static int tile_alignment_ok(int v, const T* Q, int ldq, const T* buf0, const T* buf1, int ldb,
                             const T* Y, int ldy) {
  for (const void* p : {(const void*)Q, (const void*)buf0, (const void*)buf1, (const void*)Y})
    if (reinterpret_cast<uintptr_t>(p) % (sizeof(T) * v)) return 0;
  for (int ld : {ldq, ldb, ldy})
    if (ld % v) return 0;
  return 1;
}

// A 64-bit key of the autotuner's cache from its parts (splitmix64 rounds).
static long long tune_key_mix(std::initializer_list<long long> parts) {
  uint64_t h = 0x9E3779B97F4A7C15ull;
  for (long long v : parts) {
    uint64_t z = h ^ (uint64_t)v;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    h = z ^ (z >> 31);
  }
  return (long long)h;
}

// key of the autotuner's cache: precision, width, the matrix's structure
// fingerprint and operand alignment
template <typename T>
static long long tile_key(const Op& op, int b, const T* Q, int ldq, const T* buf0, const T* buf1,
                          int ldb, const T* Y, int ldy) {
  auto al = [&](int v) { return tile_alignment_ok<T>(v, Q, ldq, buf0, buf1, ldb, Y, ldy); };
  const bool sp = split_ok(b, sizeof(T), {Q, buf0, buf1, Y}, {ldq, ldb, ldy});
  return tune_key_mix({(long long)sizeof(T), (long long)b, op.mat_id, (long long)al(8),
                       (long long)al(4), (long long)al(2), (long long)sp, 0});
}

// Every tile the autotuners consider at block width b: lane vectors of 1, 2,
// 4 or 8 elements that vec_ok admits (operand alignment, width), column
// chunks lpr·vec of at most twice the width and at most 2 KB, 1, 2 or 4 rows
// per team, and the CSR forms: per lane, staged in LDS, and (when `prefetch`)
// staged with own-row prefetch for 1 or 2 rows per team.  The prefetch form
// has only won where a step's working set exceeds the Infinity Cache, so the
// callers offer it there only.  With `split`, the split tiles follow: the
// same shapes at the half width (split_half), with 1 or 2 rows per team and
// without the prefetch form.
template <typename VecOk>
static std::vector<FilterTile> tile_candidates(int b, size_t elem, bool prefetch, bool split,
                                               VecOk&& vec_ok) {
  std::vector<FilterTile> cand;
  auto add = [&](int w0, int sp) {
    for (int v : {1, 2, 4, 8}) {
      if (!vec_ok(v) || w0 < v) continue;
      for (int l : {4, 8, 16, 32, 64}) {
        const int w = l * v;  // chunk width in elements
        if (w > 2 * std::max(w0, 4) || w * (int)elem > 2048) continue;
        for (int rt : {1, 2, 4})
          for (int lds : {0, 1, 2}) {
            if (sp == 2 && (rt > 2 || lds == 2)) continue;
            if (lds < 2 || (prefetch && rt <= 2)) cand.push_back(FilterTile{v, l, rt, lds, sp});
          }
      }
    }
  };
  add(b, 1);
  static const bool no_split = [] {
    const char* e = std::getenv("SMALGO_NO_SPLIT");
    return e && e[0] == '1';
  }();
  const int h = split && !no_split ? split_half(b, elem) : 0;
  if (h > 0) add(h, 2);
  return cand;
}

// Times step(tile) for every candidate in `sweeps` sweeps of `reps` steps, each
// after one warm step; a tile's time (ms per step, into tmin) is its fastest
// sweep, so clock ramps and one-off stalls do not pick the tile.  Sweeps after
// the first re-time only the candidates within 25% of the best time so far,
// over twice the steps; a run-off over four times the steps decides among the
// tiles within 3% of the best.
// Returns the index of the fastest.  SMALGO_TUNE_LOG=1 prints every time and
// the pick.
template <typename StepFn>
static size_t time_tiles(Ctx& cx, const std::vector<FilterTile>& cand, int sweeps, int reps,
                         StepFn&& step, std::vector<float>& tmin, const std::string& what) {
  Event e0, e1, e1s;
  // ms per step of candidate i over nrep steps after one warm step.  A split
  // tile's steps run on both streams inside one fork/join pair, as in the
  // recurrence; their time runs from e0 on the main stream to the later of
  // the two streams' end events, so the fork and the join (paid once per
  // filter, not per step) stay outside it.
  auto time_one = [&](size_t i, int nrep) {
    const bool sp = cand[i].split == 2;
    bool two = false;
    {
      SplitScope scope(cx, sp);
      step(cand[i]);  // warm
      SMALGO_HIP(hipEventRecord(e0, cx.stream));
      for (int r = 0; r < nrep; ++r) step(cand[i]);
      SMALGO_HIP(hipEventRecord(e1, cx.stream));
      two = sp && cx.split_active;
      if (two) SMALGO_HIP(hipEventRecord(e1s, cx.side));
    }
    SMALGO_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SMALGO_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (two) {
      float ms2 = 0.f;
      SMALGO_HIP(hipEventSynchronize(e1s));
      SMALGO_HIP(hipEventElapsedTime(&ms2, e0, e1s));
      ms = std::max(ms, ms2);
    }
    return ms / (float)nrep;
  };
  tmin.assign(cand.size(), 1e30f);
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    // later sweeps re-time only the tiles within 25% of the best so far, with
    // twice the steps
    const float cut = 1.25f * *std::min_element(tmin.begin(), tmin.end());
    const int nrep = sweep > 0 ? 2 * reps : reps;
    for (size_t i = 0; i < cand.size(); ++i) {
      if (sweep > 0 && tmin[i] > cut) continue;
      tmin[i] = std::min(tmin[i], time_one(i, nrep));
    }
  }
  // Run-off: the (at most four) tiles within 3% of the best are timed once
  // more over four times the steps, and the pick is the fastest of them there.
  std::vector<size_t> close;
  if (sweeps > 1) {
    const float lim = 1.03f * *std::min_element(tmin.begin(), tmin.end());
    for (size_t i = 0; i < cand.size(); ++i)
      if (tmin[i] <= lim) close.push_back(i);
    std::stable_sort(close.begin(), close.end(), [&](size_t a, size_t b) { return tmin[a] < tmin[b]; });
    if (close.size() > 4) close.resize(4);
    if (close.size() < 2) close.clear();
    for (size_t i : close) tmin[i] = time_one(i, 4 * reps);
  }
  size_t best = close.empty() ? 0 : close[0];
  if (close.empty()) {
    for (size_t i = 1; i < cand.size(); ++i)
      if (tmin[i] < tmin[best]) best = i;
  } else {
    for (size_t i : close)
      if (tmin[i] < tmin[best]) best = i;
  }
  const char* lg = std::getenv("SMALGO_TUNE_LOG");
  if (lg && lg[0] == '1') {
    for (size_t i = 0; i < cand.size(); ++i)
      std::fprintf(stderr, "[smalgo tune] %s vec=%d lpr=%d rt=%d lds=%d split=%d: %.2f us\n",
                   what.c_str(), cand[i].vec, cand[i].lpr, cand[i].rt, cand[i].lds, cand[i].split,
                   tmin[i] * 1e3f);
    std::fprintf(stderr, "[smalgo tune] %s picked vec=%d lpr=%d rt=%d lds=%d split=%d\n",
                 what.c_str(), cand[best].vec, cand[best].lpr, cand[best].rt, cand[best].lds,
                 cand[best].split);
  }
  return best;
}

template <typename T>
static FilterTile autotune_tile(Ctx& cx, const Op& op, int b, const T* Q, int ldq, T* buf0, T* buf1,
                                int ldb, T* Y, int ldy) {
  const FilterTile def = choose_tile(sizeof(T), b, {Q, buf0, buf1, Y}, {ldq, ldb, ldy});
  if (!cx.autotune) return def;
  auto al = [&](int v) { return tile_alignment_ok<T>(v, Q, ldq, buf0, buf1, ldb, Y, ldy); };
  const long long key = tile_key<T>(op, b, Q, ldq, buf0, buf1, ldb, Y, ldy);
  auto it = cx.tuned.find(key);
  if (it != cx.tuned.end()) return it->second;
  // working set of one step beyond the 240 MB Infinity-Cache budget?
  const bool beyond = 3.0 * op.n * b * sizeof(T) + (double)op.A->nnz * (sizeof(T) + 4) > 240e6;
  const bool sp = split_ok(b, sizeof(T), {Q, buf0, buf1, Y}, {ldq, ldb, ldy});
  const std::vector<FilterTile> cand =
      tile_candidates(b, sizeof(T), beyond, sp, [&](int v) { return al(v); });
  if (cand.empty()) return def;
  // The steps timed are the ones the recurrence runs from its third step on:
  // Q, X1 = buf0 and X2 = buf1 distinct, the output written in place over X2
  // (out = cq Q + a1 A X1 + b1 X1 - X2).  The two recurrence steps without X2
  // first give buf0 and buf1 defined values; the in-place step then maps buf1
  // to c - buf1, so the values stay bounded.  Two sweeps of 6 steps.
  launch_clenshaw_step<T>(cx.stream, *op.A, b, Q, ldq, T(0.5), Q, ldq, T(0.1), T(0.1), nullptr, ldb,
                          T(0), buf0, ldb, def);
  launch_clenshaw_step<T>(cx.stream, *op.A, b, Q, ldq, T(0.5), buf0, ldb, T(0.1), T(0.1), nullptr,
                          ldb, T(0), buf1, ldb, def);
  std::vector<float> tmin;
  const size_t ib = time_tiles(
      cx, cand, 2, 6,
      [&](const FilterTile& t) {
        launch_step<T>(cx, *op.A, b, Q, ldq, T(0.5), buf0, ldb, T(0.1), T(0.1), buf1, ldb, T(1),
                       buf1, ldb, t);
      },
      tmin,
      std::string(sizeof(T) == 4 ? "fp32" : "fp64") + " n=" + std::to_string(op.n) +
          " b=" + std::to_string(b));
  const FilterTile best = cand[ib];
  cx.tuned[key] = best;
  return best;
}

// The recurrence itself, written once over a "step" callback that launches
// out = cq*Q + a1*(A X1) + b1*X1 - s2*X2 with X1/X2 given as (pointer,
// implicit multiple of Q).
template <typename T, typename StepFn>
static void clenshaw_recurrence(const double* co, int K, double smin, double smax, T* buf0,
                                T* buf1, T* Y, StepFn&& step) {
  // Block Clenshaw exactly as apply_poly_block! (polfed.jl:464-501):
  //   B_{K+1} = B_{K+2} = 0
  //   for k = K+1 .. 2:  TMP = c[k-1] Q + α A B1 + β B1 - B2;  (B1,B2) <- (TMP,B1)
  //   Y = c0 Q + invd A B1 - c invd B1 - B2
  // with invd = 2/(smax-smin), α = 2 invd, β = -2 c invd.  The first two
  // steps have B1 = c_K Q and B2 ∈ {0, c_K Q}; they are folded into the
  // Q term instead of materialised, so the zero SpMM of the reference's first
  // step is never launched.  From the third step on TMP overwrites B2 in place.
  const double c = 0.5 * (smin + smax);
  const double invd = 2.0 / (smax - smin);
  const double al = 2.0 * invd, be = -2.0 * c * invd;
  FBlk<T> B1{nullptr, 0.0}, B2{nullptr, 0.0};
  T* bufs[2] = {buf0, buf1};
  for (int k = K + 1; k >= 2; --k) {
    const double ck = co[k - 1];
    if (k == K + 1) {
      B1 = {nullptr, ck};
      B2 = {nullptr, 0.0};
      continue;
    }
    T* out = B2.p ? const_cast<T*>(B2.p) : (B1.p == bufs[0] ? bufs[1] : bufs[0]);
    step(ck, al, be, B1, B2, out);
    B2 = B1;
    B1 = {out, 0.0};
  }
  step(co[0], invd, -c * invd, B1, B2, Y);
}

// Cache-aware panel width: the widest of 512 / 256 / 128-byte row segments
// whose recurrence working set (Q, B1, B2 of one panel plus the matrix) fits
// in 240 MB of the 256 MB Infinity Cache.  Measured at n = 262,144: 256-byte
// panels (201 MB) run at 5.7 TB/s, 512-byte panels (403 MB) at 3.8 TB/s.
template <typename T>
static int cache_panel_width(const Op& op) {
  // SMALGO_PANEL_BYTES=128|256|512 fixes the row-segment width (tuning override)
  static const int forced = [] {
    const char* e = std::getenv("SMALGO_PANEL_BYTES");
    const int v = e ? std::atoi(e) : 0;
    return (v == 128 || v == 256 || v == 512) ? v : 0;
  }();
  if (forced) return forced / (int)sizeof(T);
  const double mat = op.A ? (double)op.A->nnz * (sizeof(T) + 4) : 0.0;
  for (int bytes = 512; bytes > 128; bytes /= 2)
    if (3.0 * op.n * bytes + mat <= 240e6) return bytes / (int)sizeof(T);
  return 128 / (int)sizeof(T);
}

// Tuning overrides (environment): SMALGO_FILTER_TILE="vec,lpr,rt[,lds[,split]]"
// fixes the tile of row-major filters, SMALGO_PANEL_TILE likewise for panel
// filters, SMALGO_NO_PANELS=1 keeps wide blocks row-major.
static bool env_tile(const char* name, FilterTile* t) {
  const char* v = std::getenv(name);
  if (!v) return false;
  int a = 0, b = 0, c = 1, d = 0, e = 1;
  if (std::sscanf(v, "%d,%d,%d,%d,%d", &a, &b, &c, &d, &e) < 2) return false;
  *t = FilterTile{a, b, c, d, e == 2 ? 2 : 1};
  return true;
}

template <typename T>
void apply_filter(Ctx& cx, const Op& op, int b, const T* Q, int ldq, const std::vector<double>& co,
                  double smin, double smax, T* Y, int ldy, T* buf0, T* buf1, int ldb,
                  double* pbuf) {
  const char* nop = std::getenv("SMALGO_NO_PANELS");
  const bool panels_ok = !(nop && nop[0] == '1');
  // Panels only pay when one launch over the whole block would not stay
  // cache-resident; small problems run the whole width in one launch per step.
  const bool fits = !op.is_callback() &&
                    3.0 * op.n * b * sizeof(T) + (double)op.A->nnz * (sizeof(T) + 4) <= 240e6;
  if (!op.is_callback() && panels_ok && !fits && b > cache_panel_width<T>(op)) {
    // Wide blocks: repack into dense n x PW slabs (PW from the cache-aware
    // rule) so that every gather and stream of one launch chunk walks
    // contiguous rows, then unpack the result.
    const int pw = cache_panel_width<T>(op);
    const int npan = (b + pw - 1) / pw;
    const int n = op.n;
    const size_t ps = (size_t)n * pw;
    DBuf<T> Qp(ps * npan), P0(ps), P1(ps), Yp(ps * npan);
    to_panels<T>(cx.stream, n, b, pw, Q, ldq, Qp.p);
    // Panels run one after another through the whole recurrence (panel-outer,
    // step-inner): one panel's Q, B1, B2 (3 x n x PW) stay resident in the
    // 256 MB Infinity Cache across its K steps, which interleaving the panels
    // step by step would evict.  Each launch has the narrow-block shape the
    // autotuner measured.
    FilterTile tile{};
    if (!env_tile("SMALGO_PANEL_TILE", &tile))
      tile = autotune_tile<T>(cx, op, pw, Qp.p, pw, P0.p, P1.p, pw, Yp.p, pw);
    cx.last_tile = tile;
    cx.last_launches_per_step = npan;
    {
      SplitScope split(cx, tile.split == 2);  // joined before the unpacking below
      for (int pn = 0; pn < npan; ++pn) {
        const int bp = std::min(pw, b - pn * pw);
        const T* Qn = Qp.p + (size_t)pn * ps;
        auto step = [&](double cq, double a1, double b1, FBlk<T> X1, FBlk<T> X2, T* out) {
          const T* x1p = X1.p;
          double a1e = a1, b1e = b1;
          if (!x1p) {
            if (X1.qs == 0.0) {
              a1e = b1e = 0.0;
            } else {
              x1p = Qn;
              a1e = a1 * X1.qs;
              b1e = b1 * X1.qs;
            }
          }
          const T* x2p = X2.p;
          double cqe = cq, s2e = 1.0;
          if (!x2p) {
            cqe = cq - X2.qs;
            s2e = 0.0;
          }
          launch_step<T>(cx, *op.A, bp, Qn, pw, (T)cqe, x1p, pw, (T)a1e, (T)b1e, x2p, pw, (T)s2e,
                         out, pw, tile);
        };
        clenshaw_recurrence<T>(co.data(), (int)co.size() - 1, smin, smax, P0.p, P1.p,
                               Yp.p + (size_t)pn * ps, step);
      }
    }
    from_panels<T>(cx.stream, n, b, pw, Yp.p, Y, ldy);
    return;
  }
  FilterTile tile{};
  const FilterTile* tp = nullptr;
  if (!op.is_callback()) {
    if (!env_tile("SMALGO_FILTER_TILE", &tile))
      tile = autotune_tile<T>(cx, op, b, Q, ldq, buf0, buf1, ldb, Y, ldy);
    tp = &tile;
    cx.last_tile = tile;
    cx.last_launches_per_step = 1;
  }
  auto step = [&](double cq, double a1, double b1, FBlk<T> X1, FBlk<T> X2, T* out) {
    const int ldo = (out == Y) ? ldy : ldb;
    filter_step<T>(cx, op, b, Q, ldq, cq, a1, b1, X1, X2, ldb, out, ldo, pbuf, tp);
  };
  SplitScope split(cx, tp && tile.split == 2);
  clenshaw_recurrence<T>(co.data(), (int)co.size() - 1, smin, smax, buf0, buf1, Y, step);
}
template void apply_filter<double>(Ctx&, const Op&, int, const double*, int,
                                   const std::vector<double>&, double, double, double*, int,
                                   double*, double*, int, double*);
template void apply_filter<float>(Ctx&, const Op&, int, const float*, int,
                                  const std::vector<double>&, double, double, float*, int, float*,
                                  float*, int, double*);

// Y = p(Â) Q for a CSR operator with no allocation of any kind: the
// ping-pong pair of the recurrence is (Y, scratch), ordered so that the last
// step's B_2 is Y itself (the kernel allows out to alias X2 element by
// element) and never its gathered B_1; Y and scratch share the leading
// dimension ldy.  The tile is the autotuner's cached choice for this shape if
// one exists, else the default tile: nothing is timed.  Row-major launches
// only (no panel repacking).
void apply_filter_noalloc(Ctx& cx, const Op& op, int b, const double* Q, int ldq, const double* co,
                          int K, double smin, double smax, double* Y, int ldy, double* scratch) {
  if (op.is_callback()) throw Error(-1, "the no-allocation filter needs a sparse matrix");
  // which buffer of the pair is the final B_2 (and which the final B_1)?
  double* const L0 = reinterpret_cast<double*>(uintptr_t(16));
  double* const L1 = reinterpret_cast<double*>(uintptr_t(32));
  double* const LY = reinterpret_cast<double*>(uintptr_t(48));
  const double* last1 = nullptr;
  const double* last2 = nullptr;
  clenshaw_recurrence<double>(co, K, smin, smax, L0, L1, LY,
                              [&](double, double, double, FBlk<double> X1, FBlk<double> X2,
                                  double* out) {
                                if (out == LY) {
                                  last1 = X1.p;
                                  last2 = X2.p;
                                }
                              });
  const bool y_first = last2 ? last2 == L0 : last1 != L0;
  double* buf0 = y_first ? Y : scratch;
  double* buf1 = y_first ? scratch : Y;
  FilterTile tile = choose_tile(sizeof(double), b, {Q, buf0, buf1, Y}, {ldq, ldy, ldy});
  auto it = cx.tuned.find(tile_key<double>(op, b, Q, ldq, buf0, buf1, ldy, Y, ldy));
  if (it != cx.tuned.end()) tile = it->second;
  tile.split = 1;  // one stream: a split tile's half shape runs over the whole width
  cx.last_tile = tile;
  cx.last_launches_per_step = 1;
  auto step = [&](double cq, double a1, double b1, FBlk<double> X1, FBlk<double> X2, double* out) {
    filter_step<double>(cx, op, b, Q, ldq, cq, a1, b1, X1, X2, ldy, out, ldy, nullptr, &tile);
  };
  clenshaw_recurrence<double>(co, K, smin, smax, buf0, buf1, Y, step);
}

// ---------------------------------------------------------------------------
// CholQR2 with a shifted-CholQR3 fallback (Fukaya et al., SIAM J. Sci.
// Comput. 42 (2020)): R from the FP64 Cholesky of the Gram matrix, Q = W R^{-1}.
// Replaces Householder qr! in orth_block_polfed!/block_qr_factor
// (polfed.jl:561-594, :644-685) and thin_orth! (chebyshev_subspace_iteration.jl:223-231).
// Returns the count rank r = #{|R_ii| > rtol·max|R_ii|}; Q gets all b columns,
// the caller keeps the first r (the reference's rule, polfed.jl:679-684).
// ---------------------------------------------------------------------------
struct OrthWork {
  DBuf<double> G, Rinv, gemm_work, cn, red, R1, R2, status;
  DBuf<int> info2;
  std::vector<double> hG, hstatus;
};

// R (upper, zeros below the diagonal) into ow.R2 and R⁻¹ into ow.Rinv for
// the b x b Gram matrix in ow.G, which stays unchanged on the one-workgroup
// path (b <= chol_inv_max_b()) and is overwritten by rocSOLVER's potrf
// otherwise.  Returns potrf's info; R2 and Rinv are valid only for info 0.
static int chol_inv_sync(Ctx& cx, Dense& dn, OrthWork& ow, int b) {
  ow.R2.ensure((size_t)b * b);
  ow.Rinv.ensure((size_t)b * b);
  ow.info2.ensure(2);
  if (b <= chol_inv_max_b()) {
    chol_inv_upper(cx.stream, b, ow.G.p, b, ow.R2.p, b, ow.Rinv.p, b, ow.info2.p);
    int h = 0;
    SMALGO_HIP(hipMemcpyAsync(&h, ow.info2.p, sizeof(int), hipMemcpyDeviceToHost, cx.stream));
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
    return h;
  }
  const int info = dn.potrf_upper(b, ow.G.p, b);
  if (info != 0) return info;
  zero_lower_cm(cx.stream, b, ow.G.p, b);
  SMALGO_HIP(hipMemcpyAsync(ow.R2.p, ow.G.p, sizeof(double) * b * b, hipMemcpyDeviceToDevice,
                            cx.stream));
  SMALGO_HIP(hipMemcpyAsync(ow.Rinv.p, ow.G.p, sizeof(double) * b * b, hipMemcpyDeviceToDevice,
                            cx.stream));
  dn.trtri_upper(b, ow.Rinv.p, b);
  return 0;
}

template <typename T>
static int cholqr(Ctx& cx, Dense& dn, OrthWork& ow, int n, int b, const T* W, int ldw, T* Qout,
                  int ldq, T* tmp, int ldt, double rtol, std::vector<double>* Rout,
                  bool colscale = false, bool one_pass_ok = false) {
  const double u = std::is_same<T, float>::value ? 0x1p-24 : 0x1p-53;
  ow.G.ensure((size_t)b * b);
  ow.Rinv.ensure((size_t)b * b);
  ow.gemm_work.ensure(gemm_tn_work_elems(n, b, b));
  ow.hG.resize((size_t)b * b);
  // R = R_passes ... R_1 · D.  The full product is formed only when the
  // caller wants R (block Lanczos); the rank rule needs only its diagonal,
  // which is the product of the diagonals.
  const bool full = Rout != nullptr;
  std::vector<double> Racc(full ? (size_t)b * b : 0, 0.0), Rp((size_t)b * b), dacc(b, 1.0);
  const T* cur = W;
  int ldc = ldw;
  if (colscale) {
    // equilibrate: unit columns first (same span; columns of very different
    // norms, as after a filter, would make the Gram matrix ill-conditioned)
    std::vector<double> d(b), dinv(b);
    ow.cn.ensure(b);
    ow.red.ensure(colred_work_elems(n, b));
    col_dots<T>(cx.stream, n, b, W, ldw, W, ldw, ow.cn.p, ow.red.p, ow.red.n);
    SMALGO_HIP(hipMemcpyAsync(d.data(), ow.cn.p, sizeof(double) * b, hipMemcpyDeviceToHost,
                              cx.stream));
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
    for (int i = 0; i < b; ++i) {
      d[i] = d[i] > 0 ? std::sqrt(d[i]) : 1.0;
      dinv[i] = 1.0 / d[i];
      dacc[i] = d[i];
    }
    SMALGO_HIP(hipMemcpyAsync(ow.cn.p, dinv.data(), sizeof(double) * b, hipMemcpyHostToDevice,
                              cx.stream));
    copy_block<T, T>(cx.stream, n, b, W, ldw, tmp, ldt);
    scale_cols<T>(cx.stream, n, b, tmp, ldt, ow.cn.p);
    cur = tmp;
    ldc = ldt;
  }
  if (full)
    for (int i = 0; i < b; ++i) Racc[(size_t)i * b + i] = dacc[i];
  bool shifted = false;
  int passes = 2;
  for (int pass = 0; pass < passes; ++pass) {
    // never read and write the same block in one GEMM: alternate tmp <-> Qout
    T* dst = (cur == tmp) ? Qout : tmp;
    int ldd = (cur == tmp) ? ldq : ldt;
    gemm_tn<T, T, double>(cx.stream, n, b, b, 1.0, cur, ldc, cur, ldc, 0.0, ow.G.p, b,
                          ow.gemm_work.p, ow.gemm_work.n);
    SMALGO_HIP(hipMemcpyAsync(ow.hG.data(), ow.G.p, sizeof(double) * b * b, hipMemcpyDeviceToHost,
                              cx.stream));
    int info = chol_inv_sync(cx, dn, ow, b);
    double tr = 0.0;
    for (int i = 0; i < b; ++i) tr += std::max(0.0, ow.hG[(size_t)i * b + i]);
    if (one_pass_ok && colscale && pass == 0 && info == 0) {
      // Equilibrated columns: G = I + E.  With ||E||_F <= 0.1, cond(G) <= 1.1/0.9,
      // so the first pass already leaves Q orthonormal to a few units of
      // rounding (the error grows with cond(W)^2 = cond(G)); the second pass
      // is skipped.
      double e2 = 0.0;
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < b; ++i) {
          const double g = ow.hG[(size_t)j * b + i] - (i == j ? 1.0 : 0.0);
          e2 += g * g;
        }
      if (e2 <= 0.01) passes = 1;
    }
    double shift = 11.0 * ((double)n * b + (double)b * (b + 1)) * u * tr;
    int tries = 0;
    while (info != 0) {
      if (tr == 0.0 || !std::isfinite(tr) || ++tries > 6) {
        if (Rout) Rout->assign((size_t)b * b, 0.0);
        return 0;
      }
      shifted = true;
      SMALGO_HIP(hipMemcpyAsync(ow.G.p, ow.hG.data(), sizeof(double) * b * b,
                                hipMemcpyHostToDevice, cx.stream));
      add_diag(cx.stream, b, ow.G.p, b, shift);
      info = chol_inv_sync(cx, dn, ow, b);
      shift *= 100.0;
    }
    if (shifted && passes == 2) passes = 3;
    SMALGO_HIP(hipMemcpyAsync(Rp.data(), ow.R2.p, sizeof(double) * b * b, hipMemcpyDeviceToHost,
                              cx.stream));
    // dst = cur · R^{-1}; R^{-1} column-major -> transB
    gemm_nn<T, double, T>(cx.stream, n, b, b, 1.0, cur, ldc, ow.Rinv.p, b, true, 0.0, dst, ldd);
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
    if (full) {
      // Racc <- Rp · Racc   (both upper triangular, column-major)
      std::vector<double> Rn((size_t)b * b, 0.0);
      for (int j = 0; j < b; ++j)
        for (int i = 0; i <= j; ++i) {
          double s = 0.0;
          for (int k = i; k <= j; ++k) s += Rp[(size_t)k * b + i] * Racc[(size_t)j * b + k];
          Rn[(size_t)j * b + i] = s;
        }
      Racc.swap(Rn);
    }
    for (int i = 0; i < b; ++i) dacc[i] *= Rp[(size_t)i * b + i];
    cur = dst;
    ldc = ldd;
  }
  if (cur != Qout) copy_block<T, T>(cx.stream, n, b, cur, ldc, Qout, ldq);
  double r0 = 0.0;
  for (int i = 0; i < b; ++i) r0 = std::max(r0, std::fabs(dacc[i]));
  int r = 0;
  if (r0 > 0)
    for (int i = 0; i < b; ++i) r += std::fabs(dacc[i]) > rtol * r0 ? 1 : 0;
  if (Rout) *Rout = std::move(Racc);
  return r;
}

// CholQR2 of block Lanczos without host round trips: both passes, the
// product R = R_2·R_1 and the status stay on the device, and one small
// download (the two Cholesky infos and diag(R)) gives the rank.  Returns -1
// when a Cholesky factorization failed; the caller then runs cholqr(), whose
// shifted CholQR3 handles a Gram matrix that is not numerically positive
// definite.  W is only read, so that retry starts from the same block.
template <typename T>
static int cholqr_fast(Ctx& cx, Dense& dn, OrthWork& ow, int n, int b, const T* W, int ldw,
                       T* Qout, int ldq, T* tmp, int ldt, double rtol, double* Rdev) {
  ow.G.ensure((size_t)b * b);
  ow.R1.ensure((size_t)b * b);
  ow.gemm_work.ensure(gemm_tn_work_elems(n, b, b));
  ow.info2.ensure(2);
  ow.status.ensure((size_t)b + 2);
  ow.hstatus.resize((size_t)b + 2);
  if (b <= chol_inv_max_b()) {
    // one-workgroup Cholesky + triangular inverse per pass (R_1 and R_2 kept)
    ow.R2.ensure((size_t)b * b);
    ow.Rinv.ensure((size_t)b * b);
    // pass 1: G = WᵀW = R_1ᵀR_1, tmp = W R_1⁻¹
    gemm_tn<T, T, double>(cx.stream, n, b, b, 1.0, W, ldw, W, ldw, 0.0, ow.G.p, b,
                          ow.gemm_work.p, ow.gemm_work.n);
    chol_inv_upper(cx.stream, b, ow.G.p, b, ow.R1.p, b, ow.Rinv.p, b, ow.info2.p);
    gemm_nn<T, double, T>(cx.stream, n, b, b, 1.0, W, ldw, ow.Rinv.p, b, true, 0.0, tmp, ldt);
    // pass 2: tmpᵀtmp = R_2ᵀR_2, Q = tmp R_2⁻¹, R = R_2 R_1
    gemm_tn<T, T, double>(cx.stream, n, b, b, 1.0, tmp, ldt, tmp, ldt, 0.0, ow.G.p, b,
                          ow.gemm_work.p, ow.gemm_work.n);
    chol_inv_upper(cx.stream, b, ow.G.p, b, ow.R2.p, b, ow.Rinv.p, b, ow.info2.p + 1);
    tri_mul_upper(cx.stream, b, ow.R2.p, ow.R1.p, Rdev);
    gemm_nn<T, double, T>(cx.stream, n, b, b, 1.0, tmp, ldt, ow.Rinv.p, b, true, 0.0, Qout, ldq);
  } else {
    // pass 1: G = WᵀW = R_1ᵀR_1, tmp = W R_1⁻¹
    gemm_tn<T, T, double>(cx.stream, n, b, b, 1.0, W, ldw, W, ldw, 0.0, ow.G.p, b,
                          ow.gemm_work.p, ow.gemm_work.n);
    dn.potrf_upper_async(b, ow.G.p, b, ow.info2.p);
    zero_lower_cm(cx.stream, b, ow.G.p, b);
    SMALGO_HIP(hipMemcpyAsync(ow.R1.p, ow.G.p, sizeof(double) * b * b, hipMemcpyDeviceToDevice,
                              cx.stream));
    dn.trtri_upper(b, ow.G.p, b);
    gemm_nn<T, double, T>(cx.stream, n, b, b, 1.0, W, ldw, ow.G.p, b, true, 0.0, tmp, ldt);
    // pass 2: tmpᵀtmp = R_2ᵀR_2, Q = tmp R_2⁻¹, R = R_2 R_1
    gemm_tn<T, T, double>(cx.stream, n, b, b, 1.0, tmp, ldt, tmp, ldt, 0.0, ow.G.p, b,
                          ow.gemm_work.p, ow.gemm_work.n);
    dn.potrf_upper_async(b, ow.G.p, b, ow.info2.p + 1);
    zero_lower_cm(cx.stream, b, ow.G.p, b);
    tri_mul_upper(cx.stream, b, ow.G.p, ow.R1.p, Rdev);
    dn.trtri_upper(b, ow.G.p, b);
    gemm_nn<T, double, T>(cx.stream, n, b, b, 1.0, tmp, ldt, ow.G.p, b, true, 0.0, Qout, ldq);
  }
  pack_qr_status(cx.stream, b, ow.info2.p, Rdev, ow.status.p);
  SMALGO_HIP(hipMemcpyAsync(ow.hstatus.data(), ow.status.p, sizeof(double) * (b + 2),
                            hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  if (ow.hstatus[0] != 0.0 || ow.hstatus[1] != 0.0) return -1;
  double r0 = 0.0;
  for (int i = 0; i < b; ++i) {
    const double d = std::fabs(ow.hstatus[2 + i]);
    if (!std::isfinite(d)) return -1;
    r0 = std::max(r0, d);
  }
  int r = 0;
  if (r0 > 0)
    for (int i = 0; i < b; ++i) r += std::fabs(ow.hstatus[2 + i]) > rtol * r0 ? 1 : 0;
  return r;
}

// ---------------------------------------------------------------------------
// formatting of the progress line (polfed.jl:991-1000)
// ---------------------------------------------------------------------------
// Julia's print(::Float64): the shortest digit string that round-trips
// (d_1 … d_k, value = 0.d_1…d_k · 10^pt), written in fixed notation when
// -4 < pt <= 6 and otherwise as d_1.d_2…d_k e<exp> with exp = pt - 1, no
// '+' and no zero padding ("1.0e-5", "1.0927391319480254e-9", "0.0001",
// "123456.0", "1.234567e6").  A single digit keeps a trailing ".0".
std::string format_julia_float(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "Inf" : "-Inf";
  if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
  char buf[64];
  // shortest round-trip digits in scientific form: [-]d[.ddd]e±XX
  auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  std::string out;
  size_t i = 0;
  if (s[0] == '-') {
    out += '-';
    i = 1;
  }
  const size_t epos = s.find('e');
  std::string digits;
  for (size_t j = i; j < epos; ++j)
    if (s[j] != '.') digits += s[j];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  const int e10 = std::atoi(s.c_str() + epos + 1);  // value = d.ddd · 10^e10
  const int k = (int)digits.size();
  const int pt = e10 + 1;
  if (-4 < pt && pt <= 6) {
    if (pt <= 0) {
      out += "0.";
      out.append((size_t)(-pt), '0');
      out += digits;
    } else if (pt >= k) {
      out += digits;
      out.append((size_t)(pt - k), '0');
      out += ".0";
    } else {
      out += digits.substr(0, pt);
      out += '.';
      out += digits.substr(pt);
    }
    return out;
  }
  out += digits[0];
  out += '.';
  out += k > 1 ? digits.substr(1) : std::string("0");
  out += 'e';
  out += std::to_string(e10);
  return out;
}

// Julia's @printf("%.3e", x): C's %.3e, except that infinities and NaN are
// written Inf, -Inf and NaN.
static std::string printf_e3(double x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x > 0 ? "Inf" : "-Inf";
  char buf[32];
  std::snprintf(buf, sizeof buf, "%.3e", x);
  return buf;
}

std::string progress_line(int m, const std::vector<double>& lam, const std::vector<double>& res) {
  if (res.empty()) return "m=" + std::to_string(m) + "  inside=0  maxres=Inf";
  std::vector<double> rs(res);
  std::sort(rs.begin(), rs.end());
  const int L = (int)rs.size();
  auto ridx = [&](double f) {
    int k = (int)std::nearbyint(f * L);
    return std::min(std::max(k, 1), L) - 1;
  };
  return "m=" + std::to_string(m) + "  inside=" + std::to_string(lam.size()) +
         "  r90=" + format_julia_float(rs[ridx(0.9)]) + "  r95=" + format_julia_float(rs[ridx(0.95)]) +
         "  maxres=" + format_julia_float(rs.back());
}

static std::vector<int> stable_argsort(const std::vector<double>& v, bool descending = false) {
  std::vector<int> p(v.size());
  std::iota(p.begin(), p.end(), 0);
  if (descending)
    std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return v[a] > v[b]; });
  else
    std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return v[a] < v[b]; });
  return p;
}

static DBuf<int> upload_idx(Ctx& cx, const std::vector<int>& v) {
  DBuf<int> d(std::max<size_t>(v.size(), 1));
  if (!v.empty())
    SMALGO_HIP(hipMemcpyAsync(d.p, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice,
                              cx.stream));
  return d;
}

// Rayleigh quotients and residuals of the columns of X (FP64):
// λ_j = x_jᵀ A x_j, res_j = ‖A x_j − λ_j x_j‖ (rayleigh!, polfed.jl:791-830)
struct RayleighWork {
  DBuf<double> AX, lam, res, red;
};
static void rayleigh(Ctx& cx, const Op& op, int n, int k, const double* X, int ldx, RayleighWork& w,
                     std::vector<double>& lam, std::vector<double>& res) {
  const int ld = round_up(std::max(k, 1), 4);
  w.AX.ensure((size_t)n * ld);
  w.lam.ensure(std::max(k, 1));
  w.res.ensure(std::max(k, 1));
  w.red.ensure(colred_work_elems(n, k));
  apply_A<double>(cx, op, k, X, ldx, w.AX.p, ld);
  col_dots<double>(cx.stream, n, k, X, ldx, w.AX.p, ld, w.lam.p, w.red.p, w.red.n);
  col_resid<double>(cx.stream, n, k, X, ldx, w.AX.p, ld, w.lam.p, w.res.p, w.red.p, w.red.n);
  lam.resize(k);
  res.resize(k);
  if (k) {
    SMALGO_HIP(hipMemcpyAsync(lam.data(), w.lam.p, sizeof(double) * k, hipMemcpyDeviceToHost,
                              cx.stream));
    SMALGO_HIP(hipMemcpyAsync(res.data(), w.res.p, sizeof(double) * k, hipMemcpyDeviceToHost,
                              cx.stream));
  }
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}

// Original-space Rayleigh–Ritz (refine_original_rr, polfed.jl:837-880) on the
// k columns of X (row-major FP64).  Output Y (n x kq, ld ldy, sorted by λ).
struct RRWork {
  OrthWork ow;
  DBuf<double> Q, tmp, AQ, H, w, e, gw;
  RayleighWork rw;
};
static int refine_rr(Ctx& cx, Dense& dn, const Op& op, int n, int k, const double* X, int ldx,
                     RRWork& w, DBuf<double>& Y, int& ldy, std::vector<double>& lam,
                     std::vector<double>& res, double rtol = 1e-12) {
  const int ld = round_up(std::max(k, 1), 4);
  w.Q.ensure((size_t)n * ld);
  w.tmp.ensure((size_t)n * ld);
  int kq = cholqr<double>(cx, dn, w.ow, n, k, X, ldx, w.Q.p, ld, w.tmp.p, ld, rtol, nullptr,
                          true, true);
  if (kq == 0) {
    lam.clear();
    res.clear();
    return 0;
  }
  w.AQ.ensure((size_t)n * ld);
  apply_A<double>(cx, op, kq, w.Q.p, ld, w.AQ.p, ld);
  w.H.ensure((size_t)kq * kq);
  w.gw.ensure(gemm_tn_work_elems(n, kq, kq));
  gemm_tn<double, double, double>(cx.stream, n, kq, kq, 1.0, w.Q.p, ld, w.AQ.p, ld, 0.0, w.H.p, kq,
                                  w.gw.p, w.gw.n);
  symmetrize(cx.stream, kq, w.H.p, kq);
  w.w.ensure(kq);
  w.e.ensure(kq);
  dn.syevd(kq, w.H.p, kq, w.w.p, w.e.p, true);
  ldy = ld;
  Y.ensure((size_t)n * ld);
  gemm_nn<double, double, double>(cx.stream, n, kq, kq, 1.0, w.Q.p, ld, w.H.p, kq, true, 0.0, Y.p,
                                  ld);
  std::vector<double> l, r;
  rayleigh(cx, op, n, kq, Y.p, ld, w.rw, l, r);
  std::vector<int> p = stable_argsort(l);
  lam.resize(kq);
  res.resize(kq);
  for (int i = 0; i < kq; ++i) {
    lam[i] = l[p[i]];
    res[i] = r[p[i]];
  }
  DBuf<int> dp = upload_idx(cx, p);
  gather_cols_rm<double>(cx.stream, n, kq, Y.p, ld, dp.p, w.tmp.p, ld);
  SMALGO_HIP(hipMemcpyAsync(Y.p, w.tmp.p, sizeof(double) * (size_t)n * ld, hipMemcpyDeviceToDevice,
                            cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  return kq;
}

// ---------------------------------------------------------------------------
// one-call block kernels on host column-major arrays
// ---------------------------------------------------------------------------
template <typename T>
static void upload_cm(Ctx& cx, int n, int b, const double* H, int ldh, T* D, int ldd) {
  DBuf<double> cm((size_t)n * std::max(b, 1));
  for (int j = 0; j < b; ++j)
    SMALGO_HIP(hipMemcpyAsync(cm.p + (size_t)j * n, H + (size_t)j * ldh, sizeof(double) * n,
                              hipMemcpyHostToDevice, cx.stream));
  cm_to_rm<double, T>(cx.stream, n, b, cm.p, n, D, ldd);
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}
template <typename T>
static void download_cm(Ctx& cx, int n, int b, const T* D, int ldd, double* H, int ldh) {
  DBuf<double> cm((size_t)n * std::max(b, 1));
  rm_to_cm<T>(cx.stream, n, b, D, ldd, nullptr, cm.p, n);
  for (int j = 0; j < b; ++j)
    SMALGO_HIP(hipMemcpyAsync(H + (size_t)j * ldh, cm.p + (size_t)j * n, sizeof(double) * n,
                              hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}

template <typename T>
static void orth_against_t(Ctx& cx, int n, int b, double* W, int ldw, int m, const double* Q,
                           int ldq) {
  const int lb = round_up(std::max(b, 1), 4), lm = round_up(std::max(m, 1), 4);
  DBuf<T> dW((size_t)n * lb), dQ((size_t)n * lm), C((size_t)lm * lb);
  DBuf<double> gw(gemm_tn_work_elems(n, m, b));
  upload_cm<T>(cx, n, b, W, ldw, dW.p, lb);
  upload_cm<T>(cx, n, m, Q, ldq, dQ.p, lm);
  gemm_tn<T, T, T>(cx.stream, n, m, b, 1.0, dQ.p, lm, dW.p, lb, 0.0, C.p, lb, gw.p, gw.n);
  gemm_nn<T, T, T>(cx.stream, n, m, b, -1.0, dQ.p, lm, C.p, lb, false, 1.0, dW.p, lb);
  download_cm<T>(cx, n, b, dW.p, lb, W, ldw);
}
void orth_against_host(Ctx& cx, int n, int b, double* W, int ldw, int m, const double* Q, int ldq,
                       int precision) {
  if (m == 0 || b == 0) return;
  if (precision == SMALGO_PREC_FP32_REFINE64)
    orth_against_t<float>(cx, n, b, W, ldw, m, Q, ldq);
  else
    orth_against_t<double>(cx, n, b, W, ldw, m, Q, ldq);
}

template <typename T>
static void block_qr_t(Ctx& cx, int n, int b, const double* W, int ldw, double rtol, double* Q,
                       int ldq, double* R, int ldr, int* rank) {
  const int lb = round_up(std::max(b, 1), 4);
  DBuf<T> dW((size_t)n * lb), dQ((size_t)n * lb), tmp((size_t)n * lb);
  upload_cm<T>(cx, n, b, W, ldw, dW.p, lb);
  Dense dn(cx);
  OrthWork ow;
  std::vector<double> Rh;
  *rank = cholqr<T>(cx, dn, ow, n, b, dW.p, lb, dQ.p, lb, tmp.p, lb, rtol, &Rh);
  download_cm<T>(cx, n, b, dQ.p, lb, Q, ldq);
  if (R)
    for (int j = 0; j < b; ++j)
      for (int i = 0; i < b; ++i) R[(size_t)j * ldr + i] = Rh[(size_t)j * b + i];
}
void block_qr_host(Ctx& cx, int n, int b, const double* W, int ldw, double rtol, int precision,
                   double* Q, int ldq, double* R, int ldr, int* rank) {
  if (precision == SMALGO_PREC_FP32_REFINE64)
    block_qr_t<float>(cx, n, b, W, ldw, rtol, Q, ldq, R, ldr, rank);
  else
    block_qr_t<double>(cx, n, b, W, ldw, rtol, Q, ldq, R, ldr, rank);
}

void build_X_host(Ctx& cx, int n, int M, int k, const double* V, int ldv, const double* U, int ldu,
                  double* X, int ldx) {
  const int lm = round_up(std::max(M, 1), 4), lk = round_up(std::max(k, 1), 4);
  DBuf<double> dV((size_t)n * lm), dU((size_t)M * lk), dX((size_t)n * lk);
  upload_cm<double>(cx, n, M, V, ldv, dV.p, lm);
  upload_cm<double>(cx, M, k, U, ldu, dU.p, lk);
  gemm_nn<double, double, double>(cx.stream, n, M, k, 1.0, dV.p, lm, dU.p, lk, false, 0.0, dX.p, lk);
  download_cm<double>(cx, n, k, dX.p, lk, X, ldx);
}

void rayleigh_host(Ctx& cx, const Op& op, int k, const double* X, int ldx, double* lam, double* res) {
  const int n = op.n, lk = round_up(std::max(k, 1), 4);
  DBuf<double> dX((size_t)n * lk);
  upload_cm<double>(cx, n, k, X, ldx, dX.p, lk);
  RayleighWork w;
  std::vector<double> l, r;
  rayleigh(cx, op, n, k, dX.p, lk, w, l, r);
  std::copy(l.begin(), l.end(), lam);
  std::copy(r.begin(), r.end(), res);
}

void residuals_host(Ctx& cx, const Op& op, int k, const double* X, int ldx, const double* lam,
                    double* res) {
  const int n = op.n, lk = round_up(std::max(k, 1), 4);
  DBuf<double> dX((size_t)n * lk), AX((size_t)n * lk), dl(std::max(k, 1)), dr(std::max(k, 1)),
      red(colred_work_elems(n, k));
  upload_cm<double>(cx, n, k, X, ldx, dX.p, lk);
  SMALGO_HIP(hipMemcpyAsync(dl.p, lam, sizeof(double) * k, hipMemcpyHostToDevice, cx.stream));
  apply_A<double>(cx, op, k, dX.p, lk, AX.p, lk);
  col_resid<double>(cx.stream, n, k, dX.p, lk, AX.p, lk, dl.p, dr.p, red.p, red.n);
  SMALGO_HIP(hipMemcpyAsync(res, dr.p, sizeof(double) * k, hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}

// The Rayleigh–Ritz tail of the reference's subspace iterations
// (chebyshev_subspace_iteration.jl:733-748, feast.jl:540-553,
// zoloterov.jl:1698-1706, real symmetric case) on a host block Y:
// Q = thin_orth(Y) keeping all k columns, Aq = Qᵀ(A Q), eigen(Symmetric(Aq)),
// X = Q Xq, res_j = ‖A x_j - λ_j x_j‖₂.  All k pairs in ascending λ.
Result* subspace_rr_host(Ctx& cx, const Op& op, int k, const double* Y, int ldy, double* Qh,
                         int ldqh) {
  const int n = op.n;
  std::unique_ptr<Result> out(new Result());
  out->n = n;
  out->device = cx.device;
  {
    if (k == 0) return out.release();
    const int ld = round_up(k, 4);
    Dense dn(cx);
    OrthWork ow;
    DBuf<double> Yd((size_t)n * ld), Q((size_t)n * ld), T((size_t)n * ld), AX((size_t)n * ld),
        X((size_t)n * ld), Aq((size_t)k * k), wv(k), ev(k), resd(k);
    upload_cm<double>(cx, n, k, Y, ldy, Yd.p, ld);
    if (cholqr<double>(cx, dn, ow, n, k, Yd.p, ld, Q.p, ld, T.p, ld, 0.0, nullptr, true) == 0)
      throw Error(-4, "thin_orth: Cholesky QR of Y failed (zero or non-finite block)");
    apply_A<double>(cx, op, k, Q.p, ld, AX.p, ld);
    DBuf<double> gw(gemm_tn_work_elems(n, k, k));
    gemm_tn<double, double, double>(cx.stream, n, k, k, 1.0, Q.p, ld, AX.p, ld, 0.0, Aq.p, k, gw.p,
                                    gw.n);
    // Symmetric(Aq) reads the upper triangle; Aq is row-major, i.e. the lower
    // triangle of its column-major view
    dn.syevd(k, Aq.p, k, wv.p, ev.p, false);
    gemm_nn<double, double, double>(cx.stream, n, k, k, 1.0, Q.p, ld, Aq.p, k, true, 0.0, X.p, ld);
    apply_A<double>(cx, op, k, X.p, ld, AX.p, ld);
    DBuf<double> red(colred_work_elems(n, k));
    col_resid<double>(cx.stream, n, k, X.p, ld, AX.p, ld, wv.p, resd.p, red.p, red.n);
    out->lam.resize(k);
    out->res.resize(k);
    SMALGO_HIP(hipMemcpyAsync(out->lam.data(), wv.p, sizeof(double) * k, hipMemcpyDeviceToHost,
                              cx.stream));
    SMALGO_HIP(hipMemcpyAsync(out->res.data(), resd.p, sizeof(double) * k,
                              hipMemcpyDeviceToHost, cx.stream));
    out->alloc_X(sizeof(double) * (size_t)n * k);
    rm_to_cm<double>(cx.stream, n, k, X.p, ld, nullptr, out->Xdev, n);
    if (Qh) download_cm<double>(cx, n, k, Q.p, ld, Qh, ldqh);
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
    out->line = "subspace_rr k=" + std::to_string(k);
  }
  return out.release();
}

Result* refine_rr_host(Ctx& cx, const Op& op, int k, const double* X, int ldx) {
  const int n = op.n, lk = round_up(std::max(k, 1), 4);
  DBuf<double> dX((size_t)n * lk);
  upload_cm<double>(cx, n, k, X, ldx, dX.p, lk);
  Dense dn(cx);
  RRWork w;
  DBuf<double> Y;
  int ldy = 4;
  std::unique_ptr<Result> out(new Result());
  out->n = n;
  out->device = cx.device;
  const int kq = refine_rr(cx, dn, op, n, k, dX.p, lk, w, Y, ldy, out->lam, out->res);
  if (kq > 0) {
    out->alloc_X(sizeof(double) * (size_t)n * kq);
    rm_to_cm<double>(cx.stream, n, kq, Y.p, ldy, nullptr, out->Xdev, n);
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
  }
  return out.release();
}

// ---------------------------------------------------------------------------
// KPM moments with the doubling identities (Weiße et al., Rev. Mod. Phys. 78,
// 275 (2006), Sec. II.B): per Chebyshev step one fused SpMM of the R probes.
// ---------------------------------------------------------------------------
void kpm_moments(Ctx& cx, const Op& op, int nmom, int R, uint64_t seed, double smin, double smax,
                 double* mu) {
  const int n = op.n;
  if (nmom < 2 || R < 1) throw Error(-1, "need nmom >= 2 and R >= 1");
  if (!(smax > smin)) throw Error(-1, "need smin < smax");
  if (op.is_callback()) throw Error(-1, "KPM runs on device matrices");
  const int ld = round_up(R, 4);
  const double c = 0.5 * (smin + smax), invd = 2.0 / (smax - smin);
  DBuf<double> r0((size_t)n * ld), ta((size_t)n * ld), tb((size_t)n * ld);
  const int kmax = (nmom + 1) / 2;  // T_0 .. T_kmax needed
  DBuf<double> dots((size_t)(2 * kmax + 2) * ld), red(colred_work_elems(n, R));
  gaussian_block<double>(cx.stream, n, R, seed, r0.p, ld);
  const FilterTile t = choose_tile(sizeof(double), R, {r0.p, ta.p, tb.p}, {ld});
  // T_1 = Â r
  launch_clenshaw_step<double>(cx.stream, *op.A, R, nullptr, 0, 0.0, r0.p, ld, invd, -c * invd,
                               nullptr, 0, 0.0, ta.p, ld, t);
  // slot 2k: <T_k,T_k>, slot 2k+1: <T_k+1,T_k>
  col_dots<double>(cx.stream, n, R, r0.p, ld, r0.p, ld, dots.p + 0 * ld, red.p, red.n);
  col_dots<double>(cx.stream, n, R, ta.p, ld, r0.p, ld, dots.p + 1 * ld, red.p, red.n);
  double* Tk = ta.p;        // T_k   (k = 1)
  double* Tkm1 = r0.p;      // T_k-1
  double* spare = tb.p;
  for (int k = 1; k < kmax; ++k) {
    // T_k+1 = 2Â T_k − T_k-1  (written over T_k-1 unless that is the probe block)
    double* out = (Tkm1 == r0.p) ? spare : Tkm1;
    launch_clenshaw_step<double>(cx.stream, *op.A, R, nullptr, 0, 0.0, Tk, ld, 2 * invd,
                                 -2 * c * invd, Tkm1, ld, 1.0, out, ld, t);
    col_dots<double>(cx.stream, n, R, Tk, ld, Tk, ld, dots.p + (size_t)(2 * k) * ld, red.p, red.n);
    col_dots<double>(cx.stream, n, R, out, ld, Tk, ld, dots.p + (size_t)(2 * k + 1) * ld, red.p,
                     red.n);
    Tkm1 = Tk;
    Tk = out;
  }
  std::vector<double> h((size_t)(2 * kmax) * ld);
  SMALGO_HIP(hipMemcpyAsync(h.data(), dots.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                            cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  auto avg = [&](int slot) {
    double s = 0;
    for (int r = 0; r < R; ++r) s += h[(size_t)slot * ld + r];
    return s / ((double)R * n);
  };
  const double m0 = avg(0), m1 = avg(1);
  mu[0] = m0;
  if (nmom > 1) mu[1] = m1;
  for (int k = 1; k < kmax; ++k) {
    if (2 * k < nmom) mu[2 * k] = 2 * avg(2 * k) - m0;
    if (2 * k + 1 < nmom) mu[2 * k + 1] = 2 * avg(2 * k + 1) - m1;
  }
}

// ---------------------------------------------------------------------------
// Batched spectral slices (no reference counterpart: polfed, polfed.jl:1128,
// treats one interval per call).  The current blocks of many intervals are
// packed column-wise into dense n x pw panels and filtered together: every
// column carries its own Chebyshev coefficients, so one launch serves
// several slices, and narrow slices share the gathered rows of a wide panel.
// ---------------------------------------------------------------------------
struct SlicePack {
  struct Seg {
    int slice, c0, w, panel, off;  // slice columns [c0, c0+w) -> panel column off
  };
  struct Panel {
    int w = 0, K = 0;  // width (multiple of 4) and degree (max over its slices)
    size_t base = 0;   // element offset of the panel's n x w slab (set by the filter)
    FilterTile tile{};
  };
  std::vector<Seg> segs;
  std::vector<Panel> pan;
  std::vector<std::vector<int>> col_slice;  // slice of every panel column (-1: padding)
  std::vector<std::vector<int>> col_idx;    // its column within the slice
  int wmax = 0;
  long long steps = 0;  // Σ_p K_p
  double pred_ms = 0;   // modelled time of one run
};

// Packing.  Slices are taken in order of decreasing degree.  With a fixed
// width every panel is wfix wide, except that a last panel the columns do not
// fill is cut to them (rounded up to 4); a slice that does not fit the room
// left continues in the next panel.  Otherwise panels have their own widths (up
// to wmax, multiples of 4): a slice joins the open panel when the measured
// cost of widening it, K_p (t(w+s) - t(w)), is below the cost of a panel of
// its own, K_i t(s), where t(w) is the tuned time of one step at width w.
// Joining shares the matrix read of each step and widens the gathered rows;
// its price is the steps the joining columns wait through leading zero
// coefficients (K_p - K_i).
template <typename TimeFn>
static SlicePack pack_slices(const std::vector<int>& widths, const std::vector<int>& degrees,
                             int wmax, int wfix, TimeFn&& t) {
  SlicePack pk;
  pk.wmax = wfix > 0 ? wfix : wmax;
  const int ns = (int)widths.size();
  std::vector<int> ord(ns);
  for (int i = 0; i < ns; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return degrees[a] > degrees[b]; });
  int used = 0;  // columns used in the open panel (-1: none open)
  bool open = false;
  auto start = [&](int K) {
    pk.pan.push_back({});
    pk.pan.back().K = K;
    pk.col_slice.emplace_back();
    pk.col_idx.emplace_back();
    used = 0;
    open = true;
  };
  for (int i : ord) {
    int c0 = 0;
    while (c0 < widths[i]) {
      const int rem = widths[i] - c0;
      bool join = false;
      if (open && used < pk.wmax) {
        if (wfix > 0) {
          join = true;
        } else {
          const int take = std::min(rem, pk.wmax - used);
          const int Kp = pk.pan.back().K;
          const double widen = Kp * (t(round_up(used + take, 4)) - t(round_up(used, 4)));
          const double alone = degrees[i] * t(round_up(take, 4));
          join = widen <= alone;
        }
      }
      if (!join) start(degrees[i]);
      const int w = std::min(rem, pk.wmax - used);
      const int panel = (int)pk.pan.size() - 1;
      pk.segs.push_back({i, c0, w, panel, used});
      for (int c = 0; c < w; ++c) {
        pk.col_slice[panel].push_back(i);
        pk.col_idx[panel].push_back(c0 + c);
      }
      used += w;
      c0 += w;
      if (used == pk.wmax) open = false;
    }
  }
  for (size_t p = 0; p < pk.pan.size(); ++p) {
    SlicePack::Panel& P = pk.pan[p];
    P.w = round_up((int)pk.col_slice[p].size(), 4);  // wfix for every full panel
    pk.col_slice[p].resize(P.w, -1);
    pk.col_idx[p].resize(P.w, 0);
    pk.steps += P.K;
    pk.pred_ms += P.K * t(P.w);
  }
  return pk;
}

// Chebyshev coefficients of the columns of one slice: one vector shared by
// all its columns, or one vector per column.  Coefficients past a vector's
// end are zero.
struct SliceCoeffs {
  const std::vector<double>* shared = nullptr;
  const std::vector<std::vector<double>>* percol = nullptr;
  int degree() const {
    if (shared) return (int)shared->size() - 1;
    int K = 0;
    for (const auto& v : *percol) K = std::max(K, (int)v.size() - 1);
    return K;
  }
  double at(int j, int k) const {
    const std::vector<double>& v = shared ? *shared : (*percol)[j];
    return k < (int)v.size() ? v[k] : 0.0;
  }
};

template <typename T>
struct SliceFilter {
  // One fused step of a panel: out = cq∘Q + a1 (A X1) + b1 X1 - s2 X2 with a
  // per-column cq row; operands by role (0 none, 1 Q, 2 P0, 3 P1, 4 Y).
  struct Step {
    int x1, x2, out;
    size_t cq;  // table offset of the cq row
    double a1, b1, s2;
  };
  Ctx& cx;
  const Op& op;
  int n = 0;
  SlicePack pk;
  std::vector<int> widths;
  DBuf<T> Qp, Yp, P0, P1, tab;  // panels, ping-pong buffers, coefficient table
  DBuf<T> P0s, P1s;             // ping-pong buffers of the side chain (paired panels)
  std::vector<std::vector<Step>> steps;
  // Paired chains: panels whose steps do not split and whose width lets two
  // of them (Q, B1, B2 each, plus the matrix) fit the cache budget run as two
  // independent chains, one per stream (chain[1] on the side stream with
  // P0s/P1s), their launches interleaved.  The other panels (solo) run first,
  // one after another, splitting their own steps where their tile does.
  bool paired = false;
  std::vector<int> solo, chain[2];
  std::vector<hipGraphExec_t> execs;
  std::vector<hipGraph_t> graphs;
  int graph_nodes_max = 2048;  // launches per captured graph

  SliceFilter(Ctx& c, const Op& o) : cx(c), op(o), n(o.n) {}
  SliceFilter(const SliceFilter&) = delete;
  SliceFilter& operator=(const SliceFilter&) = delete;
  ~SliceFilter() { drop_graphs(); }

  void drop_graphs() {
    for (auto e : execs) (void)hipGraphExecDestroy(e);
    for (auto g : graphs) (void)hipGraphDestroy(g);
    execs.clear();
    graphs.clear();
  }

  // Build the packing, the per-column coefficient table and the step list.
  // wfix > 0 gives every full panel that width; otherwise panels are sized by the
  // measured step time of each width, up to the cache-aware panel width.
  void plan(const std::vector<int>& w, const std::vector<SliceCoeffs>& co, double smin,
            double smax, int wfix) {
    drop_graphs();
    widths = w;
    std::vector<int> deg(w.size());
    for (size_t i = 0; i < w.size(); ++i) deg[i] = co[i].degree();
    const int wmax = wfix > 0 ? wfix : cache_panel_width<T>(op);
    P0.ensure((size_t)n * wmax);
    P1.ensure((size_t)n * wmax);
    auto t = [&](int width) {
      if (width <= 0) return 0.0;
      float ms = 0;
      tune(width, &ms);
      return (double)ms;
    };
    pk = pack_slices(w, deg, wmax, wfix, t);
    // Narrower panels that pair into two chains (below) may beat the
    // cache-width plan: plan once more at the widest pairable width with the
    // step times scaled by the measured pairing gain, and keep the cheaper.
    const int wp = pair_width(wmax);
    if (wfix <= 0 && wp >= 8 && wp < wmax) {
      const double r = pair_ratio(wp);
      SlicePack alt = pack_slices(w, deg, wp, 0, [&](int width) { return r * t(width); });
      if (alt.pred_ms < pk.pred_ms) pk = std::move(alt);
    }
    for (auto& P : pk.pan) {
      float ms = 0;
      P.tile = tune(P.w, &ms);
    }
    build_table(co, smin, smax);
    size_t tot = 0;
    for (auto& P : pk.pan) {  // slabs one after another
      P.base = tot;
      tot += (size_t)n * P.w;
    }
    Qp.ensure(tot);
    Yp.ensure(tot);
    for (const auto& P : pk.pan) fill_block<T>(cx.stream, n, P.w, Qp.p + P.base, P.w, T(0));
    fill_block<T>(cx.stream, n, wmax, P0.p, wmax, T(0));
    fill_block<T>(cx.stream, n, wmax, P1.p, wmax, T(0));
    plan_chains();
    if (paired) {
      P0s.ensure((size_t)n * wmax);
      P1s.ensure((size_t)n * wmax);
      fill_block<T>(cx.stream, n, wmax, P0s.p, wmax, T(0));
      fill_block<T>(cx.stream, n, wmax, P1s.p, wmax, T(0));
    }
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
  }

  // Panels go to the chain with fewer steps so far, longest first.
  // SMALGO_SLICE_PAIR=0 keeps every panel solo.
  void plan_chains() {
    solo.clear();
    chain[0].clear();
    chain[1].clear();
    const double mat = (double)op.A->nnz * (sizeof(T) + 4);
    const bool on = pairing_enabled();
    std::vector<int> pair;
    for (size_t q = 0; q < pk.pan.size(); ++q) {
      const int w = pk.pan[q].w;
      const bool fits = 2.0 * 3.0 * n * w * sizeof(T) + mat <= 240e6;
      const bool splits = pk.pan[q].tile.split == 2 && split_ok(w, sizeof(T), {P0.p, P1.p}, {w});
      if (!splits && fits && on)
        pair.push_back((int)q);
      else
        solo.push_back((int)q);
    }
    paired = pair.size() >= 2;
    if (!paired) {
      solo.insert(solo.end(), pair.begin(), pair.end());
      std::sort(solo.begin(), solo.end());
      return;
    }
    std::stable_sort(pair.begin(), pair.end(),
                     [&](int a, int b) { return steps[a].size() > steps[b].size(); });
    size_t load[2] = {0, 0};
    for (int q : pair) {
      const int c = load[1] < load[0] ? 1 : 0;
      chain[c].push_back(q);
      load[c] += steps[q].size();
    }
  }

  // every launch of one run in issue order: (panel, step, side chain)
  template <typename F>
  void each_launch(F&& f) {
    for (int q : solo)
      for (const Step& st : steps[q]) f(q, st, false);
    std::vector<std::pair<int, size_t>> seq[2];
    for (int c = 0; c < 2; ++c)
      for (int q : chain[c])
        for (size_t k = 0; k < steps[q].size(); ++k) seq[c].push_back({q, k});
    const size_t m = std::max(seq[0].size(), seq[1].size());
    for (size_t i = 0; i < m; ++i)
      for (int c = 0; c < 2; ++c)
        if (i < seq[c].size()) f(seq[c][i].first, steps[seq[c][i].first][seq[c][i].second], c == 1);
  }

  static bool pairing_enabled() {
    const char* e = std::getenv("SMALGO_SLICE_PAIR");
    return !(e && e[0] == '0');
  }

  // widest panel (multiple of 4, at most wmax) two of which (Q, B1, B2 each)
  // fit the 240 MB cache budget with the matrix; 0 when pairing is off
  int pair_width(int wmax) const {
    if (!pairing_enabled()) return 0;
    const double mat = (double)op.A->nnz * (sizeof(T) + 4);
    const double per_col = 2.0 * 3.0 * n * sizeof(T);
    const int w = (int)std::min<double>(wmax, std::floor((240e6 - mat) / per_col));
    return w > 0 ? w / 4 * 4 : 0;
  }

  // Step time of two w-wide panels run as two concurrent chains (one per
  // stream, each with the tuned one-launch tile of width w), over twice one
  // chain's step time: the pairing gain the packer plans with (cached per
  // precision, width and structure).
  double pair_ratio(int w) {
    const long long key = tune_key_mix({(long long)sizeof(T), (long long)w, op.mat_id, 2});
    auto it = cx.tuned_ms.find(key);
    if (it != cx.tuned_ms.end()) return it->second;
    float t1 = 0.f;
    FilterTile tl = tune(w, &t1);
    double r = 1.0;
    if (tl.split != 2 || !split_ok(w, sizeof(T), {P0.p, P1.p}, {w})) {
      tl.split = 1;
      DBuf<T> co((size_t)w), qa((size_t)n * w), qb((size_t)n * w), a0((size_t)n * w),
          a1((size_t)n * w), b0((size_t)n * w), b1((size_t)n * w);
      fill_block<T>(cx.stream, 1, w, co.p, w, T(0.25));
      for (T* q : {qa.p, qb.p, a0.p, a1.p, b0.p, b1.p}) fill_block<T>(cx.stream, n, w, q, w, T(0));
      const int reps = 8;
      Event e0, e1, e1s;
      float best = 1e30f;
      for (int sweep = 0; sweep < 2; ++sweep) {
        float ms = 0.f, ms2 = 0.f;
        {
          SplitScope scope(cx, true);
          for (int r2 = 0; r2 <= reps; ++r2) {
            if (r2 == 1) SMALGO_HIP(hipEventRecord(e0, cx.stream));  // after a warm step
            launch_clenshaw_step_cols<T>(cx.stream, *op.A, w, qa.p, w, co.p, a0.p, w, T(0.1),
                                         T(0.1), a1.p, w, T(1), a1.p, w, tl);
            launch_clenshaw_step_cols<T>(cx.side, *op.A, w, qb.p, w, co.p, b0.p, w, T(0.1),
                                         T(0.1), b1.p, w, T(1), b1.p, w, tl);
          }
          SMALGO_HIP(hipEventRecord(e1, cx.stream));
          SMALGO_HIP(hipEventRecord(e1s, cx.side));
        }
        SMALGO_HIP(hipEventSynchronize(e1));
        SMALGO_HIP(hipEventSynchronize(e1s));
        SMALGO_HIP(hipEventElapsedTime(&ms, e0, e1));
        SMALGO_HIP(hipEventElapsedTime(&ms2, e0, e1s));
        best = std::min(best, std::max(ms, ms2) / (float)reps);
      }
      r = std::min(1.0, (double)best / (2.0 * t1));
    }
    cx.tuned_ms[key] = (float)r;
    return r;
  }

  // Autotuner for the per-column kernel at width w: time a few steps of
  // every (vec, lanes per row, rows per team) tile on the real matrix.
  FilterTile tune(int w, float* ms_out) {
    const long long key = tune_key_mix({(long long)sizeof(T), (long long)w, op.mat_id, 1});
    // a tile forced through the environment is timed but neither read from
    // nor written to the cache, which matrices of the same structure share
    FilterTile forced{};
    const bool is_forced = env_tile("SMALGO_SLICE_TILE", &forced) && w % forced.vec == 0;
    auto it = cx.tuned.find(key);
    if (!is_forced && it != cx.tuned.end() && cx.tuned_ms.count(key)) {
      *ms_out = cx.tuned_ms[key];
      return it->second;
    }
    DBuf<T> co((size_t)w);
    fill_block<T>(cx.stream, 1, w, co.p, w, T(0.25));
    fill_block<T>(cx.stream, n, w, P0.p, w, T(0));
    fill_block<T>(cx.stream, n, w, P1.p, w, T(0));
    std::vector<FilterTile> cand;
    if (is_forced) {
      cand.push_back(forced);
    } else if (!cx.autotune) {
      const FilterTile d = choose_tile(sizeof(T), w, {P0.p, P1.p}, {w});
      cand.push_back(FilterTile{d.vec, d.lpr, 1});
    } else {
      // panels are sized to stay cache-resident: no prefetch form; panels
      // (pool blocks, ld = w) allow the split when w does
      cand = tile_candidates(w, sizeof(T), false, split_ok(w, sizeof(T), {P0.p, P1.p}, {w}),
                             [&](int v) { return w % v == 0; });
    }
    // The step the recurrence runs: Q, X1 and X2 distinct, the output in
    // place over X2 (out = cq∘Q + a1 A X1 + b1 X1 - X2); two sweeps of 4 steps.
    DBuf<T> qt((size_t)n * w);
    fill_block<T>(cx.stream, n, w, qt.p, w, T(0));
    std::vector<float> tmin;
    const size_t ib = time_tiles(
        cx, cand, cand.size() > 1 ? 2 : 1, 4,
        [&](const FilterTile& tl) {
          launch_step_cols<T>(cx, *op.A, w, qt.p, w, co.p, P0.p, w, T(0.1), T(0.1), P1.p, w, T(1),
                              P1.p, w, tl);
        },
        tmin,
        std::string("panel ") + (sizeof(T) == 4 ? "fp32" : "fp64") + " n=" + std::to_string(n) +
            " w=" + std::to_string(w));
    const FilterTile best = cand[ib];
    const float best_ms = tmin[ib];
    if (!is_forced) {
      cx.tuned[key] = best;
      cx.tuned_ms[key] = best_ms;
    }
    *ms_out = best_ms;
    return best;
  }

  // The Clenshaw recurrence of apply_poly_block! (polfed.jl:464-501) for
  // every panel, unfolded: B_{K+1} = c_K Q is materialised by a Q-only step
  // (no gather), then TMP = c_{k-1} Q + α A B1 + β B1 - B2 for k = K..2 with
  // TMP overwriting B2 in place once B2 exists, then the final
  // Y = c_0 Q + invd A B1 - c invd B1 - B2.  Only the Q coefficient differs
  // between columns (all slices share smin/smax), so one row of cq per step.
  // A column whose slice has degree K_i below the panel's K_p gets zero
  // coefficients for k > K_i: its state stays exactly zero until step K_i,
  // so it runs exactly the degree-K_i recurrence of its own slice.  Padding
  // columns have Q = 0 and cq = 0 and stay zero.
  void build_table(const std::vector<SliceCoeffs>& co, double smin, double smax) {
    const double c = 0.5 * (smin + smax);
    const double invd = 2.0 / (smax - smin);
    const double al = 2.0 * invd, be = -2.0 * c * invd;
    std::vector<T> h;
    steps.assign(pk.pan.size(), {});
    for (size_t p = 0; p < pk.pan.size(); ++p) {
      const int pw = pk.pan[p].w;
      const std::vector<int>& cs = pk.col_slice[p];
      const std::vector<int>& ci = pk.col_idx[p];
      auto cqrow = [&](int k) {
        const size_t r = h.size();
        h.resize(r + pw, T(0));
        for (int j = 0; j < pw; ++j)
          if (cs[j] >= 0) h[r + j] = (T)co[cs[j]].at(ci[j], k);
        return r;
      };
      const int K = pk.pan[p].K;
      // roles of B1 / B2: 0 zero, 2 / 3 buffer P0 / P1
      int b1 = 0, b2 = 0;
      if (K >= 1) {
        steps[p].push_back({0, 0, 2, cqrow(K), 0.0, 0.0, 0.0});  // B1 = c_K Q
        b1 = 2;
        for (int k = K; k >= 2; --k) {
          const int out = b2 ? b2 : (b1 == 2 ? 3 : 2);
          steps[p].push_back({b1, b2, out, cqrow(k - 1), al, be, b2 ? 1.0 : 0.0});
          b2 = b1;
          b1 = out;
        }
      }
      steps[p].push_back({b1, b2, 4, cqrow(0), b1 ? invd : 0.0, b1 ? -c * invd : 0.0,
                          b2 ? 1.0 : 0.0});
    }
    tab.alloc(std::max<size_t>(h.size(), 1));
    SMALGO_HIP(hipMemcpyAsync(tab.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice,
                              cx.stream));
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
  }

  // slice i's n x widths[i] block (row-major, ld) into / out of the panels
  void load(int i, const T* src, int lds) {
    for (const auto& sg : pk.segs)
      if (sg.slice == i) {
        const SlicePack::Panel& P = pk.pan[sg.panel];
        copy_block<T, T>(cx.stream, n, sg.w, src + sg.c0, lds, Qp.p + P.base + sg.off, P.w);
      }
  }
  void store(int i, T* dst, int ldd) {
    for (const auto& sg : pk.segs)
      if (sg.slice == i) {
        const SlicePack::Panel& P = pk.pan[sg.panel];
        copy_block<T, T>(cx.stream, n, sg.w, Yp.p + P.base + sg.off, P.w, dst + sg.c0, ldd);
      }
  }

  void launch(int p, const Step& st, bool side = false) {
    const SlicePack::Panel& P = pk.pan[p];
    const int pw = P.w;
    T* Qn = Qp.p + P.base;
    T* Yn = Yp.p + P.base;
    auto ptr = [&](int role) -> T* {
      switch (role) {
        case 1: return Qn;
        case 2: return side ? P0s.p : P0.p;
        case 3: return side ? P1s.p : P1.p;
        case 4: return Yn;
        default: return nullptr;
      }
    };
    if (side) {  // paired chains: the side chain's panel, one launch on the side stream
      launch_clenshaw_step_cols<T>(cx.side, *op.A, pw, Qn, pw, tab.p + st.cq, ptr(st.x1), pw,
                                   (T)st.a1, (T)st.b1, ptr(st.x2), pw, (T)st.s2, ptr(st.out),
                                   pw, P.tile);
      return;
    }
    launch_step_cols<T>(cx, *op.A, pw, Qn, pw, tab.p + st.cq, ptr(st.x1), pw, (T)st.a1,
                        (T)st.b1, ptr(st.x2), pw, (T)st.s2, ptr(st.out), pw, P.tile);
  }

  bool any_split() const {
    for (const auto& P : pk.pan)
      if (P.tile.split == 2) return true;
    return false;
  }

  // Y-panels = p_i(Â) Q-panels for every slice.  With graph = true the
  // launches are captured once per plan into HIP graphs of at most
  // graph_nodes_max kernels (kernel nodes only) and replayed afterwards.
  void run(bool graph) {
    if (!graph) {
      SplitScope split(cx, any_split() || paired);
      each_launch([&](int p, const Step& st, bool side) { launch(p, st, side); });
      return;
    }
    if (execs.empty()) capture();
    for (auto e : execs) SMALGO_HIP(hipGraphLaunch(e, cx.stream));
  }

  void prepare() {
    if (execs.empty()) capture();
  }

  void capture() {
    struct L {
      int p;
      const Step* st;
      bool side;
    };
    std::vector<L> all;
    each_launch([&](int p, const Step& st, bool side) { all.push_back({p, &st, side}); });
    const bool two = any_split() || paired;
    for (size_t b = 0; b < all.size(); b += graph_nodes_max) {
      const size_t e = std::min(all.size(), b + (size_t)graph_nodes_max);
      hipGraph_t g = nullptr;
      if (two) side_stream(cx);
      SMALGO_HIP(hipStreamBeginCapture(cx.stream, hipStreamCaptureModeThreadLocal));
      try {
        // split panels and paired chains: the side stream joins the capture
        // through the fork event and leaves it through the join event (graph
        // edges), so each graph holds both streams' launches as two branches
        SplitScope split(cx, two);
        for (size_t k = b; k < e; ++k) launch(all[k].p, *all[k].st, all[k].side);
      } catch (...) {
        (void)hipStreamEndCapture(cx.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      SMALGO_HIP(hipStreamEndCapture(cx.stream, &g));
      graphs.push_back(g);
      hipGraphExec_t ex = nullptr;
      SMALGO_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      execs.push_back(ex);
    }
  }

  size_t launches() const {
    size_t t = 0;
    for (const auto& v : steps) t += v.size();
    return t;
  }
  // bytes one run moves by the model of a fused step (padding included)
  double bytes() const {
    double b = 0;
    for (size_t p = 0; p < pk.pan.size(); ++p)
      b += (double)steps[p].size() * ((double)op.A->nnz * (sizeof(T) + 4) + 4.0 * (n + 1) +
                                      4.0 * n * pk.pan[p].w * sizeof(T));
    return b;
  }
  int max_width() const {
    int w = 0;
    for (const auto& P : pk.pan) w = std::max(w, P.w);
    return w;
  }
};

// ---------------------------------------------------------------------------
// POLFED building blocks
// ---------------------------------------------------------------------------
// Block basis of one interval: V = [Q_1 … Q_m] row-major n x ldV (block j at
// columns [j·s, j·s + r_j)), with the reduced-matrix blocks A_j, B_j.
template <typename T>
struct Basis {
  DBuf<T> V;
  int ldV = 0, s = 0, mmax = 0;
  std::vector<int> ranks;
  std::vector<std::vector<double>> Ablk, Bblk;  // column-major r_j x r_j, r_{j+1} x r_j
  DBuf<double> Adev, Bprev;                     // s x s device blocks
  // every block's A_j (r_j x r_j) and R_j (r_j x r_j, B_j = its first r_{j+1}
  // rows), column-major at s·s stride: read by the host once, after the basis
  DBuf<double> Aall, Rall;
  std::vector<int> rnext;  // r_{j+1} of each stored R_j
  int r_cur = 0, r_prev = 0;
  bool done = false;
};

struct StepWork {
  OrthWork ow;
  DBuf<double> gw;
  std::vector<double> R;
};

// orth_block_polfed!(randn(n,s); rtol) (polfed.jl:1133-1134) into block 0
template <typename T>
static void basis_start(Ctx& cx, Dense& dn, OrthWork& ow, Basis<T>& B, int n, const double* Q1host,
                        uint64_t seed, double rtol) {
  const int s = B.s, lds = round_up(s, 4);
  DBuf<double> q64((size_t)n * lds), t64((size_t)n * lds), o64((size_t)n * lds);
  if (Q1host) {
    DBuf<double> cm((size_t)n * s);
    SMALGO_HIP(hipMemcpyAsync(cm.p, Q1host, sizeof(double) * (size_t)n * s, hipMemcpyHostToDevice,
                              cx.stream));
    cm_to_rm<double, double>(cx.stream, n, s, cm.p, n, q64.p, lds);
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
  } else {
    gaussian_block<double>(cx.stream, n, s, seed, q64.p, lds);
  }
  B.r_cur = cholqr<double>(cx, dn, ow, n, s, q64.p, lds, o64.p, lds, t64.p, lds, rtol, nullptr);
  copy_block<double, T>(cx.stream, n, B.r_cur, o64.p, lds, B.V.p, B.ldV);
  if (B.r_cur == 0) B.done = true;
}

template <typename T>
static void basis_alloc(Basis<T>& B, int n, int s, int mmax) {
  B.s = s;
  B.mmax = mmax;
  B.ldV = round_up(mmax * s, 4);
  B.V.alloc((size_t)n * B.ldV);
  B.Adev.alloc((size_t)s * s);
  B.Bprev.alloc((size_t)s * s);
  B.Aall.alloc((size_t)mmax * s * s);
  B.Rall.alloc((size_t)mmax * s * s);
}

// The reduced-matrix blocks to the host (one download after the basis is built)
template <typename T>
static void basis_download(Ctx& cx, Basis<T>& B) {
  const int nb = (int)B.ranks.size(), ss = B.s * B.s;
  std::vector<double> a((size_t)nb * ss), r((size_t)B.rnext.size() * ss);
  if (nb)
    SMALGO_HIP(hipMemcpyAsync(a.data(), B.Aall.p, sizeof(double) * a.size(),
                              hipMemcpyDeviceToHost, cx.stream));
  if (!r.empty())
    SMALGO_HIP(hipMemcpyAsync(r.data(), B.Rall.p, sizeof(double) * r.size(),
                              hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  B.Ablk.assign(nb, {});
  for (int j = 0; j < nb; ++j) {
    const int rj = B.ranks[j];
    B.Ablk[j].assign(a.begin() + (size_t)j * ss, a.begin() + (size_t)j * ss + (size_t)rj * rj);
  }
  B.Bblk.assign(B.rnext.size(), {});
  for (size_t j = 0; j < B.rnext.size(); ++j) {
    const int rj = B.ranks[j], rn = B.rnext[j];
    const double* R = r.data() + j * ss;  // r_j x r_j, ld r_j
    std::vector<double> Bj((size_t)rn * rj);
    for (int c = 0; c < rj; ++c)
      for (int i = 0; i < rn; ++i) Bj[(size_t)c * rn + i] = R[(size_t)c * rj + i];
    B.Bblk[j] = std::move(Bj);
  }
}

// One block-Lanczos step after the filter W = p(Â) Q_j (polfed.jl:1066-1097):
// W -= Q_{j-1} B_{j-1}ᵀ; A_j = sym(Q_jᵀ W); W -= Q_j A_j; reorth_full!;
// [Q_{j+1}, B_j] = block_qr_factor(W).  W (n x r, ld ldw) is overwritten; tmp
// is an n x (ld ldt) scratch block.  Marks the basis done on rank collapse or
// after block mmax.
template <typename T>
static void lanczos_step(Ctx& cx, Dense& dn, StepWork& sw, const smalgo_polfed_params& p, int n,
                         Basis<T>& B, int j, T* W, int ldw, T* tmp, int ldt, DBuf<T>& Cbuf,
                         double& gemm_flops) {
  const int s = B.s, ldV = B.ldV, r = B.r_cur, r_prev = B.r_prev;
  T* Qj = B.V.p + (size_t)j * s;
  if (sw.gw.n < gemm_tn_work_elems(n, (j + 1) * s, r) || sw.gw.n < gemm_tn_work_elems(n, r, r)) {
    size_t g = 0;
    for (int k = 1; k <= B.mmax; ++k) g = std::max(g, gemm_tn_work_elems(n, k * s, s));
    sw.gw.ensure(g);
  }
  // W -= Q_{j-1} B_{j-1}ᵀ   (polfed.jl:1073-1075)
  if (j > 0 && r_prev > 0) {
    const T* Qp = B.V.p + (size_t)(j - 1) * s;
    gemm_nn<T, double, T>(cx.stream, n, r_prev, r, -1.0, Qp, ldV, B.Bprev.p, s, false, 1.0, W, ldw);
    gemm_flops += 2.0 * n * r_prev * r;
  }
  // A_j = sym(Q_jᵀ W)   (polfed.jl:1078)
  gemm_tn<T, T, double>(cx.stream, n, r, r, 1.0, Qj, ldV, W, ldw, 0.0, B.Adev.p, r, sw.gw.p,
                        sw.gw.n);
  symmetrize(cx.stream, r, B.Adev.p, r);
  gemm_flops += 2.0 * n * r * r;
  // A_j stays on the device; the host reads all blocks after the basis
  SMALGO_HIP(hipMemcpyAsync(B.Aall.p + (size_t)j * s * s, B.Adev.p, sizeof(double) * r * r,
                            hipMemcpyDeviceToDevice, cx.stream));
  B.ranks.push_back(r);
  if (j == B.mmax - 1) {  // Q_{mmax+1} and B_mmax never enter T or the Ritz vectors
    B.done = true;
    return;
  }
  // W -= Q_j A_j   (polfed.jl:1082)
  gemm_nn<T, double, T>(cx.stream, n, r, r, -1.0, Qj, ldV, B.Adev.p, r, false, 1.0, W, ldw);
  gemm_flops += 2.0 * n * r * r;
  // reorth_full!(W, Qhist, Q): against every stored block, then once more
  // against Q_j (polfed.jl:623-631; Qhist already holds Q_j, 1067)
  const int mall = (j + 1) * s;
  const int ldc = round_up(s, 4);
  Cbuf.ensure((size_t)ldV * ldc);
  for (int pass = 0; pass < p.reorth_passes; ++pass) {
    gemm_tn<T, T, T>(cx.stream, n, mall, r, 1.0, B.V.p, ldV, W, ldw, 0.0, Cbuf.p, ldc, sw.gw.p,
                     sw.gw.n);
    gemm_nn<T, T, T>(cx.stream, n, mall, r, -1.0, B.V.p, ldV, Cbuf.p, ldc, false, 1.0, W, ldw);
    gemm_tn<T, T, T>(cx.stream, n, r, r, 1.0, Qj, ldV, W, ldw, 0.0, Cbuf.p, ldc, sw.gw.p, sw.gw.n);
    gemm_nn<T, T, T>(cx.stream, n, r, r, -1.0, Qj, ldV, Cbuf.p, ldc, false, 1.0, W, ldw);
    gemm_flops += 4.0 * n * mall * r + 4.0 * n * r * r;
  }
  // [Q_{j+1}, B_j] = block_qr_factor(W; rtol)   (polfed.jl:1090): CholQR2 on the
  // device, R_j (r x r) straight into its slot of Rall
  T* Qn = B.V.p + (size_t)(j + 1) * s;
  double* Rj = B.Rall.p + (size_t)j * s * s;
  int rn = cholqr_fast<T>(cx, dn, sw.ow, n, r, W, ldw, Qn, ldV, tmp, ldt, p.rtol, Rj);
  if (rn < 0) {  // a Cholesky factorization failed: the shifted path, R from the host
    rn = cholqr<T>(cx, dn, sw.ow, n, r, W, ldw, Qn, ldV, tmp, ldt, p.rtol, &sw.R);
    if (rn > 0)
      SMALGO_HIP(hipMemcpyAsync(Rj, sw.R.data(), sizeof(double) * r * r, hipMemcpyHostToDevice,
                                cx.stream));
  }
  gemm_flops += 8.0 * n * r * r;
  if (rn == 0) {  // rank collapse (polfed.jl:1091)
    B.done = true;
    return;
  }
  if (rn < s) fill_block<T>(cx.stream, n, s - rn, Qn + rn, ldV, T(0));
  // B_j = R[1:rn, 1:r]; the device copy the next step reads as Bᵀ: element (k, i) at k*s + i
  b_block(cx.stream, s, r, rn, Rj, B.Bprev.p);
  B.rnext.push_back(rn);
  B.r_prev = r;
  B.r_cur = rn;
}

// p(x) - p(a) = (x - a) q(x): Chebyshev coefficients of q (degree K-1) by
// the backward recurrence of Clenshaw's algorithm at x = a (x T_0 = T_1,
// x T_k = (T_{k+1} + T_{k-1})/2 give q_{m-1} = 2 c_m + 2 a q_m - q_{m+1},
// q_0 = c_1 + a q_1 - q_2 / 2); returns p(a) = c_0 + a q_0 - q_1 / 2.
static double cheb_divide(const std::vector<double>& c, double a, std::vector<double>& q) {
  const int K = (int)c.size() - 1;
  q.assign(K + 2, 0.0);
  for (int m = K; m >= 2; --m) q[m - 1] = 2.0 * c[m] + 2.0 * a * q[m] - q[m + 1];
  if (K >= 1) q[0] = c[1] + a * q[1] - 0.5 * q[2];
  const double pa = c[0] + a * q[0] - 0.5 * q[1];
  q.resize(std::max(K, 0));
  return pa;
}

struct ResidualFilterWork {
  DBuf<double> Rd, cqd, pd;
  DBuf<float> Rf, Wf;
  double host_s = 0;  // host time spent on coefficients and plans
};

// Y = p(Â) Z for unit columns z_i with Rayleigh quotients θ_i, written as
//   p(Â) z_i = p(θ̂_i) z_i + q_i(Â) r̂_i,   r̂_i = (A z_i - θ_i z_i) / d,
// q_i = (p - p(θ̂_i)) / (x - θ̂_i), θ̂ = (θ - c) / d.  The residual r̂_i is
// formed in FP64 (one SpMM); the degree-(K-1) part runs in FP32 with a
// per-column coefficient row.  Its rounding is relative to |r̂_i|, which is
// already small for Ritz vectors, so Y keeps FP64 accuracy at FP32 bytes.
static void residual_filter(Ctx& cx, const Op& op, int k, const double* Z, int ldz,
                            const std::vector<double>& th, const std::vector<double>& co,
                            double smin, double smax, double* Y, int ldy, ResidualFilterWork& w,
                            int kmain = -1, const std::vector<double>* cog = nullptr) {
  const int n = op.n;
  if (k == 0) return;
  const auto h0 = std::chrono::steady_clock::now();
  const double c = 0.5 * (smin + smax), d = 0.5 * (smax - smin);
  // columns [kmain, k) (guard vectors) may take their own, lower-degree window
  if (kmain < 0 || kmain > k || cog == nullptr) kmain = k;
  const int K = std::max((int)co.size() - 1, cog && kmain < k ? (int)cog->size() - 1 : 0);
  std::vector<std::vector<double>> q(kmain), qg(k - kmain);
  std::vector<double> pa(k), mth(k);
  for (int i = 0; i < k; ++i) {
    const double a = std::min(1.0, std::max(-1.0, (th[i] - c) / d));
    pa[i] = i < kmain ? cheb_divide(co, a, q[i]) : cheb_divide(*cog, a, qg[i - kmain]);
    mth[i] = -th[i] / d;
  }
  const int ld = round_up(k, 4);
  w.Rd.ensure((size_t)n * ld);
  w.Rf.ensure((size_t)n * ld);
  w.Wf.ensure((size_t)n * ld);
  w.cqd.ensure(ld);
  w.pd.ensure(ld);
  SMALGO_HIP(hipMemcpyAsync(w.cqd.p, mth.data(), sizeof(double) * k, hipMemcpyHostToDevice,
                            cx.stream));
  SMALGO_HIP(hipMemcpyAsync(w.pd.p, pa.data(), sizeof(double) * k, hipMemcpyHostToDevice,
                            cx.stream));
  FilterTile t = choose_tile(sizeof(double), k, {Z, w.Rd.p, Y}, {ldz, ld, ldy});
  t.rt = 1;
  if (K >= 1) {
    // r̂ = (A Z - Z diag θ) / d, then rounded to FP32
    launch_clenshaw_step_cols<double>(cx.stream, *op.A, k, Z, ldz, w.cqd.p, Z, ldz, 1.0 / d, 0.0,
                                      nullptr, 0, 0.0, w.Rd.p, ld, t);
    copy_block<double, float>(cx.stream, n, k, w.Rd.p, ld, w.Rf.p, ld);
    // W = q_i(Â) r̂_i: one coefficient vector per column; the guard columns,
    // when they take a lower degree, form a second slice with its own panels
    SliceFilter<float> sf(cx, op);
    std::vector<SliceCoeffs> cs{SliceCoeffs{nullptr, &q}};
    std::vector<int> widths{kmain};
    if (kmain < k) {
      cs.push_back(SliceCoeffs{nullptr, &qg});
      widths.push_back(k - kmain);
    }
    if (kmain == 0) {
      cs.erase(cs.begin());
      widths.erase(widths.begin());
    }
    sf.plan(widths, cs, smin, smax, cache_panel_width<float>(op));
    w.host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
    for (int i = 0, off = 0; i < (int)widths.size(); off += widths[i], ++i)
      sf.load(i, w.Rf.p + off, ld);
    sf.run(false);
    for (int i = 0, off = 0; i < (int)widths.size(); off += widths[i], ++i)
      sf.store(i, w.Wf.p + off, ld);
    copy_block<float, double>(cx.stream, n, k, w.Wf.p, ld, Y, ldy);
  } else {
    fill_block<double>(cx.stream, n, k, Y, ldy, 0.0);
  }
  // Y = p(θ̂) ∘ Z + W   (in place on W: out = cq∘Q - s2 X2 with s2 = -1)
  launch_clenshaw_step_cols<double>(cx.stream, *op.A, k, Z, ldz, w.pd.p, nullptr, 0, 0.0, 0.0, Y,
                                    ldy, -1.0, Y, ldy, t);
}

static bool refine_fp64_env() {
  const char* e = std::getenv("SMALGO_REFINE_FP64");
  return e && e[0] == '1';
}

struct ExtractStats {
  long long refine_steps = 0;
  int locked = 0;
  double gemm_flops = 0;
  int refine_iterations = 0;
  double refine_host_s = 0;
};

// extract_pairs (polfed.jl:892-984) + the return rule of polfed (:1148-1157)
// on a finished basis; the mixed-precision branch adds the FP64 refinement.
template <typename T>
static Result* polfed_extract(Ctx& cx, Dense& dn, PhaseTimer& tm, const Op& op, const Basis<T>& Bs,
                              int nev, int degree, double smin, double smax,
                              const smalgo_polfed_params& p, ExtractStats& xs) {
  const int n = op.n, s = Bs.s, ldV = Bs.ldV;
  const std::vector<int>& ranks = Bs.ranks;
  const auto& Ablk = Bs.Ablk;
  const auto& Bblk = Bs.Bblk;
  const T* Vp = Bs.V.p;
  const int nb = (int)ranks.size();
  long long refine_steps = 0;
  int locked_total = 0;
  double gemm_flops = 0;
  size_t hx = tm.begin("extract");
  std::unique_ptr<Result> out(new Result());
  out->n = n;
  out->device = cx.device;
  std::vector<int> rowmap;
  for (int j = 0; j < nb; ++j)
    for (int t = 0; t < ranks[j]; ++t) rowmap.push_back(j * s + t);
  const int M = (int)rowmap.size();
  std::vector<double> lam_f, res_f;
  DBuf<double> Xfin;
  int ldxf = 4;
  std::vector<int> final_cols;
  if (M > 0) {
    std::vector<double> Th((size_t)M * M, 0.0);
    int o = 0;
    for (int j = 0; j < nb; ++j) {
      const int r = ranks[j];
      for (int c = 0; c < r; ++c)
        for (int i = 0; i < r; ++i) Th[(size_t)(o + c) * M + o + i] = Ablk[j][(size_t)c * r + i];
      if (j < nb - 1 && j < (int)Bblk.size()) {
        const int rn = ranks[j + 1];
        for (int c = 0; c < r; ++c)
          for (int i = 0; i < rn; ++i) {
            const double v = Bblk[j][(size_t)c * rn + i];
            Th[(size_t)(o + c) * M + o + r + i] = v;  // T[J, I] = B
            Th[(size_t)(o + r + i) * M + o + c] = v;  // T[I, J] = Bᵀ
          }
      }
      o += r;
    }
    DBuf<double> Td((size_t)M * M), Dd(M), Ed(M);
    SMALGO_HIP(hipMemcpyAsync(Td.p, Th.data(), sizeof(double) * Th.size(), hipMemcpyHostToDevice,
                              cx.stream));
    dn.syevd(M, Td.p, M, Dd.p, Ed.p, true);
    std::vector<double> th(M);
    SMALGO_HIP(hipMemcpyAsync(th.data(), Dd.p, sizeof(double) * M, hipMemcpyDeviceToHost, cx.stream));
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
    const int m = std::min(M, std::max(p.overextract * nev, nev + p.extra_keep));
    std::vector<int> order = stable_argsort(th, true);
    order.resize(m);
    const int Mpad = nb * s;
    const int ldu = round_up(m, 4);
    DBuf<double> U((size_t)Mpad * ldu);
    fill_block<double>(cx.stream, Mpad, ldu, U.p, ldu, 0.0);
    DBuf<int> dsel = upload_idx(cx, order), drow = upload_idx(cx, rowmap);
    gather_cols_colmajor(cx.stream, M, m, Td.p, M, dsel.p, drow.p, U.p, ldu);
    // X = V U  (build_X!, polfed.jl:767-776): one MFMA GEMM instead of 64-wide chunks
    DBuf<double> X((size_t)n * ldu);
    gemm_nn<T, double, double>(cx.stream, n, Mpad, m, 1.0, Vp, ldV, U.p, ldu, false, 0.0, X.p, ldu);
    gemm_flops += 2.0 * n * Mpad * m;
    RayleighWork rw;
    std::vector<double> lam, res;
    rayleigh(cx, op, n, m, X.p, ldu, rw, lam, res);
    std::vector<double> lam_keep, res_keep;
    std::vector<int> idx_keep;
    for (int jj = 0; jj < m; ++jj)
      if (p.emin <= lam[jj] && lam[jj] <= p.emax) {
        lam_keep.push_back(lam[jj]);
        res_keep.push_back(res[jj]);
        idx_keep.push_back(jj);
      }
    if (!idx_keep.empty()) {
      const int npre = std::min((int)res_keep.size(), nev + p.extra_keep);
      std::vector<int> ps0 = stable_argsort(res_keep);
      ps0.resize(npre);
      std::vector<int> idxc(npre);
      std::vector<double> lamc(npre), resc(npre);
      for (int i = 0; i < npre; ++i) {
        idxc[i] = idx_keep[ps0[i]];
        lamc[i] = lam_keep[ps0[i]];
        resc[i] = res_keep[ps0[i]];
      }
      int ldc = round_up(npre, 4);
      DBuf<double> Xc;
      int kc = npre;
      const bool mixed = std::is_same<T, float>::value;
      const double anorm = std::max(std::fabs(smin), std::fabs(smax));
      if (p.refine && mixed && p.refine_maxiter > 0) {
        // ---- FP64 refinement of an FP32-built basis (no reference counterpart):
        // filtered subspace iteration on the candidates plus guard vectors (the
        // next Ritz vectors by θ), Rayleigh–Ritz in FP64, locking of pairs whose
        // residual is below refine_tol·‖A‖ (‖A‖ = max(|smin|,|smax|)) or, by
        // default, below the caller's tol, the bound of the return rule.
        const size_t hr = tm.begin("refine");
        const int guard = p.refine_guard >= 0 ? p.refine_guard : nev;
        std::vector<int> zc(idxc);
        std::vector<char> used(m, 0);
        for (int v : idxc) used[v] = 1;
        for (int jj = 0; jj < m && (int)zc.size() < npre + guard; ++jj)
          if (!used[jj]) zc.push_back(jj);
        int kz = (int)zc.size();
        int ldz = round_up(kz, 4);
        DBuf<double> Z((size_t)n * ldz), Lk((size_t)n * ldz), Ytmp((size_t)n * ldz);
        DBuf<double> Zb0, Zb1;  // ping-pong blocks of the FP64 refinement filter only
        DBuf<int> dz = upload_idx(cx, zc);
        gather_cols_rm<double>(cx.stream, n, kz, X.p, ldu, dz.p, Z.p, ldz);
        const int Kr = p.refine_degree >= 0 ? p.refine_degree : degree;
        const std::vector<double> cr =
            build_coeffs(Kr, smin, smax, p.sigma, p.emin, p.emax, p.window_type, p.jackson != 0);
        const double tol_abs = p.refine_tol >= 0 ? p.refine_tol * anorm : p.tol;
        // The guard vectors only widen the Rayleigh–Ritz subspace by directions just
        // outside the window, which the FP32 basis already carries; in the first
        // iteration they take a window of refine_guard_degree (default Kr/50).
        std::vector<double> cg;
        {
          const int Kg = p.refine_guard_degree >= 0 ? p.refine_guard_degree
                                                    : std::max(1, (Kr + 49) / 50);
          if (Kg < Kr)
            cg = build_coeffs(Kg, smin, smax, p.sigma, p.emin, p.emax, p.window_type,
                              p.jackson != 0);
        }
        RRWork rr;
        int nlock = 0;
        std::vector<double> lam_lock, res_lock;
        std::vector<double> lam_act, res_act;
        OrthWork ow2;
        DBuf<double> Cl, glw;
        DBuf<double> pb2;
        if (op.is_callback()) pb2.alloc((size_t)n * ldz);
        ResidualFilterWork rfw;
        const bool fp64_filter = refine_fp64_env() || op.is_callback();
        if (fp64_filter) {
          Zb0.alloc((size_t)n * ldz);
          Zb1.alloc((size_t)n * ldz);
        }
        for (int it = 0; it < p.refine_maxiter && kz > 0; ++it) {
          // Y = p_r(Â) Z: in residual form with an FP32 filter (SMALGO_REFINE_FP64=1:
          // the FP64 filter on Z itself)
          if (fp64_filter) {
            apply_filter<double>(cx, op, kz, Z.p, ldz, cr, smin, smax, Ytmp.p, ldz, Zb0.p, Zb1.p,
                                 ldz, pb2.p);
          } else {
            std::vector<double> th(kz);
            for (int i = 0; i < kz; ++i) th[i] = it == 0 ? lam[zc[i]] : lam_act[i];
            const bool split = it == 0 && !cg.empty() && npre < kz;
            residual_filter(cx, op, kz, Z.p, ldz, th, cr, smin, smax, Ytmp.p, ldz, rfw,
                            split ? npre : -1, split ? &cg : nullptr);
          }
          if (it == 0 && !cg.empty() && npre < kz && !fp64_filter)
            refine_steps += (long long)Kr * npre + (long long)((int)cg.size() - 1) * (kz - npre);
          else
            refine_steps += (long long)Kr * kz;
          ++xs.refine_iterations;
          // project out the locked vectors
          if (nlock > 0) {
            Cl.ensure((size_t)nlock * ldz);
            glw.ensure(gemm_tn_work_elems(n, nlock, kz));
            gemm_tn<double, double, double>(cx.stream, n, nlock, kz, 1.0, Lk.p, ldz, Ytmp.p, ldz,
                                            0.0, Cl.p, ldz, glw.p, glw.n);
            gemm_nn<double, double, double>(cx.stream, n, nlock, kz, -1.0, Lk.p, ldz, Cl.p, ldz,
                                            false, 1.0, Ytmp.p, ldz);
          }
          DBuf<double> Yr;
          int ldyr = 4;
          const int kq = refine_rr(cx, dn, op, n, kz, Ytmp.p, ldz, rr, Yr, ldyr, lam_act, res_act);
          if (kq == 0) break;
          // lock converged inside pairs; keep the unconverged inside pairs and
          // only the guard vectors nearest the interval (at most half as many
          // as unconverged pairs, at least 16) for the next iteration
          std::vector<int> keep, lock, outside;
          int inside_unconv = 0;
          for (int i = 0; i < kq; ++i) {
            const bool inside = p.emin <= lam_act[i] && lam_act[i] <= p.emax;
            if (inside && res_act[i] < tol_abs && nlock + (int)lock.size() < ldz)
              lock.push_back(i);
            else if (inside) {
              keep.push_back(i);
              ++inside_unconv;
            } else {
              outside.push_back(i);
            }
          }
          {
            auto dist = [&](int i) {
              return lam_act[i] < p.emin ? p.emin - lam_act[i] : lam_act[i] - p.emax;
            };
            std::stable_sort(outside.begin(), outside.end(),
                             [&](int a, int b) { return dist(a) < dist(b); });
            const int ng = std::min((int)outside.size(), std::max(16, inside_unconv / 2));
            keep.insert(keep.end(), outside.begin(), outside.begin() + ng);
            std::sort(keep.begin(), keep.end());
          }
          if (!lock.empty()) {
            DBuf<int> dl = upload_idx(cx, lock);
            gather_cols_rm<double>(cx.stream, n, (int)lock.size(), Yr.p, ldyr, dl.p, Lk.p + nlock,
                                   ldz);
            for (int i : lock) {
              lam_lock.push_back(lam_act[i]);
              res_lock.push_back(res_act[i]);
            }
            nlock += (int)lock.size();
          }
          std::vector<double> la2, re2;
          for (int i : keep) {
            la2.push_back(lam_act[i]);
            re2.push_back(res_act[i]);
          }
          lam_act.swap(la2);
          res_act.swap(re2);
          kz = (int)keep.size();
          if (kz > 0) {
            DBuf<int> dk = upload_idx(cx, keep);
            gather_cols_rm<double>(cx.stream, n, kz, Yr.p, ldyr, dk.p, Z.p, ldz);
          }
          SMALGO_HIP(hipStreamSynchronize(cx.stream));
          if (inside_unconv == 0) break;
        }
        locked_total = nlock;
        xs.refine_host_s = rfw.host_s;
        // candidate set = locked ∪ active; the reference's post-refine selection follows
        kc = nlock + kz;
        ldc = round_up(std::max(kc, 1), 4);
        Xc.alloc((size_t)n * ldc);
        if (nlock) copy_block<double, double>(cx.stream, n, nlock, Lk.p, ldz, Xc.p, ldc);
        if (kz) copy_block<double, double>(cx.stream, n, kz, Z.p, ldz, Xc.p + nlock, ldc);
        lamc = lam_lock;
        resc = res_lock;
        lamc.insert(lamc.end(), lam_act.begin(), lam_act.end());
        resc.insert(resc.end(), res_act.begin(), res_act.end());
        // sort the candidates by λ like refine_original_rr's output
        std::vector<int> pl = stable_argsort(lamc);
        std::vector<double> l2(kc), r2(kc);
        for (int i = 0; i < kc; ++i) {
          l2[i] = lamc[pl[i]];
          r2[i] = resc[pl[i]];
        }
        DBuf<int> dpl = upload_idx(cx, pl);
        DBuf<double> Xs((size_t)n * ldc);
        gather_cols_rm<double>(cx.stream, n, kc, Xc.p, ldc, dpl.p, Xs.p, ldc);
        Xc = std::move(Xs);
        lamc.swap(l2);
        resc.swap(r2);
        SMALGO_HIP(hipStreamSynchronize(cx.stream));
        tm.end(hr);
      } else {
        Xc.alloc((size_t)n * ldc);
        DBuf<int> dc = upload_idx(cx, idxc);
        gather_cols_rm<double>(cx.stream, n, npre, X.p, ldu, dc.p, Xc.p, ldc);
        if (p.refine) {
          const size_t hr = tm.begin("refine");
          RRWork rr;
          DBuf<double> Yr;
          int ldyr = 4;
          kc = refine_rr(cx, dn, op, n, npre, Xc.p, ldc, rr, Yr, ldyr, lamc, resc);
          Xc = std::move(Yr);
          ldc = ldyr;
          tm.end(hr);
        }
      }
      // inside2 filter (polfed.jl:971-975) — the unrefined branch already holds
      // only inside pairs
      std::vector<int> ins;
      for (int i = 0; i < kc; ++i)
        if (p.emin <= lamc[i] && lamc[i] <= p.emax) ins.push_back(i);
      if (!ins.empty()) {
        std::vector<double> l2, r2;
        for (int i : ins) {
          l2.push_back(lamc[i]);
          r2.push_back(resc[i]);
        }
        const int nkeep = std::min((int)r2.size(), nev + p.extra_keep);
        std::vector<int> ps = stable_argsort(r2);
        ps.resize(nkeep);
        std::vector<double> lf(nkeep), rf(nkeep);
        std::vector<int> cf(nkeep);
        for (int i = 0; i < nkeep; ++i) {
          lf[i] = l2[ps[i]];
          rf[i] = r2[ps[i]];
          cf[i] = ins[ps[i]];
        }
        std::vector<int> pe = stable_argsort(lf);
        for (int i = 0; i < nkeep; ++i) {
          lam_f.push_back(lf[pe[i]]);
          res_f.push_back(rf[pe[i]]);
          final_cols.push_back(cf[pe[i]]);
        }
        Xfin = std::move(Xc);
        ldxf = ldc;
      }
    }
  }
  tm.end(hx);
  out->line = progress_line(nb, lam_f, res_f);
  if (p.show_progress) {
    std::printf("%s\n", out->line.c_str());
    std::fflush(stdout);
  }
  // ---- return rule (polfed.jl:1148-1157)
  std::vector<int> pick;
  if (!res_f.empty() && (int)res_f.size() >= nev) {
    std::vector<double> rs(res_f);
    std::sort(rs.begin(), rs.end());
    if (rs[nev - 1] < p.tol)
      for (int i = 0; i < nev; ++i) pick.push_back(i);
  }
  if (pick.empty() && !res_f.empty()) {
    pick = stable_argsort(res_f);
    pick.resize(std::min(nev, (int)res_f.size()));
  }
  const int kout = (int)pick.size();
  std::vector<int> cols(kout);
  for (int i = 0; i < kout; ++i) {
    out->lam.push_back(lam_f[pick[i]]);
    out->res.push_back(res_f[pick[i]]);
    cols[i] = final_cols[pick[i]];
  }
  if (kout > 0) {
    out->alloc_X(sizeof(double) * (size_t)n * kout);
    DBuf<int> dc = upload_idx(cx, cols);
    rm_to_cm<double>(cx.stream, n, kout, Xfin.p, ldxf, dc.p, out->Xdev, n);
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
  }
  out->stats["block_steps"] = nb;
  out->stats["basis_dim"] = M;
  out->stats["iterations"] = nb;
  xs.refine_steps += refine_steps;
  xs.locked += locked_total;
  xs.gemm_flops += gemm_flops;
  return out.release();
}

// ---------------------------------------------------------------------------
// POLFED (one interval)
// ---------------------------------------------------------------------------
template <typename T>
static Result* polfed_run(Ctx& cx, const Op& op, int nev, int degree, const smalgo_polfed_params& p,
                          const double* Q1host) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const int n = op.n;
  if (nev < 1) throw Error(-1, "nev must be positive");
  if (degree < 0) throw Error(-1, "degree must be nonnegative");
  if (p.s < 1) throw Error(-1, "block size s must be positive");
  if (p.s > n) throw Error(-1, "block size s must not exceed n");
  double smin = p.smin, smax = p.smax;
  if (!p.has_bounds) {
    if (op.is_callback()) throw Error(-1, "Matrix-free POLFED requires spectral bounds smin and smax");
    smin = op.A->gersh_min;
    smax = op.A->gersh_max;
  }
  if (!(smax > smin)) throw Error(-1, "need smin < smax");
  const std::vector<double> coeffs =
      build_coeffs(degree, smin, smax, p.sigma, p.emin, p.emax, p.window_type, p.jackson != 0);
  const int s = p.s;
  const int mmax = p.mmax > 0 ? p.mmax : (int)std::ceil(2.8 * nev / (double)s);
  const int lds = round_up(s, 4);
  PhaseTimer tm(cx.stream);
  Dense dn(cx);
  StepWork sw;
  Basis<T> B;
  basis_alloc<T>(B, n, s, mmax);
  DBuf<T> W((size_t)n * lds), b0((size_t)n * lds), b1((size_t)n * lds), Cbuf;
  DBuf<double> pbuf;
  if (op.is_callback()) pbuf.alloc((size_t)n * lds);
  fill_block<T>(cx.stream, n, B.ldV, B.V.p, B.ldV, T(0));
  basis_start<T>(cx, dn, sw.ow, B, n, Q1host, p.seed, p.rtol);
  FilterTile basis_tile{};
  const long long split0 = cx.split_steps;
  int basis_lps = 1;
  long long filter_steps = 0;
  double gemm_flops = 0;
  // algorithmic bytes of the basis-build filter, per fused step and launch:
  // nnz*(w+4) + 4(n+1) for the matrix (read once per launch: once per panel
  // when a wide block is split) + 4*n*b*w for the four block streams
  double bytes = 0;
  const double mat_bytes = (double)(op.A ? op.A->nnz : 0) * (sizeof(T) + 4) + 4.0 * (n + 1);
  for (int j = 0; j < mmax && !B.done; ++j) {
    T* Qj = B.V.p + (size_t)j * s;
    size_t hf = tm.begin("filter");
    apply_filter<T>(cx, op, B.r_cur, Qj, B.ldV, coeffs, smin, smax, W.p, lds, b0.p, b1.p, lds,
                    pbuf.p);
    tm.end(hf);
    if (j == 0) {
      basis_tile = cx.last_tile;
      basis_lps = cx.last_launches_per_step;
    }
    filter_steps += degree;  // K fused steps (the all-zero first step is skipped)
    bytes += (double)degree * (cx.last_launches_per_step * mat_bytes +
                               4.0 * n * B.r_cur * sizeof(T));
    size_t ho = tm.begin("orth");
    lanczos_step<T>(cx, dn, sw, p, n, B, j, W.p, lds, b0.p, lds, Cbuf, gemm_flops);
    tm.end(ho);
  }
  ExtractStats xs;
  basis_download<T>(cx, B);
  std::unique_ptr<Result> out(polfed_extract<T>(cx, dn, tm, op, B, nev, degree, smin, smax, p, xs));
  auto ph = tm.resolve();
  const int nb = (int)B.ranks.size();
  out->stats["t_total"] = std::chrono::duration<double>(clk::now() - t0).count();
  out->stats["t_filter"] = ph["filter"];
  out->stats["t_orth"] = ph["orth"];
  out->stats["t_extract"] = ph["extract"];
  out->stats["t_refine"] = ph["refine"];
  out->stats["filter_steps"] = (double)filter_steps;
  out->stats["filter_launches"] = (double)filter_steps * basis_lps;
  out->stats["refine_filter_steps"] = (double)xs.refine_steps;
  out->stats["refine_iterations"] = xs.refine_iterations;
  out->stats["t_refine_host"] = xs.refine_host_s;
  out->stats["locked"] = xs.locked;
  out->stats["gemm_flops"] = gemm_flops + xs.gemm_flops;
  out->stats["smin"] = smin;
  out->stats["smax"] = smax;
  out->stats["tile_vec"] = basis_tile.vec;
  out->stats["tile_lpr"] = basis_tile.lpr;
  out->stats["tile_rt"] = basis_tile.rt;
  out->stats["tile_lds"] = basis_tile.lds;
  out->stats["tile_split"] = basis_tile.split;
  out->stats["filter_split_steps"] = (double)(cx.split_steps - split0);
  out->stats["spmm_bytes"] = bytes;
  return out.release();
}

Result* polfed_solve(Ctx& cx, const Op& op, int nev, int degree, const smalgo_polfed_params& p,
                     const double* Q1) {
  if (p.precision == SMALGO_PREC_FP32_REFINE64) {
    if (op.is_callback()) throw Error(-1, "matrix-free POLFED runs in FP64 only");
    return polfed_run<float>(cx, op, nev, degree, p, Q1);
  }
  return polfed_run<double>(cx, op, nev, degree, p, Q1);
}

// ---------------------------------------------------------------------------
// Chebyshev filtered subspace iteration (FP64)
// ---------------------------------------------------------------------------
Result* chebyshev_solve(Ctx& cx, const Op& op, double emin, double emax, const smalgo_cheb_params& p,
                        const double* Q0, bool matrix_free) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const int n = op.n;
  if (!(emin < emax)) throw Error(-1, "need emin<emax");
  if (!(p.m0 > 0)) throw Error(-1, "m0 must be positive");
  if (!(p.degree >= 0)) throw Error(-1, "degree must be nonnegative");
  const int m0 = std::min(p.m0, n);
  double smin = p.smin, smax = p.smax;
  if (!p.has_bounds) {
    if (op.is_callback())
      throw Error(-1, "gershgorin_bounds has no matrix-free method; provide smin and smax");
    smin = op.A->gersh_min;
    smax = op.A->gersh_max;
  }
  std::unique_ptr<Result> out(new Result());
  out->n = n;
  out->device = cx.device;
  if (!(emin >= smin && emax <= smax)) {
    out->stats["warn_interval"] = 1.0;
    if (!matrix_free)
      std::fprintf(stderr,
                   "Warning: target interval not contained in supplied spectral bounds; filter "
                   "quality may degrade\n");
  }
  const double c = 0.5 * (smin + smax), d = 0.5 * (smax - smin);
  const double a = (emin - c) / d, b = (emax - c) / d;
  if (!(-1.0 <= a && a < b && b <= 1.0)) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "mapped interval [%s,%s] not inside [-1,1]%s",
                  format_julia_float(a).c_str(), format_julia_float(b).c_str(),
                  matrix_free ? "" : "; provide better spectral bounds");
    throw Error(-1, buf);
  }
  const double sh = (p.sigma - c) / d;
  std::vector<double> coeffs;
  switch (p.window_type) {
    case SMALGO_WINDOW_STEP: coeffs = cheb_window_coeffs(a, b, p.degree, p.jackson != 0); break;
    case SMALGO_WINDOW_DELTA: coeffs = cheb_delta_coeffs(sh, p.degree, p.jackson != 0); break;
    case SMALGO_WINDOW_SMOOTH:
      coeffs = cheb_smooth_window_coeffs(a, b, p.degree, p.jackson != 0, 4096, 0.02);
      break;
    default:
      throw Error(-1, "Unknown window type, possible are :step, :delta, and :smooth");
  }
  // rows start on 128-byte lines, so the filter may split the block by columns
  const int ld = round_up(m0, 16);
  PhaseTimer tm(cx.stream);
  Dense dn(cx);
  OrthWork ow;
  DBuf<double> Q((size_t)n * ld), Y((size_t)n * ld), B0((size_t)n * ld), B1((size_t)n * ld),
      X((size_t)n * ld), AX((size_t)n * ld), Aq((size_t)m0 * m0), wv(m0), ev(m0), lamd(m0),
      resd(m0), red(colred_work_elems(n, m0)), gw(gemm_tn_work_elems(n, m0, m0));
  DBuf<double> pbuf;
  if (op.is_callback()) pbuf.alloc((size_t)n * ld);
  if (Q0) {
    DBuf<double> cm((size_t)n * m0);
    SMALGO_HIP(hipMemcpyAsync(cm.p, Q0, sizeof(double) * (size_t)n * m0, hipMemcpyHostToDevice,
                              cx.stream));
    cm_to_rm<double, double>(cx.stream, n, m0, cm.p, n, Q.p, ld);
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
  } else {
    gaussian_block<double>(cx.stream, n, m0, p.seed, Q.p, ld);
  }
  std::vector<double> lam(m0, 0.0), res(m0, std::numeric_limits<double>::infinity());
  bool prev_ok = false;
  int its = 0;
  long long filter_steps = 0;
  FilterTile cheb_tile{};
  const long long split0 = cx.split_steps;
  int cheb_lps = 1;
  for (int it = 1; it <= p.maxiter; ++it) {
    its = it;
    size_t hf = tm.begin("filter");
    apply_filter<double>(cx, op, m0, Q.p, ld, coeffs, smin, smax, Y.p, ld, B0.p, B1.p, ld, pbuf.p);
    cheb_tile = cx.last_tile;
    cheb_lps = cx.last_launches_per_step;
    tm.end(hf);
    filter_steps += p.degree;
    size_t ho = tm.begin("orth");
    // Q .= Y; thin_orth!(Q)  — all min(n,m0) columns kept (:733-734).  With
    // rtol = 0 a zero rank means the Cholesky QR gave up (a zero or non-finite
    // block, e.g. NaN from a host operator): Q would be stale, so stop here.
    if (cholqr<double>(cx, dn, ow, n, m0, Y.p, ld, Q.p, ld, B0.p, ld, 0.0, nullptr, true) == 0)
      throw Error(-4, "thin_orth: Cholesky QR of the filtered block failed (zero or "
                      "non-finite block at iteration " + std::to_string(it) + ")");
    tm.end(ho);
    size_t hr = tm.begin("rr");
    apply_A<double>(cx, op, m0, Q.p, ld, AX.p, ld);                       // R = A Q
    gemm_tn<double, double, double>(cx.stream, n, m0, m0, 1.0, Q.p, ld, AX.p, ld, 0.0, Aq.p, m0,
                                    gw.p, gw.n);                          // Aq = QᵀR
    // eigen!(Symmetric(Aq)) reads the upper triangle of Aq; Aq is row-major
    // here, i.e. the lower triangle of its column-major view.
    dn.syevd(m0, Aq.p, m0, wv.p, ev.p, false);
    gemm_nn<double, double, double>(cx.stream, n, m0, m0, 1.0, Q.p, ld, Aq.p, m0, true, 0.0, X.p,
                                    ld);                                  // X = Q Xq
    apply_A<double>(cx, op, m0, X.p, ld, AX.p, ld);
    col_resid<double>(cx.stream, n, m0, X.p, ld, AX.p, ld, wv.p, resd.p, red.p, red.n);
    SMALGO_HIP(hipMemcpyAsync(lam.data(), wv.p, sizeof(double) * m0, hipMemcpyDeviceToHost,
                              cx.stream));
    SMALGO_HIP(hipMemcpyAsync(res.data(), resd.p, sizeof(double) * m0, hipMemcpyDeviceToHost,
                              cx.stream));
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
    tm.end(hr);
    int cnt = 0;
    double maxres = -1;
    for (int j = 0; j < m0; ++j)
      if (lam[j] >= emin && lam[j] <= emax && res[j] < p.res_gate) {
        ++cnt;
        maxres = std::max(maxres, res[j]);
      }
    if (cnt == 0) maxres = std::numeric_limits<double>::infinity();
    if (p.debug) {
      if (matrix_free) {
        // the matrix-free form adds quantile(res[inside], 0.9) (:826): Julia's
        // default quantile, linear between order statistics at h = (k-1)·0.9
        std::vector<double> ri;
        for (int j = 0; j < m0; ++j)
          if (lam[j] >= emin && lam[j] <= emax && res[j] < p.res_gate) ri.push_back(res[j]);
        double r90 = std::numeric_limits<double>::infinity();
        if (!ri.empty()) {
          std::sort(ri.begin(), ri.end());
          const double h = (double)(ri.size() - 1) * 0.9;
          const size_t lo = (size_t)std::floor(h);
          const size_t hi = std::min(lo + 1, ri.size() - 1);
          r90 = ri[lo] + (h - (double)lo) * (ri[hi] - ri[lo]);
        }
        std::printf("it=%d inside=%d maxres=%s r90=%s\n", it, cnt, printf_e3(maxres).c_str(),
                    printf_e3(r90).c_str());
      } else {
        std::printf("it=%d inside=%d maxres=%s\n", it, cnt, printf_e3(maxres).c_str());
      }
      std::fflush(stdout);
    }
    if (matrix_free) {  // :827-836
      const bool ok = cnt > 0 && maxres < p.tol;
      if (p.two_hit) {
        if (ok && (it == 1 || prev_ok)) {
          if (p.debug) {
            std::printf("converged\n");
            std::fflush(stdout);
          }
          break;
        }
        prev_ok = ok;
      } else if (ok) {
        break;
      }
    }
  }
  // selection (:750-756)
  std::vector<int> idx;
  std::vector<double> ls;
  const double gate = std::max(p.res_gate, 10 * p.tol);
  for (int j = 0; j < m0; ++j)
    if (lam[j] >= emin && lam[j] <= emax && res[j] < gate) {
      idx.push_back(j);
      ls.push_back(lam[j]);
    }
  std::vector<int> pp = stable_argsort(ls);
  std::vector<int> cols;
  for (int i : pp) {
    cols.push_back(idx[i]);
    out->lam.push_back(lam[idx[i]]);
    out->res.push_back(res[idx[i]]);
  }
  if (!cols.empty()) {
    out->alloc_X(sizeof(double) * (size_t)n * cols.size());
    DBuf<int> dc = upload_idx(cx, cols);
    rm_to_cm<double>(cx.stream, n, (int)cols.size(), X.p, ld, dc.p, out->Xdev, n);
  }
  auto ph = tm.resolve();
  out->stats["t_total"] = std::chrono::duration<double>(clk::now() - t0).count();
  out->stats["t_filter"] = ph["filter"];
  out->stats["t_orth"] = ph["orth"];
  out->stats["t_extract"] = ph["rr"];
  out->stats["iterations"] = its;
  out->stats["filter_steps"] = (double)filter_steps;
  out->stats["filter_launches"] = op.is_callback() ? (double)filter_steps
                                                   : (double)filter_steps * cheb_lps;
  out->stats["tile_vec"] = cheb_tile.vec;
  out->stats["tile_lpr"] = cheb_tile.lpr;
  out->stats["tile_rt"] = cheb_tile.rt;
  out->stats["tile_lds"] = cheb_tile.lds;
  out->stats["tile_split"] = cheb_tile.split;
  out->stats["filter_split_steps"] = (double)(cx.split_steps - split0);
  out->stats["smin"] = smin;
  out->stats["smax"] = smax;
  // the matrix is read once per launch (once per panel of a wide block)
  out->stats["spmm_bytes"] = (double)filter_steps *
                             ((op.is_callback() ? 1 : cheb_lps) *
                                  ((double)(op.A ? op.A->nnz : 0) * 12.0 + 4.0 * (n + 1)) +
                              4.0 * n * m0 * 8.0);
  char buf[64];
  std::snprintf(buf, sizeof buf, "it=%d inside=%zu", its, out->lam.size());
  out->line = buf;
  return out.release();
}

static int slice_pw_env() {
  const char* e = std::getenv("SMALGO_SLICE_PW");
  return e ? std::atoi(e) : 0;
}

template <typename T>
static void apply_poly_slices_t(Ctx& cx, const Op& op, const std::vector<int>& widths,
                                const std::vector<std::vector<double>>& coeffs, double smin,
                                double smax, int pw_force, bool graph, int replays, const double* Q,
                                int ldq, double* Y, int ldy, double* stats) {
  const int n = op.n, ns = (int)widths.size();
  int wtot = 0;
  for (int w : widths) wtot += w;
  std::vector<SliceCoeffs> co(ns);
  for (int i = 0; i < ns; ++i) co[i].shared = &coeffs[i];
  SliceFilter<T> sf(cx, op);
  sf.plan(widths, co, smin, smax, pw_force);
  const int ld = round_up(std::max(wtot, 1), 4);
  DBuf<T> blk((size_t)n * ld);
  Event e0, e1;
  double ms_tot = 0;
  for (int r = 0; r < replays; ++r) {
    // the host round trip between replays is ordinary eager work
    upload_cm<T>(cx, n, wtot, Q + (size_t)r * ldq * wtot, ldq, blk.p, ld);
    int off = 0;
    for (int i = 0; i < ns; ++i) {
      sf.load(i, blk.p + off, ld);
      off += widths[i];
    }
    if (graph) sf.prepare();  // capture and instantiate outside the timed region
    SMALGO_HIP(hipEventRecord(e0, cx.stream));
    sf.run(graph);
    SMALGO_HIP(hipEventRecord(e1, cx.stream));
    off = 0;
    for (int i = 0; i < ns; ++i) {
      sf.store(i, blk.p + off, ld);
      off += widths[i];
    }
    download_cm<T>(cx, n, wtot, blk.p, ld, Y + (size_t)r * ldy * wtot, ldy);
    float ms = 0;
    SMALGO_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms_tot += ms;
  }
  if (stats) {
    stats[0] = sf.max_width();
    stats[1] = (double)sf.pk.pan.size();
    stats[2] = (double)sf.launches();
    stats[3] = replays > 0 ? ms_tot / replays : 0.0;
    stats[4] = sf.bytes();
    stats[5] = (double)sf.execs.size();
  }
}

void apply_poly_slices_host(Ctx& cx, const Op& op, const std::vector<int>& widths,
                            const std::vector<std::vector<double>>& coeffs, double smin, double smax,
                            int precision, int pw_force, bool graph, int replays, const double* Q,
                            int ldq, double* Y, int ldy, double* stats) {
  if (op.is_callback()) throw Error(-1, "batched slices need a sparse matrix");
  if (widths.size() != coeffs.size() || widths.empty()) throw Error(-1, "bad slice list");
  for (int w : widths)
    if (w < 1) throw Error(-1, "slice widths must be positive");
  if (pw_force < 0 || pw_force % 4) throw Error(-1, "panel width must be a multiple of 4");
  if (precision == SMALGO_PREC_FP32_REFINE64)
    apply_poly_slices_t<float>(cx, op, widths, coeffs, smin, smax, pw_force, graph, replays, Q,
                               ldq, Y, ldy, stats);
  else
    apply_poly_slices_t<double>(cx, op, widths, coeffs, smin, smax, pw_force, graph, replays, Q,
                                ldq, Y, ldy, stats);
}

// ---------------------------------------------------------------------------
// POLFED over many slices at once.  Every slice runs the block Lanczos of
// polfed (polfed.jl:1128-1158) with its own interval, degree, block width and
// basis; the filters of all active slices at block step j share one packed
// launch sequence (captured in HIP graphs), then each slice does its own
// orthogonalisation and, at the end, its own extraction and return rule.
// ---------------------------------------------------------------------------
template <typename T>
static std::vector<Result*> polfed_slices_run(Ctx& cx, const Op& op,
                                              const std::vector<SliceSpec>& sl,
                                              const smalgo_polfed_params& p, bool graph) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const int n = op.n, ns = (int)sl.size();
  if (op.is_callback()) throw Error(-1, "batched slices need a sparse matrix");
  if (ns < 1) throw Error(-1, "need at least one slice");
  double smin = p.smin, smax = p.smax;
  if (!p.has_bounds) {
    smin = op.A->gersh_min;
    smax = op.A->gersh_max;
  }
  if (!(smax > smin)) throw Error(-1, "need smin < smax");
  std::vector<std::vector<double>> coeffs(ns);
  std::vector<int> mmax(ns);
  int smax_w = 1;
  for (int i = 0; i < ns; ++i) {
    const SliceSpec& q = sl[i];
    if (q.nev < 1) throw Error(-1, "nev must be positive");
    if (q.degree < 0) throw Error(-1, "degree must be nonnegative");
    if (q.s < 1 || q.s > n) throw Error(-1, "block size s must be in [1, n]");
    if (!(q.emax >= q.emin)) throw Error(-1, "slice interval must have emin <= emax");
    coeffs[i] = build_coeffs(q.degree, smin, smax, q.sigma, q.emin, q.emax, p.window_type,
                             p.jackson != 0);
    mmax[i] = q.mmax > 0 ? q.mmax : (int)std::ceil(2.8 * q.nev / (double)q.s);
    smax_w = std::max(smax_w, q.s);
  }
  PhaseTimer tm(cx.stream);
  Dense dn(cx);
  StepWork sw;
  std::vector<Basis<T>> B(ns);
  for (int i = 0; i < ns; ++i) {
    basis_alloc<T>(B[i], n, sl[i].s, mmax[i]);
    fill_block<T>(cx.stream, n, B[i].ldV, B[i].V.p, B[i].ldV, T(0));
    basis_start<T>(cx, dn, sw.ow, B[i], n, nullptr, p.seed + (uint64_t)i, p.rtol);
  }
  const int ldw = round_up(smax_w, 4);
  DBuf<T> W((size_t)n * ldw), tmp((size_t)n * ldw), Cbuf;
  SliceFilter<T> sf(cx, op);
  const int pw_force = slice_pw_env();
  std::vector<int> planned;  // active slices the current plan was built for
  std::vector<long long> fsteps(ns, 0);
  double gemm_flops = 0, filter_bytes = 0, filter_launches = 0;
  int replans = 0, max_launch_graphs = 0, paired_runs = 0;
  for (int j = 0;; ++j) {
    std::vector<int> act;
    for (int i = 0; i < ns; ++i)
      if (!B[i].done && j < mmax[i]) act.push_back(i);
    if (act.empty()) break;
    if (act != planned) {
      std::vector<int> w;
      std::vector<SliceCoeffs> co;
      for (int i : act) {
        w.push_back(sl[i].s);
        co.push_back(SliceCoeffs{&coeffs[i], nullptr});
      }
      sf.plan(w, co, smin, smax, pw_force);
      planned = act;
      ++replans;
    }
    size_t hf = tm.begin("filter");
    for (size_t k = 0; k < act.size(); ++k) {
      const int i = act[k];
      sf.load((int)k, B[i].V.p + (size_t)j * sl[i].s, B[i].ldV);
    }
    sf.run(graph);
    tm.end(hf);
    paired_runs += sf.paired ? 1 : 0;
    max_launch_graphs = std::max(max_launch_graphs, (int)sf.execs.size());
    filter_bytes += sf.bytes();
    filter_launches += (double)sf.launches();
    size_t ho = tm.begin("orth");
    for (size_t k = 0; k < act.size(); ++k) {
      const int i = act[k];
      sf.store((int)k, W.p, round_up(sl[i].s, 4));
      fsteps[i] += sl[i].degree;
      lanczos_step<T>(cx, dn, sw, p, n, B[i], j, W.p, round_up(sl[i].s, 4), tmp.p,
                      round_up(sl[i].s, 4), Cbuf, gemm_flops);
    }
    tm.end(ho);
    if (p.show_progress) {
      std::fprintf(stderr, "slices: block %d, %zu active, panel width %d, %d panels, %zu launches, %.1f s\n",
                  j + 1, act.size(), sf.max_width(), (int)sf.pk.pan.size(), sf.launches(),
                  std::chrono::duration<double>(clk::now() - t0).count());
      std::fflush(stderr);
    }
  }
  const int pw_used = sf.max_width();
  std::vector<Result*> out(ns, nullptr);
  std::map<std::string, double> ph_all;
  try {
    for (int i = 0; i < ns; ++i) {
      smalgo_polfed_params pp = p;
      pp.s = sl[i].s;
      pp.emin = sl[i].emin;
      pp.emax = sl[i].emax;
      pp.sigma = sl[i].sigma;
      pp.mmax = mmax[i];
      pp.seed = p.seed + (uint64_t)i;
      ExtractStats xs;
      basis_download<T>(cx, B[i]);
      out[i] = polfed_extract<T>(cx, dn, tm, op, B[i], sl[i].nev, sl[i].degree, smin, smax, pp, xs);
      out[i]->stats["filter_steps"] = (double)fsteps[i];
      out[i]->stats["refine_filter_steps"] = (double)xs.refine_steps;
      out[i]->stats["locked"] = xs.locked;
      out[i]->stats["gemm_flops"] = xs.gemm_flops;
    }
    ph_all = tm.resolve();
  } catch (...) {
    for (Result* r : out) delete r;
    throw;
  }
  auto& ph = ph_all;
  const double ttot = std::chrono::duration<double>(clk::now() - t0).count();
  for (int i = 0; i < ns; ++i) {
    Result* r = out[i];
    r->stats["t_total"] = ttot;  // the whole batch
    r->stats["t_filter"] = ph["filter"];
    r->stats["t_orth"] = ph["orth"];
    r->stats["t_extract"] = ph["extract"];
    r->stats["t_refine"] = ph["refine"];
    r->stats["smin"] = smin;
    r->stats["smax"] = smax;
    r->stats["slice_panel_width"] = pw_used;
    r->stats["slice_replans"] = replans;
    r->stats["slice_graphs"] = max_launch_graphs;
    r->stats["slice_filter_bytes"] = filter_bytes;
    r->stats["slice_launches"] = filter_launches;
    r->stats["slice_gemm_flops"] = gemm_flops;
    // tile of the widest panel of the last plan
    FilterTile wt{};
    int ww = -1;
    for (const auto& P : sf.pk.pan)
      if (P.w > ww) {
        ww = P.w;
        wt = P.tile;
      }
    r->stats["tile_vec"] = wt.vec;
    r->stats["tile_lpr"] = wt.lpr;
    r->stats["tile_rt"] = wt.rt;
    r->stats["tile_lds"] = wt.lds;
    r->stats["tile_split"] = wt.split;
    r->stats["slice_paired"] = paired_runs;  // block steps whose panels ran as two chains
  }
  return out;
}

std::vector<Result*> polfed_slices_solve(Ctx& cx, const Op& op, const std::vector<SliceSpec>& sl,
                                         const smalgo_polfed_params& p, bool graph) {
  if (p.precision == SMALGO_PREC_FP32_REFINE64) return polfed_slices_run<float>(cx, op, sl, p, graph);
  if (p.precision != SMALGO_PREC_FP64) throw Error(-1, "unknown precision mode");
  return polfed_slices_run<double>(cx, op, sl, p, graph);
}

}  // namespace smalgo
