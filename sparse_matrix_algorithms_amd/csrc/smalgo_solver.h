// Host drivers: context, operators, POLFED and Chebyshev subspace iteration.
#pragma once

#include <rocblas/rocblas.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/smalgo.h"
#include "smalgo_internal.h"

namespace smalgo {

// Device memory of the solvers comes from a per-thread, per-device cache of
// freed blocks: a solve's workspace is handed back to the next solve instead
// of going through hipFree/hipMalloc (each of which synchronises the device).
// Safe because every entry point finishes its work on its one stream before
// it returns, so a returned block is idle, and blocks never move between
// threads.  At most kPoolCacheBytes stay cached per thread and device.
void* pool_alloc(size_t bytes, int* device);
void pool_free(void* p, size_t bytes, int device);

// RAII device buffer
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  int dev = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), dev(o.dev) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; dev = o.dev; o.p = nullptr; o.n = 0; }
    return *this;
  }
  void alloc(size_t count) {
    release();
    if (count) p = static_cast<T*>(pool_alloc(count * sizeof(T), &dev));
    n = count;
  }
  void ensure(size_t count) { if (count > n) alloc(count); }
  void release() {
    if (p) pool_free(p, n * sizeof(T), dev);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

// A timing event that is destroyed on every exit path of its scope.
struct Event {
  hipEvent_t e = nullptr;
  Event() { SMALGO_HIP(hipEventCreate(&e)); }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  operator hipEvent_t() const { return e; }
};

// Pinned host staging for host-callback operators: two grow-only pinned
// blocks, a second stream for the host->device leg and per-chunk events.
struct Staging {
  double* x = nullptr;
  double* y = nullptr;
  size_t cap = 0;  // doubles per block
  hipStream_t up = nullptr;
  std::vector<hipEvent_t> ev;
  hipEvent_t done = nullptr;
  void ensure(size_t count, int chunks);
  ~Staging();
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::string err;
  std::map<long long, FilterTile> tuned;  // (precision, width, structure, alignment) -> tile
  bool autotune = true;
  FilterTile last_tile{};  // tile of the most recent filter
  int last_launches_per_step = 1;  // kernel launches per filter step (panels)
  std::map<long long, float> tuned_ms;  // per-step time of tuned per-column tiles
  Staging staging;                      // host-callback operator transfers
  // second stream for split filter steps (FilterTile::split = 2), created on
  // first use; fork_ev / join_ev order it after / before the main stream
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  long long split_steps = 0;  // filter steps issued as two launches (statistics)
  bool split_active = false;  // inside a fork/join pair of the two streams
  int split_h = -1;           // column split of the last launch in it (0: one launch)
};

struct Mat {
  Ctx* ctx = nullptr;
  DevCSR A;
  int symmetric = -1;  // computed at upload: 1 / 0
  long long id = 0;
  // fingerprint of the sparsity structure (n, row pointers, column indices):
  // a filter tile's speed depends on the structure only, so matrices with the
  // same structure share the autotuner's picks
  long long structure = 0;
};

// Operator: device CSR or host callback (Y <- A X on column-major host arrays)
struct Op {
  const DevCSR* A = nullptr;
  smalgo_matblock_fn fn = nullptr;
  void* user = nullptr;
  int n = 0;
  long long mat_id = 0;  // the matrix's structure fingerprint (tuning cache key)
  bool is_callback() const { return fn != nullptr; }
};

struct Result {
  int64_t n = 0;
  std::vector<double> lam, res;
  double* Xdev = nullptr;  // column-major n x k on the device (from pool_alloc)
  size_t Xbytes = 0;
  int device = 0;
  std::map<std::string, double> stats;
  std::string line;
  void alloc_X(size_t bytes) { Xdev = static_cast<double*>(pool_alloc(bytes, &device)); Xbytes = bytes; }
  ~Result() {
    if (Xdev) {
      int cur = 0;
      const bool sw = hipGetDevice(&cur) == hipSuccess && cur != device;
      if (sw) (void)hipSetDevice(device);
      pool_free(Xdev, Xbytes, device);
      if (sw) (void)hipSetDevice(cur);
    }
  }
};

// Small FP64 dense LAPACK-type steps on the device (rocSOLVER).
struct Dense {
  Ctx& cx;
  DBuf<int> info;
  explicit Dense(Ctx& c) : cx(c), info(1) {}
  int potrf_upper(int b, double* A, int lda);  // returns info (0 = success)
  // the same without waiting: info is written to dinfo on the device
  void potrf_upper_async(int b, double* A, int lda, int* dinfo);
  void trtri_upper(int b, double* A, int lda);
  void syevd(int n, double* A, int lda, double* w, double* work_e, bool upper);
};

// Timing of device phases with events on the context stream.
struct PhaseTimer {
  hipStream_t st;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev;
  std::map<std::string, double> open_;
  explicit PhaseTimer(hipStream_t s) : st(s) {}
  size_t begin(const std::string& name);
  void end(size_t h);
  std::map<std::string, double> resolve();  // seconds per phase (syncs)
  ~PhaseTimer();
};

FilterTile choose_tile(size_t elem, int b, std::initializer_list<const void*> ptrs,
                       std::initializer_list<int> lds);

// block Clenshaw filter Y = p(Â) Q (coefficients on the host)
template <typename T>
void apply_filter(Ctx& cx, const Op& op, int b, const T* Q, int ldq, const std::vector<double>& co,
                  double smin, double smax, T* Y, int ldy, T* buf0, T* buf1, int ldb,
                  double* pbuf /* callback staging, n x ldb doubles, may be null for CSR */);

// The same filter for a CSR operator without allocating: Y (row-major, ld
// ldy) and `scratch` (n x ldy doubles) are the ping-pong pair; coefficients
// co[0..K] on the host.
void apply_filter_noalloc(Ctx& cx, const Op& op, int b, const double* Q, int ldq, const double* co,
                          int K, double smin, double smax, double* Y, int ldy, double* scratch);

// Device scratch of the no-allocation entry point (smalgo_noalloc.hip): a
// static device array whose address is looked up once per device at context
// creation.  noalloc_scratch returns the calling device's address and its
// size in bytes, or throws if the lookup failed.
void noalloc_scratch_init(int device);
double* noalloc_scratch(int device, size_t* bytes);

// Y = A X (CSR or callback)
template <typename T>
void apply_A(Ctx& cx, const Op& op, int b, const T* X, int ldx, T* Y, int ldy);

Result* polfed_solve(Ctx& cx, const Op& op, int nev, int degree, const smalgo_polfed_params& p,
                     const double* Q1);
Result* chebyshev_solve(Ctx& cx, const Op& op, double emin, double emax, const smalgo_cheb_params& p,
                        const double* Q0, bool matrix_free);

// One-call versions of the reference's block kernels (host column-major arrays).
void orth_against_host(Ctx& cx, int n, int b, double* W, int ldw, int m, const double* Q, int ldq,
                       int precision);
void block_qr_host(Ctx& cx, int n, int b, const double* W, int ldw, double rtol, int precision,
                   double* Q, int ldq, double* R, int ldr, int* rank);
void build_X_host(Ctx& cx, int n, int M, int k, const double* V, int ldv, const double* U, int ldu,
                  double* X, int ldx);
void rayleigh_host(Ctx& cx, const Op& op, int k, const double* X, int ldx, double* lam, double* res);
void residuals_host(Ctx& cx, const Op& op, int k, const double* X, int ldx, const double* lam,
                    double* res);
Result* refine_rr_host(Ctx& cx, const Op& op, int k, const double* X, int ldx);
Result* subspace_rr_host(Ctx& cx, const Op& op, int k, const double* Y, int ldy, double* Qh,
                         int ldqh);

// one slice of a batched POLFED run: interval, window centre, pairs wanted,
// block width, filter degree, basis blocks (<= 0: ceil(2.8 nev / s))
struct SliceSpec {
  double emin, emax, sigma;
  int nev, s, degree, mmax;
};
std::vector<Result*> polfed_slices_solve(Ctx& cx, const Op& op, const std::vector<SliceSpec>& sl,
                                         const smalgo_polfed_params& p, bool graph);
// the packed multi-slice filter alone (host column-major blocks; replays
// consecutive input blocks through one plan); stats[6]: widest panel, panels,
// launches per run, ms per run, modelled bytes per run, graphs
void apply_poly_slices_host(Ctx& cx, const Op& op, const std::vector<int>& widths,
                            const std::vector<std::vector<double>>& coeffs, double smin, double smax,
                            int precision, int pw_force, bool graph, int replays, const double* Q,
                            int ldq, double* Y, int ldy, double* stats);

void kpm_moments(Ctx& cx, const Op& op, int nmom, int R, uint64_t seed, double smin, double smax,
                 double* mu);

std::string format_julia_float(double v);
std::string progress_line(int m, const std::vector<double>& lam, const std::vector<double>& res);

}  // namespace smalgo
