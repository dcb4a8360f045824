/*
 * smalgo.h — C ABI of the MI355X (gfx950) solve-free interior eigensolver.
 *
 * Drop-in boundary for the two solve-free drivers of
 * MathiasTensor/sparse_matrix_algorithms:
 *   chebyshev_poly_real_symm_sparse  (chebyshev_subspace_iteration.jl:667-757, :771-848)
 *   polfed                           (polfed.jl:1128-1158, :1183-1186)
 * and of the helpers those drivers call (bounds, coefficients, the block
 * Clenshaw filter).  A Julia module `ccall`s these symbols with the
 * reference's own argument meaning; a Python ctypes mirror does the same.
 *
 * Conventions
 *   - Host arrays are column-major (Julia layout), FP64, owned by the caller.
 *   - The library owns all device memory (per context / matrix / result).
 *   - Every call returns an int status: SMALGO_OK (0) or a negative code;
 *     smalgo_last_error(ctx) gives the message of the last failure on ctx.
 *   - One context per host thread; callbacks run on the calling thread only.
 */
#ifndef SMALGO_H
#define SMALGO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMALGO_OK 0
#define SMALGO_ERR_ARG -1       /* ArgumentError / error(...) in the reference */
#define SMALGO_ERR_HIP -2       /* HIP runtime failure */
#define SMALGO_ERR_CONFIG -3    /* unsupported tile / configuration */
#define SMALGO_ERR_INTERNAL -4  /* workspace or library failure */
#define SMALGO_ERR_CALLBACK -5  /* user matblock callback returned nonzero */

/* window_type (polfed.jl:404-414 `window_type` Symbol) */
#define SMALGO_WINDOW_STEP 0
#define SMALGO_WINDOW_DELTA 1
#define SMALGO_WINDOW_SMOOTH 2

/* precision of the filter / basis (FP32 basis + FP64 refinement; FP64 = reference) */
#define SMALGO_PREC_FP64 0
#define SMALGO_PREC_FP32_REFINE64 1

typedef struct smalgo_ctx smalgo_ctx;
typedef struct smalgo_mat smalgo_mat;
typedef struct smalgo_result smalgo_result;

/* User block operator Y <- A*X on host arrays (n x b, column-major, leading
 * dimensions ldx/ldy).  Replaces MatrixFreeRealSymOp.matblock!
 * (polfed.jl:148-165, chebyshev_subspace_iteration.jl:151-167) and, through a
 * column loop, matvec!.  Returns 0 on success. */
typedef int (*smalgo_matblock_fn)(void* user, int64_t n, int64_t b, const double* X, int64_t ldx,
                                  double* Y, int64_t ldy);

/* ---- context ---------------------------------------------------------- */
/* One GPU, one HIP stream (default priority) for all work of the context; a
 * second compute stream is created on the first filter step the autotuner
 * runs as two concurrent launches over the column halves, and every entry
 * point joins it back into the first before it returns.  Calls on one
 * context must not overlap; contexts are independent. */
int smalgo_ctx_create(int device, smalgo_ctx** out);
void smalgo_ctx_destroy(smalgo_ctx* ctx);
const char* smalgo_last_error(const smalgo_ctx* ctx);
/* Library version string, e.g. "smalgo 0.1 gfx950". */
const char* smalgo_version(void);
/* v as Julia's print(::Float64) writes it, the form of every float in the
 * progress line of polfed.jl:991-1000 (println of Float64 values): shortest
 * round-trip digits, fixed notation for 1e-4 <= |v| < 1e6, otherwise
 * d.ddde<exp> with no zero-padded exponent ("1.0e-5", "3.1117078311317e-14").
 * buf receives a NUL-terminated string of at most len-1 characters.  No GPU. */
int smalgo_julia_float_string(double v, char* buf, int64_t len);

/* ---- sparse matrix ------------------------------------------------------
 * Upload a compressed sparse matrix.  ptr has n+1 entries, idx/val have nnz.
 * For the symmetric matrices of this path, Julia's SparseMatrixCSC
 * (colptr,rowval,nzval) of A is the CSR of A, so both layouts are accepted
 * as CSR.  index_base is 1 for Julia arrays, 0 for C/Python.
 * Replaces the SparseMatrixCSC{Float64,Int} operand of apply_A_block!
 * (polfed.jl:196-199) and mul!(AX,A,B1) (chebyshev_subspace_iteration.jl:574). */
int smalgo_csr_upload(smalgo_ctx* ctx, int64_t n, const int64_t* ptr, const int64_t* idx,
                      const double* val, int index_base, smalgo_mat** out);
void smalgo_mat_free(smalgo_mat* mat);
int64_t smalgo_mat_n(const smalgo_mat* mat);
int64_t smalgo_mat_nnz(const smalgo_mat* mat);
/* exact symmetry test of the uploaded pattern and values
 * (issymmetric, chebyshev_subspace_iteration.jl:669) */
int smalgo_mat_is_symmetric(const smalgo_mat* mat, int* is_symmetric);

/* gershgorin_bounds (polfed.jl:254-266, chebyshev_subspace_iteration.jl:351-367) */
int smalgo_gershgorin_bounds(const smalgo_mat* mat, double* smin, double* smax);

/* ---- filter coefficients (host) -----------------------------------------
 * jackson_kernel (polfed.jl:275-284): g[0..K] (K<=0 gives g[0]=1). */
int smalgo_jackson_kernel(int K, double* g);
/* cheb_delta_coeffs (polfed.jl:299-313): c[0..K] */
int smalgo_cheb_delta_coeffs(double sigma_hat, int K, int jackson, double* c);
/* cheb_window_coeffs (polfed.jl:329-347): c[0..K] */
int smalgo_cheb_window_coeffs(double a, double b, int K, int jackson, double* c);
/* cheb_smooth_window_coeffs (polfed.jl:364-388): c[0..K], nsamp<=0 -> 4096, delta<=0 -> 0.02 */
int smalgo_cheb_smooth_window_coeffs(double a, double b, int K, int jackson, int nsamp,
                                     double delta, double* c);
/* _build_coeffs (polfed.jl:404-414): maps sigma/emin/emax by c=(smin+smax)/2,
 * d=(smax-smin)/2 and dispatches on window_type. */
int smalgo_build_coeffs(int degree, double smin, double smax, double sigma, double emin,
                        double emax, int window_type, int jackson, double* c);

/* ---- block Clenshaw filter ----------------------------------------------
 * Y = p(Â) Q with Â=(A-cI)/d, p = sum_k coeffs[k] T_k
 * (apply_poly_block!, polfed.jl:464-501 / chebyshev_subspace_iteration.jl:554-591).
 * Q, Y: host n x b column-major.  precision: SMALGO_PREC_FP64 or
 * SMALGO_PREC_FP32_REFINE64 (filter in FP32, result returned as FP64). */
int smalgo_apply_poly_block(smalgo_ctx* ctx, const smalgo_mat* A, int64_t b, const double* Q,
                            int64_t ldq, const double* coeffs, int K, double smin, double smax,
                            int precision, double* Y, int64_t ldy);
/* No counterpart in the reference.  The filter of smalgo_apply_poly_block
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): it calls no
 * allocator.  dQ and dY are DEVICE arrays, FP64, row-major n x b (element
 * (i,c) at i*ld + c, ld >= b), not overlapping; coeffs[0..K] is a host
 * array.  The recurrence's second block lives in a fixed per-device scratch
 * of 64 MiB: n*ldy*8 bytes above that is SMALGO_ERR_ARG.  Calls that overlap
 * on one GPU share that scratch and race.  A sparse matrix only; returns
 * after the filter has finished on the context stream. */
int smalgo_apply_poly_block_noalloc(smalgo_ctx* ctx, const smalgo_mat* A, int64_t b,
                                    const double* dQ, int64_t ldq, const double* coeffs, int K,
                                    double smin, double smax, double* dY, int64_t ldy);
/* Same with a host operator (matrix-free form, polfed.jl:508-545). */
int smalgo_apply_poly_block_mf(smalgo_ctx* ctx, int64_t n, smalgo_matblock_fn fn, void* user,
                               int64_t b, const double* Q, int64_t ldq, const double* coeffs,
                               int K, double smin, double smax, double* Y, int64_t ldy);

/* ---- block kernels of the reference, one call each ----------------------
 * All blocks are host arrays, n x b column-major; precision as above (the
 * FP32 mode rounds the inputs to FP32 and computes the products in FP32).
 *
 * orth_against!(W, Q): W <- W - Q (Qᵀ W)   (polfed.jl:605-609) */
int smalgo_orth_against(smalgo_ctx* ctx, int64_t n, int64_t b, double* W, int64_t ldw, int64_t m,
                        const double* Q, int64_t ldq, int precision);
/* block_qr_factor(W; rtol) / orth_block_polfed!(W; rtol) (polfed.jl:644-685,
 * :561-594): W = Q R by CholQR2 (shifted CholQR3 when the Gram matrix is not
 * numerically positive definite); rank = #{|R_ii| > rtol·max|R_ii|}.  Q gets
 * all b columns (the caller keeps the first rank), R is b x b upper
 * triangular, column-major (ldr >= b). */
int smalgo_block_qr(smalgo_ctx* ctx, int64_t n, int64_t b, const double* W, int64_t ldw,
                    double rtol, int precision, double* Q, int64_t ldq, double* R, int64_t ldr,
                    int* rank);
/* build_X!(X, Qhist, U): X = V U with V = [Q_1 … Q_m] as one n x M block
 * (polfed.jl:767-776); U is M x k column-major. */
int smalgo_build_X(smalgo_ctx* ctx, int64_t n, int64_t M, int64_t k, const double* V, int64_t ldv,
                   const double* U, int64_t ldu, double* X, int64_t ldx);
/* rayleigh!(λ, res, AX, A, X) (polfed.jl:791-830): λ_j = x_jᵀ A x_j,
 * res_j = ‖A x_j − λ_j x_j‖₂.  Exactly one of A or fn. */
int smalgo_rayleigh(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                    int64_t n, int64_t k, const double* X, int64_t ldx, double* lam, double* res);
/* residuals!(res, A, X, λ, AX) (chebyshev_subspace_iteration.jl:311-344) */
int smalgo_residuals(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                     int64_t n, int64_t k, const double* X, int64_t ldx, const double* lam,
                     double* res);
/* refine_original_rr(A, X) (polfed.jl:837-880): result holds (λ, Y, res)
 * sorted by λ; read it with smalgo_result_*. */
int smalgo_refine_original_rr(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn,
                              void* user, int64_t n, int64_t k, const double* X, int64_t ldx,
                              smalgo_result** out);
/* The Rayleigh–Ritz tail of the reference's subspace iterations, real
 * symmetric case: chebyshev_subspace_iteration.jl:733-748 and the same steps
 * in feast.jl:540-553 and zoloterov.jl:1698-1706.  Q = thin_orth(Y) (all k
 * columns), Aq = Qᵀ(A Q), eigen(Symmetric(Aq)), X = Q·Xq,
 * res_j = ‖A x_j − λ_j x_j‖₂.  Y: host n x k column-major (the filtered or
 * contour-integrated block).  The result holds all k pairs in ascending λ
 * (eigen's order); the caller applies its own inside/res_gate selection.  Q
 * (optional, host n x k) receives the orthonormal basis for the next
 * iteration's solves. */
int smalgo_subspace_rr(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                       int64_t n, int64_t k, const double* Y, int64_t ldy, double* Q,
                       int64_t ldq, smalgo_result** out);

/* ---- kernel polynomial method (no reference counterpart; sizes intervals and
 * feeds the learned filter parameters) ----------------------------------
 * mu[k] = (1/(R·n)) Σ_r r_rᵀ T_k(Â) r_r, k = 0..nmom-1, with Â=(A-cI)/d,
 * c=(smin+smax)/2, d=(smax-smin)/2, over R Gaussian probes drawn from seed
 * (splitmix64/Box-Muller).  Uses the same fused SpMM kernel as the filter and
 * the doubling identities μ_2k = 2⟨T_k,T_k⟩-μ_0, μ_2k+1 = 2⟨T_k+1,T_k⟩-μ_1. */
int smalgo_kpm_moments(smalgo_ctx* ctx, const smalgo_mat* A, int nmom, int R, uint64_t seed,
                       double smin, double smax, double* mu);

/* ---- measurement helper (no reference counterpart) -----------------------
 * Runs `reps` device-to-device copies of `bytes` bytes with a plain 16-byte
 * vector grid-stride copy kernel (stream_copy_kernel) on the context stream and returns the mean seconds per
 * copy; used to calibrate HBM byte counters (FETCH_SIZE / WRITE_SIZE). */
int smalgo_copy_bench(smalgo_ctx* ctx, int64_t bytes, int reps, double* seconds);

/* ---- learned filter parameters (no reference counterpart) -----------------
 * FP32 two-layer perceptron nin -> nh (ReLU) -> nout with hand-written HIP
 * forward, backward (mean squared error) and Adam kernels.  Parameters are
 * one FP32 vector [W1 (nin x nh, row-major), b1 (nh), W2 (nh x nout), b2 (nout)].
 * X: B x nin row-major, T/Y: B x nout row-major (host FP32 arrays). */
typedef struct smalgo_mlp smalgo_mlp;
int smalgo_mlp_create(smalgo_ctx* ctx, int nin, int nh, int nout, uint64_t seed, smalgo_mlp** out);
void smalgo_mlp_free(smalgo_mlp* m);
int64_t smalgo_mlp_num_params(const smalgo_mlp* m);
int smalgo_mlp_get_params(const smalgo_mlp* m, float* params);
int smalgo_mlp_set_params(smalgo_mlp* m, const float* params);
int smalgo_mlp_predict(smalgo_mlp* m, const float* X, int B, float* Y);
/* loss = mean((Y - T)^2); grad (optional, num_params floats) = dloss/dparams */
int smalgo_mlp_gradient(smalgo_mlp* m, const float* X, const float* T, int B, float* grad,
                        float* loss);
/* `epochs` full-batch Adam steps (β1=0.9, β2=0.999, ε=1e-8); loss_hist optional */
int smalgo_mlp_train(smalgo_mlp* m, const float* X, const float* T, int B, int epochs, float lr,
                     float* loss_hist, float* loss);

/* ---- POLFED ---------------------------------------------------------------
 * Keyword arguments of polfed(A::MatrixFreeRealSymOp,nev,degree;...)
 * (polfed.jl:1128) plus the mixed-precision controls.  has_* flags stand in
 * for Julia's `nothing` defaults. */
typedef struct smalgo_polfed_params {
  int s;                 /* block size (default 8) */
  int has_bounds;        /* smin/smax given; otherwise Gershgorin (sparse form only) */
  double smin, smax;
  double sigma;          /* default 0.0 */
  double emin, emax;     /* default -Inf, +Inf */
  double tol;            /* default 1e-10 */
  int overextract;       /* default 2 */
  int refine;            /* default 1 */
  int window_type;       /* default SMALGO_WINDOW_STEP */
  int jackson;           /* default 1 */
  int mmax;              /* <=0: ceil(2.8*nev/s) */
  double rtol;           /* default 1e-12 */
  int reorth_passes;     /* default 1 */
  int extra_keep;        /* default 8 */
  uint64_t seed;         /* default 1 (start block when Q1 == NULL) */
  int chunk;             /* default 64 (kept for signature parity; the device lifts all at once) */
  int show_progress;     /* default 0: print the progress line (polfed.jl:991-1000) */
  /* mixed precision (no reference counterpart) */
  int precision;         /* SMALGO_PREC_* (default FP64) */
  int refine_guard;      /* guard vectors in FP64 refinement (default -1: nev, at most the
                            Ritz vectors left after the candidates) */
  int refine_degree;     /* degree of the FP64 refinement filter (default -1: degree) */
  int refine_maxiter;    /* FP64 refinement iterations with locking (default 3) */
  double refine_tol;     /* lock a pair when ‖r‖ < refine_tol·‖A‖; default -1: when ‖r‖ < tol,
                            the caller's absolute tolerance of the return rule */
  int refine_guard_degree; /* degree of the guard vectors' filter in the first refinement
                              iteration (default -1: a fiftieth of the refinement degree,
                              rounded up; the guards only widen the Rayleigh–Ritz subspace) */
} smalgo_polfed_params;

void smalgo_polfed_default_params(smalgo_polfed_params* p);

/* polfed(A, nev, degree; kwargs...).  Q1 (optional, n x s column-major) is the
 * start block BEFORE orthonormalisation — the reference's randn(n,s) after
 * Random.seed!(seed) (polfed.jl:1133-1134); NULL draws it from the library's
 * splitmix64/Box-Muller stream with p->seed.  Exactly one of A or fn is set
 * (fn: matrix-free form, polfed.jl:1183-1186, FP64 only). */
int smalgo_polfed(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                  int64_t n, int nev, int degree, const smalgo_polfed_params* p,
                  const double* Q1, smalgo_result** out);

/* ---- batched spectral slices (no reference counterpart) -----------------
 * The reference solves one interval per polfed call (polfed.jl:1128).  Here
 * many intervals of one matrix run together: slice i is
 *   polfed(A, nev_i, degree_i; emin=emin_i, emax=emax_i, sigma=sigma_i, s=s_i, mmax=mmax_i, ...)
 * with every other option from p (window, tol, precision, refinement, ...),
 * start block from seed p->seed + i.  At each block step the filters of all
 * active slices share one packed launch sequence: their blocks are packed
 * column-wise into dense panels with per-column Chebyshev coefficients
 * (slices of lower degree wait through leading zero coefficients; slices
 * share a panel only where the measured step times say it pays), and the
 * sequence is captured in HIP graphs once and replayed (use_graph = 1) or
 * launched eagerly (0).  SMALGO_SLICE_PW fixes the panel width. */
typedef struct smalgo_slice {
  double emin, emax;  /* target interval */
  double sigma;       /* window centre (delta window) */
  int nev;            /* eigenpairs wanted */
  int s;              /* block width of this slice */
  int degree;         /* filter degree K */
  int mmax;           /* basis blocks; <= 0: ceil(2.8 nev / s) */
} smalgo_slice;

/* out: array of nslices result handles (each freed with smalgo_result_free);
 * t_total / t_filter / t_orth statistics are those of the whole batch. */
int smalgo_polfed_slices(smalgo_ctx* ctx, const smalgo_mat* A, int nslices,
                         const smalgo_slice* slices, const smalgo_polfed_params* p, int use_graph,
                         smalgo_result** out);

/* The packed multi-slice filter alone: Y_i = p_i(Â) Q_i for nslices blocks of
 * widths[i] columns, coefficients concatenated (degrees[i] + 1 each).  Q and Y
 * hold `replays` consecutive host blocks, each n x Σwidths column-major
 * (leading dimension ldq / ldy, slice i at the running column offset); block r
 * goes through the same plan (and, with use_graph, the same instantiated
 * graphs) as block 0.  panel_width > 0 gives every panel that width (a
 * multiple of 4); 0 sizes each panel (rows of at most 512 B, less where three
 * panel blocks would not fit the 256 MB Infinity Cache) from the measured
 * step time of each width.  stats (optional, 6 doubles): widest panel,
 * panels, launches per run, ms per run, modelled bytes per run, graphs per run. */
int smalgo_apply_poly_slices(smalgo_ctx* ctx, const smalgo_mat* A, int nslices, const int* widths,
                             const int* degrees, const double* coeffs, double smin, double smax,
                             int precision, int panel_width, int use_graph, int replays,
                             const double* Q, int64_t ldq, double* Y, int64_t ldy, double* stats);

/* ---- Chebyshev filtered subspace iteration ------------------------------ */
typedef struct smalgo_cheb_params {
  int has_bounds;
  double smin, smax;
  double sigma;      /* σ, default 0.0 */
  int m0;            /* default 200 */
  int degree;        /* default 80 */
  int maxiter;       /* default 8 */
  double tol;        /* default 1e-10 */
  double res_gate;   /* default 1e-6 */
  int jackson;       /* default 1 */
  int debug;         /* default 0 */
  int two_hit;       /* default 1 */
  uint64_t seed;     /* default 1 */
  int window_type;   /* default SMALGO_WINDOW_STEP */
} smalgo_cheb_params;

void smalgo_cheb_default_params(smalgo_cheb_params* p);

/* chebyshev_poly_real_symm_sparse(A, emin, emax; ...) (chebyshev_subspace_iteration.jl:667-757):
 * always runs maxiter iterations.  With fn set instead of A it is the
 * matvec! form (:771-848): bounds required, two_hit early exit.  Q0
 * (optional, n x m0 column-major) replaces randn(n,m0) (:701-702). */
int smalgo_chebyshev(smalgo_ctx* ctx, const smalgo_mat* A, smalgo_matblock_fn fn, void* user,
                     int64_t n, double emin, double emax, const smalgo_cheb_params* p,
                     const double* Q0, smalgo_result** out);

/* ---- on-disk fixtures (no reference counterpart) --------------------------
 * One little-endian binary format for matrices and eigenvalue sets, shared
 * by the C, Python (sparse_matrix_algorithms_amd.fixtures) and Julia front
 * ends.  64-byte header, then the payload; the header's checksum is
 * Σ_i (w_i xor i·0x9E3779B97F4A7C15)·0xBF58476D1CE4E5B9 mod 2^64 over the
 * payload's 8-byte words w_i.
 *   CSR: "SMALGO-CSR-v1" (16 bytes, NUL-padded) | int64 n | int64 nnz |
 *        int32 index_base = 0 | int32 flags (bit 0 symmetric) | uint64 checksum |
 *        16 zero bytes | int64 rowptr[n+1] | int64 col[nnz] | float64 val[nnz]
 *   Eig: "SMALGO-EIG-v1" | int64 count | int64 has_res | uint64 checksum |
 *        24 zero bytes | float64 lam[count] | float64 res[count] if has_res
 * These calls need no context and no GPU; smalgo_io_last_error() gives the
 * message of the calling thread's last failure. */
int smalgo_csr_save(const char* path, int64_t n, const int64_t* rowptr, const int64_t* colidx,
                    const double* val, int index_base, int symmetric);
int smalgo_csr_info(const char* path, int64_t* n, int64_t* nnz, int* symmetric);
/* rowptr (n+1), colidx (nnz), val (nnz), 0-based; any may be NULL to skip */
int smalgo_csr_load(const char* path, int64_t* rowptr, int64_t* colidx, double* val);
/* load a CSR fixture and upload it (smalgo_csr_upload with index_base 0) */
int smalgo_csr_upload_file(smalgo_ctx* ctx, const char* path, smalgo_mat** out);
int smalgo_eig_save(const char* path, int64_t count, const double* lam, const double* res);
/* lam and res NULL: only count and has_res */
int smalgo_eig_load(const char* path, int64_t* count, int* has_res, double* lam, double* res);
const char* smalgo_io_last_error(void);

/* ---- results ---------------------------------------------------------------
 * (λ::Vector, X::Matrix n x k column-major, res::Vector) as returned by the
 * drivers (polfed.jl:1148-1157, chebyshev_subspace_iteration.jl:750-756). */
int64_t smalgo_result_count(const smalgo_result* r);
int64_t smalgo_result_n(const smalgo_result* r);
/* any of lam / X / res may be NULL to skip it */
int smalgo_result_copy(const smalgo_result* r, double* lam, double* X, int64_t ldx, double* res);
/* named run statistics (seconds, counts); returns SMALGO_ERR_ARG for unknown names.
 * Names: "t_total", "t_filter", "t_orth", "t_extract", "t_refine", "filter_steps",
 * "block_steps", "basis_dim", "iterations", "spmm_bytes", "gemm_flops",
 * "refine_filter_steps", "locked". */
int smalgo_result_stat(const smalgo_result* r, const char* name, double* value);
/* the progress line of polfed.jl:991-1000 for this result */
const char* smalgo_result_progress_line(const smalgo_result* r);
void smalgo_result_free(smalgo_result* r);

#ifdef __cplusplus
}
#endif

#endif /* SMALGO_H */
