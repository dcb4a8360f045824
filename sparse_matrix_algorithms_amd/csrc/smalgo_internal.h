// Internal declarations shared by the HIP kernels and the host drivers.
// Nothing here is part of the public C ABI (include/smalgo.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace smalgo {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SMALGO_HIP(call)                                                              \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw ::smalgo::Error(-2, std::string("HIP error ") + hipGetErrorString(e_) +  \
                                    " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

// ---------------------------------------------------------------------------
// CSR matrix resident in HBM (0-based int32 indices, values in FP64 and FP32)
// ---------------------------------------------------------------------------
struct DevCSR {
  int n = 0;
  int nnz = 0;
  int* rowptr = nullptr;   // n+1
  int* col = nullptr;      // nnz
  double* val64 = nullptr; // nnz
  float* val32 = nullptr;  // nnz
  double gersh_min = 0, gersh_max = 0;
  template <typename T> const T* vals() const;
};
template <> inline const double* DevCSR::vals<double>() const { return val64; }
template <> inline const float* DevCSR::vals<float>() const { return val32; }

// ---------------------------------------------------------------------------
// fused Clenshaw step / SpMM
//   out[i, c] = cq*Q[i,c] + a1*(A X1)[i,c] + b1*X1[i,c] - s2*X2[i,c]     c < b
// Row-major blocks: element (i,c) of a block with leading dimension ld sits
// at i*ld + c.  Q and X2 may be null (their terms are dropped).  out may alias
// X2 (same pointer and ld): every element of X2 is read by the lane that
// writes the same element of out.
// ---------------------------------------------------------------------------
struct FilterTile {
  int vec = 1;   // elements per lane per load
  int lpr = 64;  // lanes per row (rows per wavefront = 64/lpr)
  int rt = 1;    // consecutive rows per lane team
  int lds = 0;   // 1: the workgroup's CSR tile is staged in LDS first; 2: staged, with the
                 // own-row loads issued before the staging barrier (rt <= 2)
  int split = 1; // 2: each step runs as two launches over the two column halves, one on each
                 // of the context's two streams (vec/lpr/rt/lds then describe a half)
};

template <typename T>
void launch_clenshaw_step(hipStream_t st, const DevCSR& A, int b, const T* Q, int ldq, T cq,
                          const T* X1, int ld1, T a1, T b1, const T* X2, int ld2, T s2, T* out,
                          int ldo, FilterTile tile);

// Per-column variant for batched slices: the Q coefficient is read from a
// per-column device row (cq[c]); a1, b1 and s2 are shared by all columns.
template <typename T>
void launch_clenshaw_step_cols(hipStream_t st, const DevCSR& A, int b, const T* Q, int ldq,
                               const T* cq, const T* X1, int ld1, T a1, T b1, const T* X2, int ld2,
                               T s2, T* out, int ldo, FilterTile tile);

// Panel layout for wide blocks: ceil(b/PW) dense n x PW slabs (slab stride n*PW).
template <typename T>
void to_panels(hipStream_t st, int n, int b, int pw, const T* S, int lds, T* D);
template <typename T>
void from_panels(hipStream_t st, int n, int b, int pw, const T* S, T* D, int ldd);

// ---------------------------------------------------------------------------
// tall-skinny MFMA GEMMs
//   gemm_tn:  C[m x b] = alpha * A[n x m]^T B[n x b] + beta * C     (split-K over n)
//   gemm_nn:  C[n x b] = alpha * A[n x m] B[m x b] + beta * C
// A, B, C row-major with leading dimensions; in gemm_nn, transB=true reads
// B(k,j) from B[j*ldb + k] (a column-major m x b matrix).  TC is the output
// type; the MFMA runs in double when either input or the output is double.
// ---------------------------------------------------------------------------
template <typename TA, typename TB, typename TC>
void gemm_tn(hipStream_t st, int n, int m, int b, double alpha, const TA* A, int lda, const TB* B,
             int ldb, double beta, TC* C, int ldc, double* work, size_t work_elems);

size_t gemm_tn_work_elems(int n, int m, int b);

template <typename TA, typename TB, typename TC>
void gemm_nn(hipStream_t st, int n, int m, int b, double alpha, const TA* A, int lda, const TB* B,
             int ldb, bool transB, double beta, TC* C, int ldc);

// ---------------------------------------------------------------------------
// column reductions and small helpers
// ---------------------------------------------------------------------------
// out[c] = sum_i X[i,c] * Y[i,c]   (double accumulation), c < b
template <typename TX>
void col_dots(hipStream_t st, int n, int b, const TX* X, int ldx, const TX* Y, int ldy, double* out,
              double* work, size_t work_elems);
// out[c] = || AX[:,c] - lam[c] X[:,c] ||_2
template <typename TX>
void col_resid(hipStream_t st, int n, int b, const TX* X, int ldx, const TX* AX, int ldax,
               const double* lam, double* out, double* work, size_t work_elems);
size_t colred_work_elems(int n, int b);

template <typename TS, typename TD>
void copy_block(hipStream_t st, int n, int b, const TS* S, int lds, TD* D, int ldd);
template <typename T>
void fill_block(hipStream_t st, int n, int b, T* D, int ldd, T v);
template <typename T>
void scale_cols(hipStream_t st, int n, int b, T* D, int ldd, const double* s);
// D[i, j] = S[perm_row(i), sel[j]] where S is column-major (ld lds) and the
// rows of D are scattered by rowmap (D row rowmap[i] <- S row i).
void gather_cols_colmajor(hipStream_t st, int M, int m, const double* S, int lds, const int* sel,
                          const int* rowmap, double* D, int ldd);
// symmetric part, upper-mirrored: C := sym_upper(0.5*(C + C^T)), row-major b x b
void symmetrize(hipStream_t st, int b, double* C, int ldc);
// Gaussian start block from the counter-based splitmix64 stream (host RNG
// twin lives in the Python lattices module).
template <typename T>
void gaussian_block(hipStream_t st, int n, int b, uint64_t seed, T* D, int ldd);
template <typename TS, typename TD>
void cm_to_rm(hipStream_t st, int n, int b, const TS* S, int lds, TD* D, int ldd);
template <typename TS>
void rm_to_cm(hipStream_t st, int n, int k, const TS* S, int lds, const int* sel, double* D, int ldd);
template <typename T>
void gather_cols_rm(hipStream_t st, int n, int k, const T* S, int lds, const int* sel, T* D, int ldd);
void add_diag(hipStream_t st, int b, double* C, int ldc, double v);
// C = A·B for upper-triangular column-major b x b (ld b) matrices
void tri_mul_upper(hipStream_t st, int b, const double* A, const double* B, double* C);
// G = RᵀR (upper Cholesky, upper triangle of G read) and Ri = R⁻¹ in one
// workgroup, b <= chol_inv_max_b(); info as potrf's; G must not alias R, Ri
int chol_inv_max_b();
void chol_inv_upper(hipStream_t st, int b, const double* G, int ldg, double* R, int ldr,
                    double* Ri, int ldri, int* info);
// out[0..1] = info[0..1], out[2..2+b) = diag(R)  (R column-major, ld b)
void pack_qr_status(hipStream_t st, int b, const int* info, const double* R, double* out);
// s x s block, column-major: rows < rn and columns < r of R (ld r), zero elsewhere
void b_block(hipStream_t st, int s, int r, int rn, const double* R, double* Bt);
void zero_lower_cm(hipStream_t st, int b, double* C, int ldc);
void launch_clenshaw_step_pre(hipStream_t st, int n, int b, const double* Q, int ldq, double cq,
                              const double* P, int ldp, const double* X1, int ld1, double a1,
                              double b1, const double* X2, int ld2, double s2, double* out,
                              int ldo);
// 16-byte-vector grid-stride copy of `bytes` (multiple of 16)
void stream_copy(hipStream_t st, const void* src, void* dst, size_t bytes);
// number of compute units, queried once per device
int device_cu_count();

// on-disk fixtures (smalgo_io.cpp)
void csr_save(const char* path, int64_t n, const int64_t* rowptr, const int64_t* col,
              const double* val, int index_base, int symmetric);
void csr_info(const char* path, int64_t* n, int64_t* nnz, int* symmetric);
void csr_load(const char* path, int64_t* rowptr, int64_t* col, double* val);
void eig_save(const char* path, int64_t count, const double* lam, const double* res);
void eig_load(const char* path, int64_t* count, int* has_res, double* lam, double* res);

}  // namespace smalgo
