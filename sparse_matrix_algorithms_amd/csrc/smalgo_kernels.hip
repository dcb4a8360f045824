// HIP kernels for gfx950 (MI355X): fused CSR Clenshaw step, MFMA tall-skinny
// GEMMs, column reductions and small block helpers.
//
// Block layout: row-major n x b blocks (element (i,c) at i*ld + c), so the
// b values of one row are contiguous and the CSR gather A·X reads whole rows
// of X with coalesced vector loads.
#include "smalgo_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace smalgo {

// ---------------------------------------------------------------------------
// vector load/store helpers
// ---------------------------------------------------------------------------
template <typename T, int V> struct VecT;
template <typename T> struct VecT<T, 1> { using type = T; };
template <> struct VecT<float, 2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct VecT<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct VecT<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct VecT<double, 4> { typedef double type __attribute__((ext_vector_type(4))); };
template <> struct VecT<float, 8> { typedef float type __attribute__((ext_vector_type(8))); };
template <> struct VecT<double, 8> { typedef double type __attribute__((ext_vector_type(8))); };

template <typename T, int V>
__device__ __forceinline__ void ldv(const T* __restrict__ p, T (&r)[V]) {
  if constexpr (V == 1) {
    r[0] = *p;
  } else {
    using VT = typename VecT<T, V>::type;
    VT v = *reinterpret_cast<const VT*>(p);
#pragma unroll
    for (int e = 0; e < V; ++e) r[e] = v[e];
  }
}
template <typename T, int V>
__device__ __forceinline__ void ldv_alias(const T* p, T (&r)[V]) {
  if constexpr (V == 1) {
    r[0] = *p;
  } else {
    using VT = typename VecT<T, V>::type;
    VT v = *reinterpret_cast<const VT*>(p);
#pragma unroll
    for (int e = 0; e < V; ++e) r[e] = v[e];
  }
}
template <typename T, int V>
__device__ __forceinline__ void stv(T* p, const T (&r)[V]) {
  if constexpr (V == 1) {
    *p = r[0];
  } else {
    using VT = typename VecT<T, V>::type;
    VT v;
#pragma unroll
    for (int e = 0; e < V; ++e) v[e] = r[e];
    *reinterpret_cast<VT*>(p) = v;
  }
}

// XCD-aware block index: the dispatcher deals workgroups round-robin over the
// 8 XCDs (workgroup w lands on XCD w % 8).  Mapping logical tile
// (w % 8) * per_xcd + w / 8 gives each XCD one contiguous range of rows, so
// the rows it gathers for its CSR neighbours stay in its own L2.
__device__ __forceinline__ int xcd_tile(int per_xcd) {
  const int w = blockIdx.x;
  return (w & 7) * per_xcd + (w >> 3);
}

// SMALGO_XCD_REMAP=0 turns the remapping off (tile = workgroup id) for A/B tests
static bool xcd_remap_enabled() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("SMALGO_XCD_REMAP");
    v = (e && e[0] == '0') ? 0 : 1;
  }
  return v == 1;
}

static inline int grid_xcd(int ntiles, int* per_xcd) {
  const int g = (ntiles + 7) / 8 * 8;
  *per_xcd = g / 8;
  return g;
}

// ---------------------------------------------------------------------------
// fused Clenshaw step
//   out = cq*Q + a1*(A X1 or P) + b1*X1 - s2*X2
// A team of LPR lanes owns one row of one column chunk of width LPR*V; each
// lane covers V consecutive columns; 64/LPR rows per wavefront, four
// wavefronts per workgroup.  Tiles run column-chunk-major, so with the
// XCD-aware mapping every XCD streams a contiguous row range of one column
// chunk and keeps the rows its CSR neighbours gather within its own L2.
// ---------------------------------------------------------------------------
template <typename T, int V, int LPR, int RT, bool PERCOL>
__global__ __launch_bounds__(256) void clenshaw_kernel(
    int n, int b, int nrowtiles, int ntiles, int per_xcd, const int* __restrict__ rowptr,
    const int* __restrict__ colidx, const T* __restrict__ val, const T* __restrict__ Q, int ldq,
    const T* __restrict__ X1, int ld1, const T* __restrict__ P, int ldp, const T* X2, int ld2,
    T* out, int ldo, T cq, T a1, T b1, T s2, const T* __restrict__ pcq, long long chunk_stride) {
  constexpr int RPW = 64 / LPR;           // teams per wavefront
  constexpr int RPB = 4 * RPW * RT;       // rows per workgroup tile
  const int tile = per_xcd ? xcd_tile(per_xcd) : (int)blockIdx.x;
  if (tile >= ntiles) return;
  const int chunk = tile / nrowtiles;
  const int rtile = tile - chunk * nrowtiles;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int row0 = rtile * RPB + (wave * RPW + lane / LPR) * RT;  // RT consecutive rows
  if (row0 >= n) return;
  const int cic = (lane % LPR) * V;           // column inside the chunk
  const int c = chunk * (LPR * V) + cic;      // global column
  if (c >= b) return;
  // Row-major blocks: chunk_stride == 0 and the chunk is a column offset.
  // Panel blocks (dense n x LPR*V slabs): chunk_stride = n*LPR*V, ld = LPR*V.
  const long long coff = chunk_stride ? (long long)chunk * chunk_stride + cic : c;
  if (Q) Q += coff;
  if (X1) X1 += coff;
  if (P) P += coff;
  if (X2) X2 += coff;
  out += coff;
  const bool gather = (X1 != nullptr) && (P == nullptr);
  const bool full = (c + V <= b);
  int p0[RT], len[RT];
  int maxlen = 0;
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    const int row = min(row0 + r, n - 1);
    p0[r] = gather ? rowptr[row] : 0;
    len[r] = gather ? (row0 + r < n ? rowptr[row + 1] - p0[r] : 0) : 0;
    maxlen = max(maxlen, len[r]);
  }
  T acc[RT][V];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int e = 0; e < V; ++e) acc[r][e] = T(0);
  if (full) {
    if (gather) {
      // all index, value and gather loads unconditional (clamped index, zero
      // weight past the row end) so the RT*U load chains of a lane overlap
      constexpr int U = RT >= 4 ? 1 : (RT == 2 ? 2 : 4);
      for (int j = 0; j < maxlen; j += U) {
        int col[U][RT];
        T w[U][RT];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int r = 0; r < RT; ++r) {
            const int jj = j + u;
            const int q = len[r] > 0 ? p0[r] + min(jj, len[r] - 1) : 0;
            col[u][r] = len[r] > 0 ? colidx[q] : 0;
            w[u][r] = (jj < len[r]) ? val[q] : T(0);
          }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int r = 0; r < RT; ++r) {
            T x[V];
            ldv<T, V>(X1 + (size_t)col[u][r] * ld1, x);
#pragma unroll
            for (int e = 0; e < V; ++e) acc[r][e] += w[u][r] * x[e];
          }
      }
    }
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      const int row = row0 + r;
      if (row >= n) break;
      if (P != nullptr) ldv<T, V>(P + (size_t)row * ldp, acc[r]);
      T res[V];
#pragma unroll
      for (int e = 0; e < V; ++e) res[e] = a1 * acc[r][e];
      if (Q != nullptr) {
        T q[V];
        ldv<T, V>(Q + (size_t)row * ldq, q);
#pragma unroll
        for (int e = 0; e < V; ++e) res[e] = (PERCOL ? pcq[c + e] : cq) * q[e] + res[e];
      }
      if (X1 != nullptr) {
        T x[V];
        ldv<T, V>(X1 + (size_t)row * ld1, x);
#pragma unroll
        for (int e = 0; e < V; ++e) res[e] += b1 * x[e];
      }
      if (X2 != nullptr) {
        T y[V];
        ldv_alias<T, V>(X2 + (size_t)row * ld2, y);
#pragma unroll
        for (int e = 0; e < V; ++e) res[e] -= s2 * y[e];
      }
      stv<T, V>(out + (size_t)row * ldo, res);
    }
  } else {
    // ragged tail of the block width: element-wise
    for (int r = 0; r < RT; ++r) {
      const int row = row0 + r;
      if (row >= n) break;
      for (int e = 0; e < V && c + e < b; ++e) {
        const int cc = c + e;  // global column (per-column scalars)
        T a = T(0);
        if (gather) {
          for (int p = p0[r]; p < p0[r] + len[r]; ++p) a += val[p] * X1[(size_t)colidx[p] * ld1 + e];
        } else if (P != nullptr) {
          a = P[(size_t)row * ldp + e];
        }
        T res = a1 * a;
        if (Q != nullptr) res = (PERCOL ? pcq[cc] : cq) * Q[(size_t)row * ldq + e] + res;
        if (X1 != nullptr) res += b1 * X1[(size_t)row * ld1 + e];
        if (X2 != nullptr) res -= s2 * X2[(size_t)row * ld2 + e];
        out[(size_t)row * ldo + e] = res;
      }
    }
  }
}

// The same fused step with the workgroup's CSR tile staged in LDS first:
// the rows' pointers, column indices and values of one row tile are loaded
// once per workgroup with coalesced loads, so the LPR lanes of a team read
// them from LDS (broadcast reads) instead of each issuing its own global
// index and value loads per nonzero.  Tiles with more than 12 nonzeros per
// row on average take the global path.  Gather form only (X1 set, P null).
//
// PF (prefetch): each lane issues its own rows' Q, X1 and X2 loads before the
// staging barrier and gathers 8/RT neighbour rows per loop trip, so a row of
// up to eight nonzeros costs two memory round trips per wave (staging with
// the own-row loads, then the whole gather) instead of four.
template <typename T, int V, int LPR, int RT, bool PERCOL, bool PF>
__global__ __launch_bounds__(256) void clenshaw_lds_kernel(
    int n, int b, int nrowtiles, int ntiles, int per_xcd, const int* __restrict__ rowptr,
    const int* __restrict__ colidx, const T* __restrict__ val, const T* __restrict__ Q, int ldq,
    const T* __restrict__ X1, int ld1, const T* X2, int ld2, T* out, int ldo, T cq, T a1, T b1,
    T s2, const T* __restrict__ pcq, long long chunk_stride) {
  constexpr int RPW = 64 / LPR;
  constexpr int RPB = 4 * RPW * RT;
  constexpr int MAXE = RPB * 12;
  __shared__ int s_ptr[RPB + 1];
  __shared__ int s_col[MAXE];
  __shared__ T s_val[MAXE];
  const int tile = per_xcd ? xcd_tile(per_xcd) : (int)blockIdx.x;
  if (tile >= ntiles) return;  // whole workgroup
  const int chunk = tile / nrowtiles;
  const int rtile = tile - chunk * nrowtiles;
  const int rbase = rtile * RPB;
  const int nrows = min(RPB, n - rbase);  // >= 1: rtile < nrowtiles
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int row0 = rbase + (wave * RPW + lane / LPR) * RT;
  const int cic = (lane % LPR) * V;
  const int c = chunk * (LPR * V) + cic;
  const bool active = row0 < n && c < b;  // lane owns at least one element
  const bool full = (c + V <= b);
  const long long coff = chunk_stride ? (long long)chunk * chunk_stride + cic : c;
  T qv[PF ? RT : 1][V], xv[PF ? RT : 1][V], yv[PF ? RT : 1][V];
  if constexpr (PF) {
    if (active && full) {
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        const size_t row = (size_t)min(row0 + r, n - 1);  // clamped: unused past n
        if (Q != nullptr) ldv<T, V>(Q + coff + row * ldq, qv[r]);
        ldv<T, V>(X1 + coff + row * ld1, xv[r]);
        if (X2 != nullptr) ldv_alias<T, V>(X2 + coff + row * ld2, yv[r]);
      }
    }
  }
  const int e0 = rowptr[rbase];
  const int ne = rowptr[rbase + nrows] - e0;
  const bool staged = ne <= MAXE;  // uniform over the workgroup
  if (staged) {
    for (int t = threadIdx.x; t <= nrows; t += 256) s_ptr[t] = rowptr[rbase + t] - e0;
    for (int t = threadIdx.x; t < ne; t += 256) {
      s_col[t] = colidx[e0 + t];
      s_val[t] = val[e0 + t];
    }
  }
  __syncthreads();
  if (!active) return;
  if (Q) Q += coff;
  X1 += coff;
  if (X2) X2 += coff;
  out += coff;
  int p0[RT], len[RT];
  int maxlen = 0;
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    const int row = min(row0 + r, n - 1);
    if (staged) {
      p0[r] = s_ptr[row - rbase];
      len[r] = row0 + r < n ? s_ptr[row - rbase + 1] - p0[r] : 0;
    } else {
      p0[r] = rowptr[row];
      len[r] = row0 + r < n ? rowptr[row + 1] - p0[r] : 0;
    }
    maxlen = max(maxlen, len[r]);
  }
  auto colv = [&](int q, bool in, int& cidx, T& w) {
    if (staged) {
      cidx = s_col[q];
      w = in ? s_val[q] : T(0);
    } else {
      cidx = colidx[q];
      w = in ? val[q] : T(0);
    }
  };
  T acc[RT][V];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int e = 0; e < V; ++e) acc[r][e] = T(0);
  if (full) {
    constexpr int U = PF ? (RT >= 4 ? 2 : 8 / RT) : (RT >= 4 ? 1 : (RT == 2 ? 2 : 4));
    for (int j = 0; j < maxlen; j += U) {
      int col[U][RT];
      T w[U][RT];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int r = 0; r < RT; ++r) {
          const int jj = j + u;
          if (len[r] > 0) {
            colv(p0[r] + min(jj, len[r] - 1), jj < len[r], col[u][r], w[u][r]);
          } else {
            col[u][r] = 0;
            w[u][r] = T(0);
          }
        }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int r = 0; r < RT; ++r) {
          T x[V];
          ldv<T, V>(X1 + (size_t)col[u][r] * ld1, x);
#pragma unroll
          for (int e = 0; e < V; ++e) acc[r][e] += w[u][r] * x[e];
        }
    }
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      const int row = row0 + r;
      if (row >= n) break;
      const int rp = PF ? r : 0;
      T res[V];
#pragma unroll
      for (int e = 0; e < V; ++e) res[e] = a1 * acc[r][e];
      if (Q != nullptr) {
        if constexpr (!PF) ldv<T, V>(Q + (size_t)row * ldq, qv[0]);
#pragma unroll
        for (int e = 0; e < V; ++e) res[e] = (PERCOL ? pcq[c + e] : cq) * qv[rp][e] + res[e];
      }
      if constexpr (!PF) ldv<T, V>(X1 + (size_t)row * ld1, xv[0]);
#pragma unroll
      for (int e = 0; e < V; ++e) res[e] += b1 * xv[rp][e];
      if (X2 != nullptr) {
        if constexpr (!PF) ldv_alias<T, V>(X2 + (size_t)row * ld2, yv[0]);
#pragma unroll
        for (int e = 0; e < V; ++e) res[e] -= s2 * yv[rp][e];
      }
      stv<T, V>(out + (size_t)row * ldo, res);
    }
  } else {
    for (int r = 0; r < RT; ++r) {
      const int row = row0 + r;
      if (row >= n) break;
      for (int e = 0; e < V && c + e < b; ++e) {
        const int cc = c + e;
        T a = T(0);
        for (int q = p0[r]; q < p0[r] + len[r]; ++q) {
          int ci;
          T w;
          colv(q, true, ci, w);
          a += w * X1[(size_t)ci * ld1 + e];
        }
        T res = a1 * a;
        if (Q != nullptr) res = (PERCOL ? pcq[cc] : cq) * Q[(size_t)row * ldq + e] + res;
        res += b1 * X1[(size_t)row * ld1 + e];
        if (X2 != nullptr) res -= s2 * X2[(size_t)row * ld2 + e];
        out[(size_t)row * ldo + e] = res;
      }
    }
  }
}

template <typename T, int V, int LPR, int RT, bool PERCOL>
static void clenshaw_launch_t(hipStream_t st, const DevCSR& A, int b, const T* Q, int ldq,
                              const T* X1, int ld1, const T* P, int ldp, const T* X2, int ld2,
                              T* out, int ldo, T cq, T a1, T b1, T s2, const T* pcq, bool panels,
                              int lds) {
  constexpr int RPB = 4 * (64 / LPR) * RT;
  const int nrowtiles = (A.n + RPB - 1) / RPB;
  const int nchunks = (b + LPR * V - 1) / (LPR * V);
  const long long nt = (long long)nrowtiles * nchunks;
  if (nt > 0x7fffff00LL) throw Error(-3, "filter grid too large");
  int per_xcd = 0;
  int grid = grid_xcd((int)nt, &per_xcd);
  if (!xcd_remap_enabled()) {
    per_xcd = 0;
    grid = (int)nt;
  }
  const long long cs = panels ? (long long)A.n * (LPR * V) : 0LL;
  if (lds == 2 && RT <= 2 && X1 != nullptr && P == nullptr)
    hipLaunchKernelGGL((clenshaw_lds_kernel<T, V, LPR, RT, PERCOL, (RT <= 2)>), dim3(grid), dim3(256),
                       0, st, A.n, b, nrowtiles, (int)nt, per_xcd, A.rowptr, A.col, A.vals<T>(), Q,
                       ldq, X1, ld1, X2, ld2, out, ldo, cq, a1, b1, s2, pcq, cs);
  else if (lds && X1 != nullptr && P == nullptr)
    hipLaunchKernelGGL((clenshaw_lds_kernel<T, V, LPR, RT, PERCOL, false>), dim3(grid), dim3(256), 0,
                       st, A.n, b, nrowtiles, (int)nt, per_xcd, A.rowptr, A.col, A.vals<T>(), Q, ldq,
                       X1, ld1, X2, ld2, out, ldo, cq, a1, b1, s2, pcq, cs);
  else
    hipLaunchKernelGGL((clenshaw_kernel<T, V, LPR, RT, PERCOL>), dim3(grid), dim3(256), 0, st, A.n,
                       b, nrowtiles, (int)nt, per_xcd, A.rowptr, A.col, A.vals<T>(), Q, ldq, X1,
                       ld1, P, ldp, X2, ld2, out, ldo, cq, a1, b1, s2, pcq, cs);
  SMALGO_HIP(hipGetLastError());
}

template <typename T, bool PERCOL>
static void clenshaw_dispatch(hipStream_t st, const DevCSR& A, int b, const T* Q, int ldq,
                              const T* X1, int ld1, const T* P, int ldp, const T* X2, int ld2,
                              T* out, int ldo, T cq, T a1, T b1, T s2, const T* pcq, FilterTile t,
                              bool panels = false) {
#define SMALGO_CL(VV, LL, RR)                                                                  \
  if (t.vec == VV && t.lpr == LL && t.rt == RR)                                                \
    return clenshaw_launch_t<T, VV, LL, RR, PERCOL>(st, A, b, Q, ldq, X1, ld1, P, ldp, X2, ld2, \
                                                    out, ldo, cq, a1, b1, s2, pcq, panels,  \
                                                    t.lds);
#define SMALGO_CL_V(VV, RR) \
  SMALGO_CL(VV, 4, RR) SMALGO_CL(VV, 8, RR) SMALGO_CL(VV, 16, RR) SMALGO_CL(VV, 32, RR) SMALGO_CL(VV, 64, RR)
  SMALGO_CL_V(1, 1) SMALGO_CL_V(2, 1) SMALGO_CL_V(4, 1) SMALGO_CL_V(8, 1)
  SMALGO_CL_V(1, 2) SMALGO_CL_V(2, 2) SMALGO_CL_V(4, 2) SMALGO_CL_V(8, 2)
  SMALGO_CL_V(1, 4) SMALGO_CL_V(2, 4) SMALGO_CL_V(4, 4) SMALGO_CL_V(8, 4)
#undef SMALGO_CL_V
#undef SMALGO_CL
  throw Error(-3, "unsupported filter tile vec=" + std::to_string(t.vec) +
                      " lpr=" + std::to_string(t.lpr) + " rt=" + std::to_string(t.rt));
}

static void check_tile_alignment(FilterTile t, size_t elem, std::initializer_list<const void*> ptrs,
                                 std::initializer_list<int> lds) {
  const size_t al = elem * (size_t)t.vec;
  for (const void* p : ptrs)
    if (p && (reinterpret_cast<uintptr_t>(p) % al) != 0)
      throw Error(-3, "filter tile vec=" + std::to_string(t.vec) + " needs aligned block pointers");
  for (int ld : lds)
    if (ld % t.vec != 0)
      throw Error(-3, "filter tile vec=" + std::to_string(t.vec) + " needs ld % vec == 0");
}

template <typename T>
void launch_clenshaw_step(hipStream_t st, const DevCSR& A, int b, const T* Q, int ldq, T cq,
                          const T* X1, int ld1, T a1, T b1, const T* X2, int ld2, T s2, T* out,
                          int ldo, FilterTile tile) {
  if (A.n == 0 || b == 0) return;
  check_tile_alignment(tile, sizeof(T), {Q, X1, X2, out}, {Q ? ldq : 0, X1 ? ld1 : 0, X2 ? ld2 : 0, ldo});
  clenshaw_dispatch<T, false>(st, A, b, Q, ldq, X1, ld1, nullptr, 0, X2, ld2, out, ldo, cq, a1, b1,
                              s2, nullptr, tile);
}

template <typename T>
void launch_clenshaw_step_cols(hipStream_t st, const DevCSR& A, int b, const T* Q, int ldq,
                               const T* cq, const T* X1, int ld1, T a1, T b1, const T* X2, int ld2,
                               T s2, T* out, int ldo, FilterTile tile) {
  if (A.n == 0 || b == 0) return;
  if (!cq) throw Error(-1, "per-column filter step needs its coefficient row");
  check_tile_alignment(tile, sizeof(T), {Q, X1, X2, out}, {Q ? ldq : 0, X1 ? ld1 : 0, X2 ? ld2 : 0, ldo});
  clenshaw_dispatch<T, true>(st, A, b, Q, ldq, X1, ld1, nullptr, 0, X2, ld2, out, ldo, T(0), a1, b1,
                             s2, cq, tile);
}

// pack a row-major block (ld) into panels of width pw; unpack the reverse
template <typename T>
__global__ void to_panels_kernel(int n, int b, int pw, const T* __restrict__ S, int lds,
                                 T* __restrict__ D) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int npan = (b + pw - 1) / pw;
  if (idx >= (size_t)n * npan * pw) return;
  const int cc = (int)(idx % pw);
  const size_t r = idx / pw;
  const int i = (int)(r % n), pan = (int)(r / n);
  const int c = pan * pw + cc;
  D[idx] = c < b ? S[(size_t)i * lds + c] : T(0);
}
template <typename T>
__global__ void from_panels_kernel(int n, int b, int pw, const T* __restrict__ S,
                                   T* __restrict__ D, int ldd) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int npan = (b + pw - 1) / pw;
  if (idx >= (size_t)n * npan * pw) return;
  const int cc = (int)(idx % pw);
  const size_t r = idx / pw;
  const int i = (int)(r % n), pan = (int)(r / n);
  const int c = pan * pw + cc;
  if (c < b) D[(size_t)i * ldd + c] = S[idx];
}
template <typename T>
void to_panels(hipStream_t st, int n, int b, int pw, const T* S, int lds, T* D) {
  const size_t tot = (size_t)n * ((b + pw - 1) / pw) * pw;
  if (!tot) return;
  hipLaunchKernelGGL((to_panels_kernel<T>), dim3((tot + 255) / 256), dim3(256), 0, st, n, b, pw, S,
                     lds, D);
  SMALGO_HIP(hipGetLastError());
}
template <typename T>
void from_panels(hipStream_t st, int n, int b, int pw, const T* S, T* D, int ldd) {
  const size_t tot = (size_t)n * ((b + pw - 1) / pw) * pw;
  if (!tot) return;
  hipLaunchKernelGGL((from_panels_kernel<T>), dim3((tot + 255) / 256), dim3(256), 0, st, n, b, pw,
                     S, D, ldd);
  SMALGO_HIP(hipGetLastError());
}
template void to_panels<float>(hipStream_t, int, int, int, const float*, int, float*);
template void to_panels<double>(hipStream_t, int, int, int, const double*, int, double*);
template void from_panels<float>(hipStream_t, int, int, int, const float*, float*, int);
template void from_panels<double>(hipStream_t, int, int, int, const double*, double*, int);

// Axpby form with a precomputed operator product P = A·X1 (host-callback
// operators): out = cq*Q + a1*P + b1*X1 - s2*X2.
void launch_clenshaw_step_pre(hipStream_t st, int n, int b, const double* Q, int ldq, double cq,
                              const double* P, int ldp, const double* X1, int ld1, double a1,
                              double b1, const double* X2, int ld2, double s2, double* out,
                              int ldo) {
  if (n == 0 || b == 0) return;
  DevCSR fake;
  fake.n = n;
  FilterTile t{1, 64, 1};
  clenshaw_dispatch<double, false>(st, fake, b, Q, ldq, X1, ld1, P, ldp, X2, ld2, out, ldo, cq, a1,
                                   b1, s2, nullptr, t);
}

template void launch_clenshaw_step<float>(hipStream_t, const DevCSR&, int, const float*, int, float,
                                          const float*, int, float, float, const float*, int,
                                          float, float*, int, FilterTile);
template void launch_clenshaw_step<double>(hipStream_t, const DevCSR&, int, const double*, int,
                                           double, const double*, int, double, double,
                                           const double*, int, double, double*, int, FilterTile);
template void launch_clenshaw_step_cols<float>(hipStream_t, const DevCSR&, int, const float*, int,
                                               const float*, const float*, int, float, float,
                                               const float*, int, float, float*, int, FilterTile);
template void launch_clenshaw_step_cols<double>(hipStream_t, const DevCSR&, int, const double*, int,
                                                const double*, const double*, int, double, double,
                                                const double*, int, double, double*, int,
                                                FilterTile);

// ---------------------------------------------------------------------------
// MFMA tall-skinny GEMMs (16x16x4 tiles; 4 wavefronts, each a 32x32 sub-tile
// of a 64x64 workgroup tile; K staged through LDS 16 at a time).
//
// v_mfma_*_16x16x4 operand layout used below: lane l supplies A[l%16][l/16]
// and B[l/16][l%16].  Accumulator element r of lane l is D[row][l%16] with
// row = 4*(l/16)+r for f32 and row = l/16 + 4*r for f64 (measured on gfx950).
// ---------------------------------------------------------------------------
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename TM> struct Mfma;
template <> struct Mfma<double> {
  using acc_t = d4_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  using acc_t = f4_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename TA, typename TB, typename TC>
using mfma_t = typename std::conditional<std::is_same<TA, double>::value ||
                                             std::is_same<TB, double>::value ||
                                             std::is_same<TC, double>::value,
                                         double, float>::type;

// K depth staged per LDS round: 32 for FP32 tiles, 16 for FP64 ones, so that
// both take ~17 KB of LDS per workgroup (occupancy); GPT = elements of one
// 64-wide operand tile per thread.  D = K tiles whose global loads are in
// flight at once (a ring of registers, raw operand types): 4 for FP64 tiles,
// whose K = 64 products then issue every load before the first MFMA, 1 for
// FP32 ones (twice the elements per tile; a deeper ring halves their
// occupancy).
template <typename TM> struct GemmDepth {
  static constexpr int K = sizeof(TM) == 4 ? 32 : 16;
  static constexpr int D = sizeof(TM) == 4 ? 1 : 4;
};

// C_part[z][i][j] = sum_{k in chunk z} A[k][i] * B[k][j]
// Tile t + D is loaded into the ring slot of tile t while tile t is in the
// MFMAs, so D tiles of global loads overlap the arithmetic.
template <typename TA, typename TB, typename TM>
__global__ __launch_bounds__(256) void gemm_tn_kernel(int n, int m, int b, int kchunk,
                                                      const TA* __restrict__ A, int lda,
                                                      const TB* __restrict__ B, int ldb,
                                                      TM* __restrict__ part) {
  constexpr int GKT = GemmDepth<TM>::K, D = GemmDepth<TM>::D;
  constexpr int GPT = GKT * 64 / 256;
  __shared__ TM As[GKT][64 + 4];
  __shared__ TM Bs[GKT][64 + 4];
  using acc_t = typename Mfma<TM>::acc_t;
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64, z = blockIdx.z;
  const int kb = z * kchunk;
  const int ke = min(n, kb + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = acc_t{0, 0, 0, 0};
  TA ra[D][GPT];
  TB rb[D][GPT];
  auto load = [&](TA* xa, TB* xb, int k0) {
#pragma unroll
    for (int q = 0; q < GPT; ++q) {
      const int e = tid + 256 * q;
      const int kk = e >> 6, ii = e & 63;
      const int k = k0 + kk;
      xa[q] = (k < ke && i0 + ii < m) ? A[(size_t)k * lda + i0 + ii] : TA(0);
      xb[q] = (k < ke && j0 + ii < b) ? B[(size_t)k * ldb + j0 + ii] : TB(0);
    }
  };
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (kb + d * GKT < ke) load(ra[d], rb[d], kb + d * GKT);
  for (int k0 = kb; k0 < ke; k0 += D * GKT) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int kt = k0 + d * GKT;
      if (kt < ke) {  // the same in every thread
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
          const int e = tid + 256 * q;
          As[e >> 6][e & 63] = (TM)ra[d][q];
          Bs[e >> 6][e & 63] = (TM)rb[d][q];
        }
        __syncthreads();
        if (kt + D * GKT < ke) load(ra[d], rb[d], kt + D * GKT);
#pragma unroll
        for (int kk = 0; kk < GKT; kk += 4) {
          const int kr = kk + (lane >> 4), cl = lane & 15;
          const TM a0 = As[kr][wi + cl], a1 = As[kr][wi + 16 + cl];
          const TM b0 = Bs[kr][wj + cl], b1 = Bs[kr][wj + 16 + cl];
          acc[0][0] = Mfma<TM>::op(a0, b0, acc[0][0]);
          acc[0][1] = Mfma<TM>::op(a0, b1, acc[0][1]);
          acc[1][0] = Mfma<TM>::op(a1, b0, acc[1][0]);
          acc[1][1] = Mfma<TM>::op(a1, b1, acc[1][1]);
        }
        __syncthreads();
      }
    }
  }
  TM* P = part + (size_t)z * m * b;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi + ti * 16 + Mfma<TM>::drow(lane, r);
        const int j = j0 + wj + tj * 16 + (lane & 15);
        if (i < m && j < b) P[(size_t)i * b + j] = acc[ti][tj][r];
      }
}

// Sum of the split-K partials.  A workgroup covers 256/zg consecutive elements
// with zg threads per element; thread group g sums partials g, g+zg, ... with
// two independent accumulators, then group sums combine in LDS in a fixed
// order (deterministic).  Small Gram matrices have few elements and many
// partials, so the partials, not the elements, carry the parallelism there.
template <typename TM, typename TC>
__global__ __launch_bounds__(256) void gemm_tn_reduce(int m, int b, int nsplit, int zg,
                                                      const TM* __restrict__ part, double alpha,
                                                      double beta, TC* C, int ldc) {
  __shared__ double red[256];
  const int epb = 256 / zg;
  const int le = threadIdx.x % epb, g = threadIdx.x / epb;
  const size_t tot = (size_t)m * b;
  const size_t idx = (size_t)blockIdx.x * epb + le;
  double s0 = 0.0, s1 = 0.0;
  if (idx < tot) {
    int z = g;
    for (; z + zg < nsplit; z += 2 * zg) {
      s0 += (double)part[(size_t)z * tot + idx];
      s1 += (double)part[(size_t)(z + zg) * tot + idx];
    }
    if (z < nsplit) s0 += (double)part[(size_t)z * tot + idx];
  }
  red[threadIdx.x] = s0 + s1;
  __syncthreads();
  if (g != 0 || idx >= tot) return;
  double s = 0.0;
  for (int q = 0; q < zg; ++q) s += red[q * epb + le];
  const int i = (int)(idx / b), j = (int)(idx % b);
  double v = alpha * s;
  if (beta != 0.0) v += beta * (double)C[(size_t)i * ldc + j];
  C[(size_t)i * ldc + j] = (TC)v;
}

static int gemm_tn_nsplit(int n, int m, int b) {
  const int tiles = ((m + 63) / 64) * ((b + 63) / 64);
  const int want = std::max(1, (4 * device_cu_count() + tiles - 1) / tiles);
  const int maxs = std::max(1, n / 256);
  return std::max(1, std::min(want, maxs));
}

size_t gemm_tn_work_elems(int n, int m, int b) {
  return (size_t)gemm_tn_nsplit(n, m, b) * (size_t)m * (size_t)b;
}

template <typename TA, typename TB, typename TC>
void gemm_tn(hipStream_t st, int n, int m, int b, double alpha, const TA* A, int lda, const TB* B,
             int ldb, double beta, TC* C, int ldc, double* work, size_t work_elems) {
  if (m == 0 || b == 0) return;
  using TM = mfma_t<TA, TB, TC>;
  const int nsplit = gemm_tn_nsplit(n, m, b);
  const size_t need = (size_t)nsplit * m * b;
  if (need > work_elems) throw Error(-4, "gemm_tn workspace too small");
  int kchunk = (n + nsplit - 1) / nsplit;
  constexpr int GKT = GemmDepth<TM>::K;
  kchunk = (kchunk + GKT - 1) / GKT * GKT;
  TM* part = reinterpret_cast<TM*>(work);
  if (n > 0) {
    dim3 grid((m + 63) / 64, (b + 63) / 64, nsplit);
    hipLaunchKernelGGL((gemm_tn_kernel<TA, TB, TM>), grid, dim3(256), 0, st, n, m, b, kchunk, A,
                       lda, B, ldb, part);
    SMALGO_HIP(hipGetLastError());
  } else {
    SMALGO_HIP(hipMemsetAsync(part, 0, sizeof(TM) * (size_t)m * b, st));
  }
  const size_t tot = (size_t)m * b;
  const int ns = n > 0 ? nsplit : 1;
  const int zg = ns >= 16 ? 16 : (ns >= 4 ? 4 : 1);  // threads per element
  const size_t epb = 256 / zg;
  hipLaunchKernelGGL((gemm_tn_reduce<TM, TC>), dim3((unsigned)((tot + epb - 1) / epb)), dim3(256), 0,
                     st, m, b, ns, zg, part, alpha, beta, C, ldc);
  SMALGO_HIP(hipGetLastError());
}

// C[i][j] = alpha * sum_k A[i][k] B(k, j) + beta * C[i][j]
// K tiles go through the same register ring as in gemm_tn_kernel; with
// FP64 tiles and beta != 0 the C elements are loaded before the K loop, so
// their latency overlaps the arithmetic too.
template <typename TA, typename TB, typename TC, typename TM, bool TRANSB>
__global__ __launch_bounds__(256) void gemm_nn_kernel(int n, int m, int b, double alpha,
                                                      const TA* __restrict__ A, int lda,
                                                      const TB* __restrict__ B, int ldb,
                                                      double beta, TC* C, int ldc) {
  constexpr int GKT = GemmDepth<TM>::K, D = GemmDepth<TM>::D;
  constexpr int GPT = GKT * 64 / 256;
  __shared__ TM As[64][GKT + 1];
  __shared__ TM Bs[GKT][64 + 4];
  using acc_t = typename Mfma<TM>::acc_t;
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = acc_t{0, 0, 0, 0};
  constexpr bool early_c = D > 1;
  TC cold[2][2][4];
  if (early_c && beta != 0.0) {
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = r0 + wi + ti * 16 + Mfma<TM>::drow(lane, r);
          const int j = c0 + wj + tj * 16 + (lane & 15);
          cold[ti][tj][r] = (i < n && j < b) ? C[(size_t)i * ldc + j] : TC(0);
        }
  }
  TA ra[D][GPT];
  TB rb[D][GPT];
  auto load = [&](TA* xa, TB* xb, int k0) {
#pragma unroll
    for (int q = 0; q < GPT; ++q) {
      const int e = tid + 256 * q;
      const int rr = e / GKT, kk = e % GKT;
      const int r = r0 + rr, k = k0 + kk;
      xa[q] = (r < n && k < m) ? A[(size_t)r * lda + k] : TA(0);
      const int kb = e >> 6, cc = e & 63;
      const int k2 = k0 + kb, c = c0 + cc;
      TB v = TB(0);
      if (k2 < m && c < b) v = TRANSB ? B[(size_t)c * ldb + k2] : B[(size_t)k2 * ldb + c];
      xb[q] = v;
    }
  };
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (d * GKT < m) load(ra[d], rb[d], d * GKT);
  for (int k0 = 0; k0 < m; k0 += D * GKT) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int kt = k0 + d * GKT;
      if (kt < m) {
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
          const int e = tid + 256 * q;
          As[e / GKT][e % GKT] = (TM)ra[d][q];
          Bs[e >> 6][e & 63] = (TM)rb[d][q];
        }
        __syncthreads();
        if (kt + D * GKT < m) load(ra[d], rb[d], kt + D * GKT);
#pragma unroll
        for (int kk = 0; kk < GKT; kk += 4) {
          const int kr = kk + (lane >> 4), cl = lane & 15;
          const TM a0 = As[wi + cl][kr], a1 = As[wi + 16 + cl][kr];
          const TM b0 = Bs[kr][wj + cl], b1 = Bs[kr][wj + 16 + cl];
          acc[0][0] = Mfma<TM>::op(a0, b0, acc[0][0]);
          acc[0][1] = Mfma<TM>::op(a0, b1, acc[0][1]);
          acc[1][0] = Mfma<TM>::op(a1, b0, acc[1][0]);
          acc[1][1] = Mfma<TM>::op(a1, b1, acc[1][1]);
        }
        __syncthreads();
      }
    }
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = r0 + wi + ti * 16 + Mfma<TM>::drow(lane, r);
        const int j = c0 + wj + tj * 16 + (lane & 15);
        if (i < n && j < b) {
          double v = alpha * (double)acc[ti][tj][r];
          if (beta != 0.0)
            v += beta * (double)(early_c ? cold[ti][tj][r] : C[(size_t)i * ldc + j]);
          C[(size_t)i * ldc + j] = (TC)v;
        }
      }
}

template <typename TA, typename TB, typename TC>
void gemm_nn(hipStream_t st, int n, int m, int b, double alpha, const TA* A, int lda, const TB* B,
             int ldb, bool transB, double beta, TC* C, int ldc) {
  if (n == 0 || b == 0) return;
  using TM = mfma_t<TA, TB, TC>;
  dim3 grid((n + 63) / 64, (b + 63) / 64);
  if (transB)
    hipLaunchKernelGGL((gemm_nn_kernel<TA, TB, TC, TM, true>), grid, dim3(256), 0, st, n, m, b,
                       alpha, A, lda, B, ldb, beta, C, ldc);
  else
    hipLaunchKernelGGL((gemm_nn_kernel<TA, TB, TC, TM, false>), grid, dim3(256), 0, st, n, m, b,
                       alpha, A, lda, B, ldb, beta, C, ldc);
  SMALGO_HIP(hipGetLastError());
}

#define SMALGO_GEMM_INST(TA, TB, TC)                                                          \
  template void gemm_tn<TA, TB, TC>(hipStream_t, int, int, int, double, const TA*, int,       \
                                    const TB*, int, double, TC*, int, double*, size_t);       \
  template void gemm_nn<TA, TB, TC>(hipStream_t, int, int, int, double, const TA*, int,       \
                                    const TB*, int, bool, double, TC*, int);
SMALGO_GEMM_INST(double, double, double)
SMALGO_GEMM_INST(float, float, float)
SMALGO_GEMM_INST(float, float, double)
SMALGO_GEMM_INST(float, double, double)
SMALGO_GEMM_INST(double, float, double)
SMALGO_GEMM_INST(float, double, float)
#undef SMALGO_GEMM_INST

// ---------------------------------------------------------------------------
// column reductions
// ---------------------------------------------------------------------------
static void colred_shape(int n, int b, int* nblk, int* rows_per_blk, int* nrt) {
  const int nct = std::min(b, 256);
  *nrt = std::max(1, 256 / std::max(nct, 1));
  // at most ~2048 partial sums per column
  int want = std::max(1, std::min(4 * device_cu_count(), 2048 / *nrt));
  int rpb = std::max(64, (n + want - 1) / want);
  *rows_per_blk = rpb;
  *nblk = std::max(1, (n + rpb - 1) / rpb);
}

size_t colred_work_elems(int n, int b) {
  int nblk, rpb, nrt;
  colred_shape(n, b, &nblk, &rpb, &nrt);
  return (size_t)nblk * nrt * b;
}

// mode 0: sum x*y; mode 1: sum (y - lam*x)^2
template <typename TX, int MODE>
__global__ __launch_bounds__(256) void colred_kernel(int n, int b, int rows_per_blk, int nrt,
                                                     const TX* __restrict__ X, int ldx,
                                                     const TX* __restrict__ Y, int ldy,
                                                     const double* __restrict__ lam,
                                                     double* __restrict__ part) {
  const int nct = min(b, 256);
  const int tid = threadIdx.x;
  const int rs = tid / nct;
  if (rs >= nrt) return;
  const int r0 = blockIdx.x * rows_per_blk;
  const int r1 = min(n, r0 + rows_per_blk);
  auto term = [&](int i, int c, double l) {
    const double x = (double)X[(size_t)i * ldx + c];
    const double y = (double)Y[(size_t)i * ldy + c];
    if (MODE == 0) return x * y;
    const double d = y - l * x;
    return d * d;
  };
  for (int c = tid % nct; c < b; c += nct) {
    const double l = MODE == 1 ? lam[c] : 0.0;
    // two independent accumulators keep two rows' loads in flight
    double s0 = 0.0, s1 = 0.0;
    int i = r0 + rs;
    for (; i + nrt < r1; i += 2 * nrt) {
      s0 += term(i, c, l);
      s1 += term(i + nrt, c, l);
    }
    if (i < r1) s0 += term(i, c, l);
    part[((size_t)blockIdx.x * nrt + rs) * b + c] = s0 + s1;
  }
}

// one workgroup per group of 16 columns: 16 threads per column stride over the
// partial sums with four independent accumulators (so four loads are in
// flight per thread), then a fixed-order tree in LDS (deterministic)
template <int MODE>
__global__ __launch_bounds__(256) void colred_finish(int b, int nparts,
                                                     const double* __restrict__ part, double* out) {
  __shared__ double red[16][17];
  const int cl = threadIdx.x & 15, pl = threadIdx.x >> 4;
  const int c = blockIdx.x * 16 + cl;
  double s = 0.0;
  if (c < b) {
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    int p = pl;
    for (; p + 48 < nparts; p += 64) {
#pragma unroll
      for (int k = 0; k < 4; ++k) s4[k] += part[(size_t)(p + 16 * k) * b + c];
    }
    for (int k = 0; p < nparts; p += 16, ++k) s4[k & 3] += part[(size_t)p * b + c];
    s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
  }
  red[pl][cl] = s;
  __syncthreads();
  for (int h = 8; h > 0; h >>= 1) {
    if (pl < h) red[pl][cl] += red[pl + h][cl];
    __syncthreads();
  }
  if (pl == 0 && c < b) out[c] = MODE == 1 ? sqrt(red[0][cl]) : red[0][cl];
}

template <typename TX, int MODE>
static void colred(hipStream_t st, int n, int b, const TX* X, int ldx, const TX* Y, int ldy,
                   const double* lam, double* out, double* work, size_t work_elems) {
  if (b == 0) return;
  int nblk, rpb, nrt;
  colred_shape(n, b, &nblk, &rpb, &nrt);
  if ((size_t)nblk * nrt * b > work_elems) throw Error(-4, "column reduction workspace too small");
  hipLaunchKernelGGL((colred_kernel<TX, MODE>), dim3(nblk), dim3(256), 0, st, n, b, rpb, nrt, X,
                     ldx, Y, ldy, lam, work);
  SMALGO_HIP(hipGetLastError());
  hipLaunchKernelGGL((colred_finish<MODE>), dim3((b + 15) / 16), dim3(256), 0, st, b, nblk * nrt,
                     work, out);
  SMALGO_HIP(hipGetLastError());
}

template <typename TX>
void col_dots(hipStream_t st, int n, int b, const TX* X, int ldx, const TX* Y, int ldy, double* out,
              double* work, size_t work_elems) {
  colred<TX, 0>(st, n, b, X, ldx, Y, ldy, nullptr, out, work, work_elems);
}
template <typename TX>
void col_resid(hipStream_t st, int n, int b, const TX* X, int ldx, const TX* AX, int ldax,
               const double* lam, double* out, double* work, size_t work_elems) {
  colred<TX, 1>(st, n, b, X, ldx, AX, ldax, lam, out, work, work_elems);
}
template void col_dots<double>(hipStream_t, int, int, const double*, int, const double*, int,
                               double*, double*, size_t);
template void col_dots<float>(hipStream_t, int, int, const float*, int, const float*, int, double*,
                              double*, size_t);
template void col_resid<double>(hipStream_t, int, int, const double*, int, const double*, int,
                                const double*, double*, double*, size_t);
template void col_resid<float>(hipStream_t, int, int, const float*, int, const float*, int,
                               const double*, double*, double*, size_t);

// ---------------------------------------------------------------------------
// block helpers
// ---------------------------------------------------------------------------
template <typename TS, typename TD>
__global__ void copy_block_kernel(int n, int b, const TS* __restrict__ S, int lds, TD* __restrict__ D,
                                  int ldd) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * b) return;
  const int i = (int)(idx / b), j = (int)(idx % b);
  D[(size_t)i * ldd + j] = (TD)S[(size_t)i * lds + j];
}
template <typename TS, typename TD>
void copy_block(hipStream_t st, int n, int b, const TS* S, int lds, TD* D, int ldd) {
  const size_t tot = (size_t)n * b;
  if (!tot) return;
  hipLaunchKernelGGL((copy_block_kernel<TS, TD>), dim3((tot + 255) / 256), dim3(256), 0, st, n, b,
                     S, lds, D, ldd);
  SMALGO_HIP(hipGetLastError());
}
template void copy_block<double, double>(hipStream_t, int, int, const double*, int, double*, int);
template void copy_block<float, float>(hipStream_t, int, int, const float*, int, float*, int);
template void copy_block<double, float>(hipStream_t, int, int, const double*, int, float*, int);
template void copy_block<float, double>(hipStream_t, int, int, const float*, int, double*, int);

template <typename T>
__global__ void fill_block_kernel(int n, int b, T* D, int ldd, T v) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * b) return;
  D[(size_t)(idx / b) * ldd + idx % b] = v;
}
template <typename T>
void fill_block(hipStream_t st, int n, int b, T* D, int ldd, T v) {
  const size_t tot = (size_t)n * b;
  if (!tot) return;
  hipLaunchKernelGGL((fill_block_kernel<T>), dim3((tot + 255) / 256), dim3(256), 0, st, n, b, D,
                     ldd, v);
  SMALGO_HIP(hipGetLastError());
}
template void fill_block<double>(hipStream_t, int, int, double*, int, double);
template void fill_block<float>(hipStream_t, int, int, float*, int, float);

template <typename T>
__global__ void scale_cols_kernel(int n, int b, T* D, int ldd, const double* s) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * b) return;
  const int j = (int)(idx % b);
  T* p = D + (size_t)(idx / b) * ldd + j;
  *p = (T)((double)*p * s[j]);
}
template <typename T>
void scale_cols(hipStream_t st, int n, int b, T* D, int ldd, const double* s) {
  const size_t tot = (size_t)n * b;
  if (!tot) return;
  hipLaunchKernelGGL((scale_cols_kernel<T>), dim3((tot + 255) / 256), dim3(256), 0, st, n, b, D,
                     ldd, s);
  SMALGO_HIP(hipGetLastError());
}
template void scale_cols<double>(hipStream_t, int, int, double*, int, const double*);
template void scale_cols<float>(hipStream_t, int, int, float*, int, const double*);

__global__ void gather_cols_kernel(int M, int m, const double* __restrict__ S, int lds,
                                   const int* __restrict__ sel, const int* __restrict__ rowmap,
                                   double* __restrict__ D, int ldd) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)M * m) return;
  const int i = (int)(idx / m), j = (int)(idx % m);
  const int dr = rowmap ? rowmap[i] : i;
  D[(size_t)dr * ldd + j] = S[(size_t)i + (size_t)sel[j] * lds];
}
void gather_cols_colmajor(hipStream_t st, int M, int m, const double* S, int lds, const int* sel,
                          const int* rowmap, double* D, int ldd) {
  const size_t tot = (size_t)M * m;
  if (!tot) return;
  hipLaunchKernelGGL(gather_cols_kernel, dim3((tot + 255) / 256), dim3(256), 0, st, M, m, S, lds,
                     sel, rowmap, D, ldd);
  SMALGO_HIP(hipGetLastError());
}

__global__ void symmetrize_kernel(int b, double* C, int ldc) {
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b || j >= b || j <= i) return;
  const double v = 0.5 * (C[(size_t)i * ldc + j] + C[(size_t)j * ldc + i]);
  C[(size_t)i * ldc + j] = v;
  C[(size_t)j * ldc + i] = v;
}
void symmetrize(hipStream_t st, int b, double* C, int ldc) {
  if (b <= 1) return;
  dim3 blk(16, 16), grid((b + 15) / 16, (b + 15) / 16);
  hipLaunchKernelGGL(symmetrize_kernel, grid, blk, 0, st, b, C, ldc);
  SMALGO_HIP(hipGetLastError());
}

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ void gaussian_kernel(int n, int b, uint64_t seed, T* D, int ldd) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * b) return;
  const int i = (int)(idx % n), j = (int)(idx / n);  // column-major counter j*n + i
  const double u1 = (double)(splitmix_at(seed, idx) >> 11) * (1.0 / 9007199254740992.0);
  const double u2 =
      (double)(splitmix_at(seed ^ 0x5DEECE66Dull, idx) >> 11) * (1.0 / 9007199254740992.0);
  const double g = sqrt(-2.0 * log1p(-u1)) * cos(6.283185307179586 * u2);
  D[(size_t)i * ldd + j] = (T)g;
}
template <typename T>
void gaussian_block(hipStream_t st, int n, int b, uint64_t seed, T* D, int ldd) {
  const size_t tot = (size_t)n * b;
  if (!tot) return;
  hipLaunchKernelGGL((gaussian_kernel<T>), dim3((tot + 255) / 256), dim3(256), 0, st, n, b, seed, D,
                     ldd);
  SMALGO_HIP(hipGetLastError());
}
template void gaussian_block<double>(hipStream_t, int, int, uint64_t, double*, int);
template void gaussian_block<float>(hipStream_t, int, int, uint64_t, float*, int);

// ---------------------------------------------------------------------------
// layout conversion between the caller's column-major arrays and the
// device's row-major blocks (32x32 tiles through LDS)
// ---------------------------------------------------------------------------
template <typename TS, typename TD>
__global__ void cm_to_rm_kernel(int n, int b, const TS* __restrict__ S, int lds, TD* __restrict__ D,
                                int ldd) {
  __shared__ TD tileb[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int jj = ty; jj < 32; jj += 8) {
    const int i = i0 + tx, j = j0 + jj;
    if (i < n && j < b) tileb[jj][tx] = (TD)S[(size_t)j * lds + i];
  }
  __syncthreads();
  for (int ii = ty; ii < 32; ii += 8) {
    const int i = i0 + ii, j = j0 + tx;
    if (i < n && j < b) D[(size_t)i * ldd + j] = tileb[tx][ii];
  }
}
template <typename TS, typename TD>
void cm_to_rm(hipStream_t st, int n, int b, const TS* S, int lds, TD* D, int ldd) {
  if (!n || !b) return;
  dim3 grid((n + 31) / 32, (b + 31) / 32), blk(32, 8);
  hipLaunchKernelGGL((cm_to_rm_kernel<TS, TD>), grid, blk, 0, st, n, b, S, lds, D, ldd);
  SMALGO_HIP(hipGetLastError());
}
template void cm_to_rm<double, double>(hipStream_t, int, int, const double*, int, double*, int);
template void cm_to_rm<double, float>(hipStream_t, int, int, const double*, int, float*, int);

// D (column-major, ld ldd)[:, j] = S (row-major, ld lds)[:, sel[j]]  (sel may be null)
template <typename TS>
__global__ void rm_to_cm_kernel(int n, int k, const TS* __restrict__ S, int lds,
                                const int* __restrict__ sel, double* __restrict__ D, int ldd) {
  __shared__ double tileb[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int ii = ty; ii < 32; ii += 8) {
    const int i = i0 + ii, j = j0 + tx;
    if (i < n && j < k) tileb[ii][tx] = (double)S[(size_t)i * lds + (sel ? sel[j] : j)];
  }
  __syncthreads();
  for (int jj = ty; jj < 32; jj += 8) {
    const int i = i0 + tx, j = j0 + jj;
    if (i < n && j < k) D[(size_t)j * ldd + i] = tileb[tx][jj];
  }
}
template <typename TS>
void rm_to_cm(hipStream_t st, int n, int k, const TS* S, int lds, const int* sel, double* D, int ldd) {
  if (!n || !k) return;
  dim3 grid((n + 31) / 32, (k + 31) / 32), blk(32, 8);
  hipLaunchKernelGGL((rm_to_cm_kernel<TS>), grid, blk, 0, st, n, k, S, lds, sel, D, ldd);
  SMALGO_HIP(hipGetLastError());
}
template void rm_to_cm<double>(hipStream_t, int, int, const double*, int, const int*, double*, int);
template void rm_to_cm<float>(hipStream_t, int, int, const float*, int, const int*, double*, int);

// D[i, j] = S[i, sel[j]]  (row-major both)
template <typename T>
__global__ void gather_cols_rm_kernel(int n, int k, const T* __restrict__ S, int lds,
                                      const int* __restrict__ sel, T* __restrict__ D, int ldd) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * k) return;
  const int i = (int)(idx / k), j = (int)(idx % k);
  D[(size_t)i * ldd + j] = S[(size_t)i * lds + sel[j]];
}
template <typename T>
void gather_cols_rm(hipStream_t st, int n, int k, const T* S, int lds, const int* sel, T* D, int ldd) {
  const size_t tot = (size_t)n * k;
  if (!tot) return;
  hipLaunchKernelGGL((gather_cols_rm_kernel<T>), dim3((tot + 255) / 256), dim3(256), 0, st, n, k, S,
                     lds, sel, D, ldd);
  SMALGO_HIP(hipGetLastError());
}
template void gather_cols_rm<double>(hipStream_t, int, int, const double*, int, const int*, double*, int);
template void gather_cols_rm<float>(hipStream_t, int, int, const float*, int, const int*, float*, int);

__global__ void add_diag_kernel(int b, double* C, int ldc, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b) C[(size_t)i * ldc + i] += v;
}
void add_diag(hipStream_t st, int b, double* C, int ldc, double v) {
  if (b <= 0) return;
  hipLaunchKernelGGL(add_diag_kernel, dim3((b + 255) / 256), dim3(256), 0, st, b, C, ldc, v);
  SMALGO_HIP(hipGetLastError());
}

// C = A·B for upper-triangular column-major b x b matrices (ld b): entry
// (i, j), j >= i, is the sum over k in [i, j] of A(i,k)·B(k,j); C may not
// alias A or B.
__global__ void tri_mul_upper_kernel(int b, const double* __restrict__ A,
                                     const double* __restrict__ B, double* __restrict__ C) {
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b || j >= b) return;
  double acc = 0.0;
  for (int k = i; k <= j; ++k) acc += A[(size_t)k * b + i] * B[(size_t)j * b + k];
  C[(size_t)j * b + i] = acc;  // zero below the diagonal (empty sum)
}
void tri_mul_upper(hipStream_t st, int b, const double* A, const double* B, double* C) {
  if (b <= 0) return;
  dim3 blk(16, 16), grid((b + 15) / 16, (b + 15) / 16);
  hipLaunchKernelGGL(tri_mul_upper_kernel, grid, blk, 0, st, b, A, B, C);
  SMALGO_HIP(hipGetLastError());
}

// Upper Cholesky factor and its inverse of a symmetric positive definite
// b x b matrix in one workgroup: G = RᵀR (upper triangle of G read,
// column-major) and Ri = R⁻¹, both written in full with zeros below the
// diagonal.  Thread (ty, tx) of a TG x TG grid holds rows 4ty..4ty+3 and
// columns 4tx..4tx+3 of the matrix under elimination A and of Y = R⁻ᵀ.
// Step k publishes row k of both, which is final by then, and updates the
// rows below it:
//   A_i· -= (A_ki / A_kk) A_k·,   Y_i· -= (A_ki / √A_kk) (Y_k· / √A_kk),
// so that row k of R is A_k· / √A_kk and Y_k· / √A_kk is row k of R⁻ᵀ, i.e.
// column k of R⁻¹.  Rows at or above k take a zero multiplier, so they keep
// their final values; entries left of the diagonal of A are updated too but
// never read again.  One barrier per step (the published rows are double
// buffered); R and Ri are written from registers at the end.  info = k + 1
// at the first pivot that is not positive (NaN included), as potrf reports
// it, else 0; R and Ri are then not written.  G must not alias R or Ri.
template <int TG>
__global__ __launch_bounds__(TG * TG) void chol_inv_upper_kernel(
    int b, const double* __restrict__ G, int ldg, double* __restrict__ R, int ldr,
    double* __restrict__ Ri, int ldri, int* __restrict__ info) {
  constexpr int NB = 4;
  __shared__ double arow[2][NB * TG], yrow[2][NB * TG];
  const int tid = threadIdx.x, ty = tid / TG, tx = tid % TG;
  const int r0 = ty * NB, c0 = tx * NB;
  double a[NB][NB], y[NB][NB], sk[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    sk[r] = 1.0;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int i = r0 + r, j = c0 + c;
      double v = i == j ? 1.0 : 0.0;
      if (i < b && j < b) v = i <= j ? G[(size_t)j * ldg + i] : G[(size_t)i * ldg + j];
      a[r][c] = v;
      y[r][c] = i == j ? 1.0 : 0.0;
    }
  }
  int bad = 0;
  for (int kb = 0; kb * NB < b && bad == 0; ++kb) {
#pragma unroll
    for (int kr = 0; kr < NB; ++kr) {
      const int k = kb * NB + kr;
      if (k >= b) break;
      const int buf = k & 1;
      if (ty == kb) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          arow[buf][c0 + c] = a[kr][c];
          yrow[buf][c0 + c] = y[kr][c];
        }
      }
      __syncthreads();
      const double piv = arow[buf][k];  // the same value in every thread
      if (!(piv > 0.0)) {
        bad = k + 1;
        break;
      }
      const double s = rsqrt(piv), ip = s * s;
      if (ty == kb) sk[kr] = s;
      double ac[NB], yk[NB], ak[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        ac[c] = arow[buf][c0 + c];
        yk[c] = yrow[buf][c0 + c] * s;
        ak[c] = arow[buf][r0 + c];
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const bool below = r0 + r > k;
        const double f = below ? ak[r] * ip : 0.0, l = below ? ak[r] * s : 0.0;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          a[r][c] -= f * ac[c];
          y[r][c] -= l * yk[c];
        }
      }
    }
  }
  if (tid == 0) *info = bad;
  if (bad != 0) return;
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    const int i = r0 + r;
    if (i >= b) continue;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int j = c0 + c;
      if (j >= b) continue;
      R[(size_t)j * ldr + i] = j >= i ? a[r][c] * sk[r] : 0.0;
      Ri[(size_t)i * ldri + j] = y[r][c] * sk[r];  // (R⁻¹)_ji = Y_ij, zero for j > i
    }
  }
}

int chol_inv_max_b() { return 128; }

void chol_inv_upper(hipStream_t st, int b, const double* G, int ldg, double* R, int ldr,
                    double* Ri, int ldri, int* info) {
  if (b < 1 || b > chol_inv_max_b()) throw Error(-1, "chol_inv_upper: b out of range");
  if (b <= 64)
    hipLaunchKernelGGL(chol_inv_upper_kernel<16>, dim3(1), dim3(256), 0, st, b, G, ldg, R, ldr,
                       Ri, ldri, info);
  else
    hipLaunchKernelGGL(chol_inv_upper_kernel<32>, dim3(1), dim3(1024), 0, st, b, G, ldg, R, ldr,
                       Ri, ldri, info);
  SMALGO_HIP(hipGetLastError());
}

// out = [info[0], info[1], R(0,0), R(1,1), ..., R(b-1,b-1)] (R column-major, ld b)
__global__ void pack_qr_status_kernel(int b, const int* __restrict__ info,
                                      const double* __restrict__ R, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2) out[i] = (double)info[i];
  if (i < b) out[2 + i] = R[(size_t)i * b + i];
}
void pack_qr_status(hipStream_t st, int b, const int* info, const double* R, double* out) {
  hipLaunchKernelGGL(pack_qr_status_kernel, dim3((b + 2 + 255) / 256), dim3(256), 0, st, b, info, R,
                     out);
  SMALGO_HIP(hipGetLastError());
}

// Bt(i, c) = R(i, c) for i < rn, else 0: the s x s device block (column-major,
// ld s) that the next block step reads as B_jᵀ, from R (r x r, ld r)
__global__ void b_block_kernel(int s, int r, int rn, const double* __restrict__ R,
                               double* __restrict__ Bt) {
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s || c >= s) return;
  Bt[(size_t)c * s + i] = (i < rn && c < r) ? R[(size_t)c * r + i] : 0.0;
}
void b_block(hipStream_t st, int s, int r, int rn, const double* R, double* Bt) {
  dim3 blk(16, 16), grid((s + 15) / 16, (s + 15) / 16);
  hipLaunchKernelGGL(b_block_kernel, grid, blk, 0, st, s, r, rn, R, Bt);
  SMALGO_HIP(hipGetLastError());
}

// zero the strictly-lower triangle of a column-major b x b matrix (LAPACK
// routines leave the unused triangle untouched)
__global__ void zero_lower_cm_kernel(int b, double* C, int ldc) {
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b && j < b && i > j) C[(size_t)j * ldc + i] = 0.0;
}
void zero_lower_cm(hipStream_t st, int b, double* C, int ldc) {
  if (b <= 1) return;
  dim3 blk(16, 16), grid((b + 15) / 16, (b + 15) / 16);
  hipLaunchKernelGGL(zero_lower_cm_kernel, grid, blk, 0, st, b, C, ldc);
  SMALGO_HIP(hipGetLastError());
}

// plain streaming copy (16-byte vectors, grid-stride) for bandwidth calibration
// Each workgroup copies one contiguous 16 KiB chunk per trip: every lane
// issues its four 16-byte loads (256 lanes apart, so each is coalesced)
// before its four stores, keeping four loads per lane in flight.
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4* __restrict__ a,
                                                          float4* __restrict__ b, size_t n4) {
  constexpr int U = 4;
  const size_t per = (size_t)U * 256;
  for (size_t base = (size_t)blockIdx.x * per; base < n4; base += (size_t)gridDim.x * per) {
    float4 v[U];
    if (base + per <= n4) {
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = a[base + (size_t)u * 256 + threadIdx.x];
#pragma unroll
      for (int u = 0; u < U; ++u) b[base + (size_t)u * 256 + threadIdx.x] = v[u];
    } else {
      for (int u = 0; u < U; ++u) {
        const size_t i = base + (size_t)u * 256 + threadIdx.x;
        if (i < n4) b[i] = a[i];
      }
    }
  }
}
void stream_copy(hipStream_t st, const void* src, void* dst, size_t bytes) {
  const size_t n4 = bytes / 16;
  if (!n4) return;
  // one trip per workgroup up to 32 workgroups per CU, grid-stride beyond
  const size_t chunks = (n4 + 1023) / 1024;
  const int grid = (int)std::min<size_t>(chunks, (size_t)32 * device_cu_count());
  hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, st,
                     reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
  SMALGO_HIP(hipGetLastError());
}

int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  SMALGO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) dev = 0;
  if (!cached[dev]) {
    int cu = 0;
    SMALGO_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    cached[dev] = cu > 0 ? cu : 256;
  }
  return cached[dev];
}

}  // namespace smalgo
