// On-disk fixtures shared by the C ABI, Python and Julia: a CSR matrix and a
// (λ, residual) set, little-endian, with a 64-byte header and a 64-bit
// checksum of the payload: Σ_i (w_i xor i·K1)·K2 mod 2^64 over the payload's
// 8-byte words w_i (K1 = 0x9E3779B97F4A7C15, K2 = 0xBF58476D1CE4E5B9), so
// NumPy computes it vectorised.  Host code only; no reference counterpart (the
// reference keeps its fixtures as source code and RESULTS comments).
//
// CSR file:  "SMALGO-CSR-v1\0\0\0" | int64 n | int64 nnz | int32 index_base (0) |
//            int32 flags (bit 0: symmetric) | uint64 checksum | 16 zero bytes |
//            int64 rowptr[n+1] | int64 col[nnz] | float64 val[nnz]
// Eig file:  "SMALGO-EIG-v1\0\0\0" | int64 count | int64 has_res (0/1) |
//            uint64 checksum | 24 zero bytes | float64 lam[count] | float64 res[count]?
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/smalgo.h"
#include "smalgo_internal.h"

namespace smalgo {

static const char kCsrMagic[16] = {'S', 'M', 'A', 'L', 'G', 'O', '-', 'C',
                                   'S', 'R', '-', 'v', '1', 0,   0,   0};
static const char kEigMagic[16] = {'S', 'M', 'A', 'L', 'G', 'O', '-', 'E',
                                   'I', 'G', '-', 'v', '1', 0,   0,   0};

struct Sum64 {
  uint64_t h = 0, i = 0;
  void add(const void* p, size_t nbytes) {  // nbytes is a multiple of 8
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t o = 0; o + 8 <= nbytes; o += 8, ++i) {
      uint64_t w;
      std::memcpy(&w, b + o, 8);
      h += (w ^ (i * 0x9E3779B97F4A7C15ull)) * 0xBF58476D1CE4E5B9ull;
    }
  }
};

struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) : f(std::fopen(path, mode)) {
    if (!f) throw Error(-1, std::string("cannot open ") + path);
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void write(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f) != n) throw Error(-1, "short write");
  }
  void read(void* p, size_t n) {
    if (n && std::fread(p, 1, n, f) != n) throw Error(-1, "file is truncated");
  }
};

struct CsrHeader {
  char magic[16];
  int64_t n, nnz;
  int32_t index_base, flags;
  uint64_t checksum;
  char reserved[16];
};
struct EigHeader {
  char magic[16];
  int64_t count, has_res;
  uint64_t checksum;
  char reserved[24];
};
static_assert(sizeof(CsrHeader) == 64 && sizeof(EigHeader) == 64, "64-byte headers");

void csr_save(const char* path, int64_t n, const int64_t* rowptr, const int64_t* col,
              const double* val, int index_base, int symmetric) {
  if (n < 0 || !rowptr) throw Error(-1, "bad arguments");
  const int64_t nnz = rowptr[n] - rowptr[0];
  if (nnz < 0 || (nnz && (!col || !val))) throw Error(-1, "bad arguments");
  // stored 0-based
  std::vector<int64_t> rp(rowptr, rowptr + n + 1), ci(col, col + nnz);
  const int64_t b0 = rowptr[0];
  for (auto& v : rp) v -= b0;
  if (index_base)
    for (auto& v : ci) v -= index_base;
  for (int64_t v : ci)
    if (v < 0 || v >= n) throw Error(-1, "column index out of range");
  CsrHeader h{};
  std::memcpy(h.magic, kCsrMagic, 16);
  h.n = n;
  h.nnz = nnz;
  h.flags = symmetric ? 1 : 0;
  Sum64 c;
  c.add(rp.data(), sizeof(int64_t) * rp.size());
  c.add(ci.data(), sizeof(int64_t) * ci.size());
  c.add(val, sizeof(double) * (size_t)nnz);
  h.checksum = c.h;
  File f(path, "wb");
  f.write(&h, sizeof h);
  f.write(rp.data(), sizeof(int64_t) * rp.size());
  f.write(ci.data(), sizeof(int64_t) * ci.size());
  f.write(val, sizeof(double) * (size_t)nnz);
}

static CsrHeader csr_header(File& f) {
  CsrHeader h{};
  f.read(&h, sizeof h);
  if (std::memcmp(h.magic, kCsrMagic, 16) != 0) throw Error(-1, "not a CSR fixture");
  if (h.n < 0 || h.nnz < 0 || h.index_base != 0) throw Error(-1, "bad CSR header");
  // the payload the header announces must be in the file: 8 (n+1) + 16 nnz bytes
  const long long limit = (long long)1 << 60;
  if (h.n > limit / 64 || h.nnz > limit / 64) throw Error(-1, "bad CSR header");
  const long long need = 64 + 8 * (h.n + 1) + 16 * h.nnz;
  if (std::fseek(f.f, 0, SEEK_END) != 0) throw Error(-1, "cannot size the file");
  const long long size = std::ftell(f.f);
  if (size < need) throw Error(-1, "file is truncated");
  if (std::fseek(f.f, (long)sizeof h, SEEK_SET) != 0) throw Error(-1, "cannot seek in the file");
  return h;
}

void csr_info(const char* path, int64_t* n, int64_t* nnz, int* symmetric) {
  File f(path, "rb");
  const CsrHeader h = csr_header(f);
  if (n) *n = h.n;
  if (nnz) *nnz = h.nnz;
  if (symmetric) *symmetric = h.flags & 1;
}

void csr_load(const char* path, int64_t* rowptr, int64_t* col, double* val) {
  File f(path, "rb");
  const CsrHeader h = csr_header(f);
  std::vector<int64_t> rp(h.n + 1), ci(h.nnz);
  std::vector<double> v(h.nnz);
  f.read(rp.data(), sizeof(int64_t) * rp.size());
  f.read(ci.data(), sizeof(int64_t) * ci.size());
  f.read(v.data(), sizeof(double) * v.size());
  Sum64 c;
  c.add(rp.data(), sizeof(int64_t) * rp.size());
  c.add(ci.data(), sizeof(int64_t) * ci.size());
  c.add(v.data(), sizeof(double) * v.size());
  if (c.h != h.checksum) throw Error(-1, "checksum mismatch");
  if (rp[0] != 0 || rp[h.n] != h.nnz) throw Error(-1, "bad row pointers");
  if (rowptr) std::memcpy(rowptr, rp.data(), sizeof(int64_t) * rp.size());
  if (col) std::memcpy(col, ci.data(), sizeof(int64_t) * ci.size());
  if (val) std::memcpy(val, v.data(), sizeof(double) * v.size());
}

void eig_save(const char* path, int64_t count, const double* lam, const double* res) {
  if (count < 0 || (count && !lam)) throw Error(-1, "bad arguments");
  EigHeader h{};
  std::memcpy(h.magic, kEigMagic, 16);
  h.count = count;
  h.has_res = res ? 1 : 0;
  Sum64 c;
  c.add(lam, sizeof(double) * (size_t)count);
  if (res) c.add(res, sizeof(double) * (size_t)count);
  h.checksum = c.h;
  File f(path, "wb");
  f.write(&h, sizeof h);
  f.write(lam, sizeof(double) * (size_t)count);
  if (res) f.write(res, sizeof(double) * (size_t)count);
}

void eig_load(const char* path, int64_t* count, int* has_res, double* lam, double* res) {
  File f(path, "rb");
  EigHeader h{};
  f.read(&h, sizeof h);
  if (std::memcmp(h.magic, kEigMagic, 16) != 0) throw Error(-1, "not an eigenvalue fixture");
  if (h.count < 0 || (h.has_res != 0 && h.has_res != 1)) throw Error(-1, "bad header");
  if (count) *count = h.count;
  if (has_res) *has_res = (int)h.has_res;
  if (!lam && !res) return;
  std::vector<double> l(h.count), r(h.has_res ? h.count : 0);
  f.read(l.data(), sizeof(double) * l.size());
  f.read(r.data(), sizeof(double) * r.size());
  Sum64 c;
  c.add(l.data(), sizeof(double) * l.size());
  c.add(r.data(), sizeof(double) * r.size());
  if (c.h != h.checksum) throw Error(-1, "checksum mismatch");
  if (lam) std::memcpy(lam, l.data(), sizeof(double) * l.size());
  if (res && h.has_res) std::memcpy(res, r.data(), sizeof(double) * r.size());
}

}  // namespace smalgo
