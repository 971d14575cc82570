// Binary files for CSR matrices and AMG hierarchies.
//
// Layout, little-endian, no padding:
//   CSR file:        "GSCSR001" (8 bytes), then one matrix record.
//   Hierarchy file:  "GSHIER01" (8 bytes), int64 levels L, int64 transfers T
//                    (L - 1 for a complete hierarchy), then L records of A,
//                    T records of P, T records of R, then int64 count of
//                    aggregate maps followed by, for each, int64 n and
//                    int32[n].
//   Matrix record:   int64 rows, int64 cols, int64 nnz, int32 ptr[rows + 1],
//                    int32 col[nnz], float64 val[nnz].
// Every record is validated on load (sizes against the bytes left in the
// file before anything is allocated, then the CSR structure itself).
#include "setup.h"

#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

namespace gsx {
namespace {

constexpr char kCsrMagic[8] = {'G', 'S', 'C', 'S', 'R', '0', '0', '1'};
constexpr char kHierMagic[8] = {'G', 'S', 'H', 'I', 'E', 'R', '0', '1'};

struct File {
  std::FILE *f = nullptr;
  std::string path;
  int64_t left = 0;  // bytes not yet read (reading only)
  File(const std::string &p, const char *mode) : path(p) {
    f = std::fopen(p.c_str(), mode);
    if (!f) throw std::runtime_error("cannot open " + p);
    if (mode[0] == 'r') {
      if (std::fseek(f, 0, SEEK_END) != 0) throw std::runtime_error("cannot seek " + p);
      left = static_cast<int64_t>(std::ftell(f));
      std::rewind(f);
    }
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void put(const void *p, size_t bytes) {
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("write failed: " + path);
  }
  void get(void *p, size_t bytes) {
    if (static_cast<int64_t>(bytes) > left || (bytes && std::fread(p, 1, bytes, f) != bytes))
      throw std::runtime_error("truncated file: " + path);
    left -= static_cast<int64_t>(bytes);
  }
  int64_t i64() {
    int64_t v;
    get(&v, sizeof v);
    return v;
  }
  void close_write() {
    const int rc = std::fclose(f);
    f = nullptr;
    if (rc != 0) throw std::runtime_error("write failed: " + path);
  }
};

void put_i64(File &w, int64_t v) { w.put(&v, sizeof v); }

void put_csr(File &w, const Csr &A) {
  put_i64(w, A.rows);
  put_i64(w, A.cols);
  put_i64(w, A.nnz());
  w.put(A.ptr.data(), A.ptr.size() * sizeof(int32_t));
  w.put(A.col.data(), A.col.size() * sizeof(int32_t));
  w.put(A.val.data(), A.val.size() * sizeof(double));
}

Csr get_csr(File &r) {
  Csr A;
  A.rows = r.i64();
  A.cols = r.i64();
  const int64_t nnz = r.i64();
  if (A.rows <= 0 || A.cols <= 0 || nnz < 0 || A.rows > INT32_MAX || A.cols > INT32_MAX ||
      nnz > INT32_MAX)
    throw std::runtime_error("corrupt matrix header in " + r.path);
  const int64_t need = (A.rows + 1) * 4 + nnz * 12;
  if (need > r.left) throw std::runtime_error("truncated file: " + r.path);
  A.ptr.resize(A.rows + 1);
  A.col.resize(nnz);
  A.val.resize(nnz);
  r.get(A.ptr.data(), A.ptr.size() * sizeof(int32_t));
  r.get(A.col.data(), A.col.size() * sizeof(int32_t));
  r.get(A.val.data(), A.val.size() * sizeof(double));
  A.check();
  return A;
}

void expect_magic(File &r, const char (&magic)[8], const char *what) {
  char m[8];
  r.get(m, sizeof m);
  if (std::memcmp(m, magic, sizeof m) != 0)
    throw std::runtime_error(r.path + " is not a " + what + " file");
}

void expect_end(File &r) {
  if (r.left != 0) throw std::runtime_error("trailing bytes in " + r.path);
}

}  // namespace

void save_csr(const Csr &A, const std::string &path) {
  File w(path, "wb");
  w.put(kCsrMagic, sizeof kCsrMagic);
  put_csr(w, A);
  w.close_write();
}

Csr load_csr(const std::string &path) {
  File r(path, "rb");
  expect_magic(r, kCsrMagic, "CSR");
  Csr A = get_csr(r);
  expect_end(r);
  return A;
}

void save_hierarchy(const Hierarchy &H, const std::string &path) {
  if (H.P.size() != H.R.size())
    throw std::runtime_error("hierarchy has different numbers of P and R matrices.");
  File w(path, "wb");
  w.put(kHierMagic, sizeof kHierMagic);
  put_i64(w, static_cast<int64_t>(H.A.size()));
  put_i64(w, static_cast<int64_t>(H.P.size()));
  for (const Csr &A : H.A) put_csr(w, A);
  for (const Csr &P : H.P) put_csr(w, P);
  for (const Csr &R : H.R) put_csr(w, R);
  put_i64(w, static_cast<int64_t>(H.aggregates.size()));
  for (const auto &a : H.aggregates) {
    put_i64(w, static_cast<int64_t>(a.size()));
    w.put(a.data(), a.size() * sizeof(int32_t));
  }
  w.close_write();
}

Hierarchy load_hierarchy(const std::string &path) {
  File r(path, "rb");
  expect_magic(r, kHierMagic, "hierarchy");
  const int64_t L = r.i64(), T = r.i64();
  // Each record takes at least 28 bytes, which bounds the counts.
  if (L < 0 || T < 0 || L + 2 * T > r.left / 28)
    throw std::runtime_error("corrupt hierarchy header in " + path);
  Hierarchy H;
  for (int64_t k = 0; k < L; ++k) H.A.push_back(get_csr(r));
  for (int64_t k = 0; k < T; ++k) H.P.push_back(get_csr(r));
  for (int64_t k = 0; k < T; ++k) H.R.push_back(get_csr(r));
  const int64_t na = r.i64();
  if (na < 0 || na > r.left / 8) throw std::runtime_error("corrupt hierarchy header in " + path);
  for (int64_t k = 0; k < na; ++k) {
    const int64_t n = r.i64();
    if (n < 0 || n * 4 > r.left) throw std::runtime_error("truncated file: " + path);
    std::vector<int32_t> a(n);
    r.get(a.data(), a.size() * sizeof(int32_t));
    H.aggregates.push_back(std::move(a));
  }
  expect_end(r);
  return H;
}

}  // namespace gsx
