// CSR containers, permutations, transposes, SpGEMM, ELL and BSR3-ELL
// packing, diagonal (block) inverses and the dense SPD inverse used by the
// coarsest-level direct solve.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace gsx {

void Csr::check() const {
  if (rows <= 0 || cols <= 0)
    throw std::runtime_error("HostCSRMatrix: matrix dimensions must be nonzero.");
  if (static_cast<int64_t>(ptr.size()) != rows + 1)
    throw std::runtime_error("HostCSRMatrix: row pointer has the wrong length.");
  if (ptr[0] != 0 || static_cast<int64_t>(ptr[rows]) != nnz() ||
      col.size() != val.size())
    throw std::runtime_error("HostCSRMatrix: inconsistent nonzero count.");
  for (int64_t r = 0; r < rows; ++r)
    if (ptr[r + 1] < ptr[r])
      throw std::runtime_error("HostCSRMatrix: invalid CSR row pointer.");
  bool ok = true;
  const int64_t nz = nnz();
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int64_t k = 0; k < nz; ++k) ok = ok && col[k] >= 0 && col[k] < cols;
  if (!ok) throw std::runtime_error("HostCSRMatrix: column index out of bounds.");
}

Perm Perm::identity(int64_t n) {
  Perm p;
  p.old_to_new.resize(n);
  std::iota(p.old_to_new.begin(), p.old_to_new.end(), 0);
  p.new_to_old = p.old_to_new;
  return p;
}

Perm Perm::from_new_to_old(const std::vector<int32_t> &n2o) {
  Perm p;
  p.new_to_old = n2o;
  p.old_to_new.assign(n2o.size(), -1);
  for (size_t i = 0; i < n2o.size(); ++i) {
    if (n2o[i] < 0 || static_cast<size_t>(n2o[i]) >= n2o.size())
      throw std::runtime_error("Permutation: newToOld contains invalid index.");
    p.old_to_new[n2o[i]] = static_cast<int32_t>(i);
  }
  p.check();
  return p;
}

void Perm::check() const {
  const size_t n = old_to_new.size();
  if (n == 0) throw std::runtime_error("Permutation: size must be nonzero.");
  if (new_to_old.size() != n)
    throw std::runtime_error("Permutation: mappings have different sizes.");
  for (size_t o = 0; o < n; ++o) {
    const int32_t nw = old_to_new[o];
    if (nw < 0 || static_cast<size_t>(nw) >= n)
      throw std::runtime_error("Permutation: oldToNew contains invalid index.");
    if (new_to_old[nw] != static_cast<int32_t>(o))
      throw std::runtime_error("Permutation: mappings are not inverses.");
  }
}

Csr permute_keep_order(const Csr &A, const Perm &p) {
  if (A.rows != A.cols)
    throw std::runtime_error(
        "HostCSRMatrix::applyPermutation: matrix must be square.");
  if (static_cast<int64_t>(p.old_to_new.size()) != A.rows)
    throw std::runtime_error(
        "HostCSRMatrix::applyPermutation: permutation size does not match "
        "matrix dimensions.");
  Csr B;
  B.rows = A.rows;
  B.cols = A.cols;
  B.ptr.assign(A.rows + 1, 0);
  for (int64_t r = 0; r < A.rows; ++r) {
    const int32_t o = p.new_to_old[r];
    B.ptr[r + 1] = B.ptr[r] + (A.ptr[o + 1] - A.ptr[o]);
  }
  B.col.resize(A.nnz());
  B.val.resize(A.nnz());
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A.rows; ++r) {
    const int32_t o = p.new_to_old[r];
    int64_t d = B.ptr[r];
    for (int32_t k = A.ptr[o]; k < A.ptr[o + 1]; ++k, ++d) {
      B.col[d] = p.old_to_new[A.col[k]];
      B.val[d] = A.val[k];
    }
  }
  return B;
}

// Sort each row by column and sum duplicates in their stored order.
void sort_rows(Csr &A) {
  std::vector<int32_t> newptr(A.rows + 1, 0);
  std::vector<int32_t> counts(A.rows, 0);
#pragma omp parallel
  {
    // Sort key: column (sign bit flipped, so int32 order is unsigned order)
    // above the entry's position in the row; keys are unique, so the order
    // is that of a stable sort by column.
    std::vector<uint64_t> key;
    std::vector<double> v;
#pragma omp for schedule(dynamic, 256)
    for (int64_t r = 0; r < A.rows; ++r) {
      const int64_t b = A.ptr[r], e = A.ptr[r + 1];
      key.resize(e - b);
      for (int64_t k = b; k < e; ++k)
        key[k - b] = (static_cast<uint64_t>(static_cast<uint32_t>(A.col[k]) ^ 0x80000000u) << 32) |
                     static_cast<uint64_t>(k - b);
      std::sort(key.begin(), key.end());
      // Collapse duplicates (summed in stored order), buffering values.
      v.clear();
      int64_t n = 0;
      int32_t last = 0;
      for (size_t i = 0; i < key.size(); ++i) {
        const int32_t c = static_cast<int32_t>(static_cast<uint32_t>(key[i] >> 32) ^ 0x80000000u);
        const double x = A.val[b + static_cast<int64_t>(key[i] & 0xffffffffu)];
        if (n > 0 && last == c) {
          v.back() += x;
        } else {
          key[n++] = static_cast<uint64_t>(static_cast<uint32_t>(c));
          v.push_back(x);
          last = c;
        }
      }
      for (int64_t i = 0; i < n; ++i) {
        A.col[b + i] = static_cast<int32_t>(static_cast<uint32_t>(key[i]));
        A.val[b + i] = v[i];
      }
      counts[r] = static_cast<int32_t>(n);
    }
  }
  for (int64_t r = 0; r < A.rows; ++r) newptr[r + 1] = newptr[r] + counts[r];
  if (newptr[A.rows] != A.ptr[A.rows]) {
    RawVec<int32_t> col(newptr[A.rows]);
    RawVec<double> val(newptr[A.rows]);
    for (int64_t r = 0; r < A.rows; ++r) {
      std::memcpy(&col[newptr[r]], &A.col[A.ptr[r]], counts[r] * sizeof(int32_t));
      std::memcpy(&val[newptr[r]], &A.val[A.ptr[r]], counts[r] * sizeof(double));
    }
    A.col.swap(col);
    A.val.swap(val);
  }
  A.ptr.swap(newptr);
}

bool rows_sorted_unique(const Csr &A) {
  bool ok = true;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int64_t r = 0; r < A.rows; ++r)
    for (int32_t k = A.ptr[r] + 1; k < A.ptr[r + 1]; ++k) ok = ok && A.col[k - 1] < A.col[k];
  return ok;
}

Csr permute_sorted(const Csr &A, const int32_t *row_o2n,
                   const int32_t *col_o2n) {
  Csr B;
  B.rows = A.rows;
  B.cols = A.cols;
  B.ptr.assign(A.rows + 1, 0);
  std::vector<int32_t> src_row(A.rows);
  for (int64_t r = 0; r < A.rows; ++r) src_row[row_o2n[r]] = static_cast<int32_t>(r);
  for (int64_t r = 0; r < A.rows; ++r) {
    const int32_t o = src_row[r];
    B.ptr[r + 1] = B.ptr[r] + (A.ptr[o + 1] - A.ptr[o]);
  }
  B.col.resize(A.nnz());
  B.val.resize(A.nnz());
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A.rows; ++r) {
    const int32_t o = src_row[r];
    int64_t d = B.ptr[r];
    for (int32_t k = A.ptr[o]; k < A.ptr[o + 1]; ++k, ++d) {
      B.col[d] = col_o2n[A.col[k]];
      B.val[d] = A.val[k];
    }
  }
  sort_rows(B);
  return B;
}

Csr relabel(const Csr &A, const int32_t *rq, const int32_t *cmap) {
  Csr B;
  B.rows = A.rows;
  B.cols = A.cols;
  B.ptr.assign(A.rows + 1, 0);
  for (int64_t i = 0; i < A.rows; ++i) B.ptr[i + 1] = B.ptr[i] + (A.ptr[rq[i] + 1] - A.ptr[rq[i]]);
  B.col.resize(A.nnz());
  B.val.resize(A.nnz());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < A.rows; ++i) {
    int64_t d = B.ptr[i];
    for (int32_t k = A.ptr[rq[i]]; k < A.ptr[rq[i] + 1]; ++k, ++d) {
      B.col[d] = cmap[A.col[k]];
      B.val[d] = A.val[k];
    }
  }
  return B;
}

Csr take_rows(const Csr &A, const int32_t *rows, int64_t nrows) {
  Csr B;
  B.rows = nrows;
  B.cols = A.cols;
  B.ptr.assign(nrows + 1, 0);
  for (int64_t i = 0; i < nrows; ++i) {
    if (rows[i] < 0 || rows[i] >= A.rows) throw std::runtime_error("take_rows: row out of range.");
    B.ptr[i + 1] = B.ptr[i] + (A.ptr[rows[i] + 1] - A.ptr[rows[i]]);
  }
  B.col.resize(B.ptr[nrows]);
  B.val.resize(B.ptr[nrows]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nrows; ++i) {
    const int32_t r = rows[i];
    std::copy(A.col.begin() + A.ptr[r], A.col.begin() + A.ptr[r + 1], B.col.begin() + B.ptr[i]);
    std::copy(A.val.begin() + A.ptr[r], A.val.begin() + A.ptr[r + 1], B.val.begin() + B.ptr[i]);
  }
  return B;
}

void check_columns(const Csr &A, const char *what) {
  bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
  for (int64_t k = 0; k < static_cast<int64_t>(A.col.size()); ++k)
    bad = bad || A.col[k] < 0 || A.col[k] >= A.cols;
  if (bad) throw std::runtime_error(what);
}

Csr transpose(const Csr &A) {
  Csr T;
  T.rows = A.cols;
  T.cols = A.rows;
  T.ptr.assign(A.cols + 1, 0);
  for (int32_t c : A.col) T.ptr[c + 1]++;
  for (int64_t i = 0; i < A.cols; ++i) T.ptr[i + 1] += T.ptr[i];
  T.col.resize(A.nnz());
  T.val.resize(A.nnz());
  std::vector<int32_t> next(T.ptr.begin(), T.ptr.end() - 1);
  for (int64_t r = 0; r < A.rows; ++r)
    for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
      const int32_t d = next[A.col[k]]++;
      T.col[d] = static_cast<int32_t>(r);
      T.val[d] = A.val[k];
    }
  return T;  // rows of T are visited in ascending r: already sorted
}

Csr spgemm(const Csr &A, const Csr &B) {
  if (A.cols != B.rows) throw std::runtime_error("spgemm: shape mismatch.");
  Csr C;
  C.rows = A.rows;
  C.cols = B.cols;
  C.ptr.assign(A.rows + 1, 0);
  std::vector<int32_t> rowlen(A.rows, 0);
#pragma omp parallel
  {
    std::vector<int32_t> mark(B.cols, -1);
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < A.rows; ++r) {
      int32_t cnt = 0;
      for (int32_t ka = A.ptr[r]; ka < A.ptr[r + 1]; ++ka) {
        const int32_t j = A.col[ka];
        for (int32_t kb = B.ptr[j]; kb < B.ptr[j + 1]; ++kb) {
          const int32_t c = B.col[kb];
          if (mark[c] != static_cast<int32_t>(r)) {
            mark[c] = static_cast<int32_t>(r);
            ++cnt;
          }
        }
      }
      rowlen[r] = cnt;
    }
  }
  for (int64_t r = 0; r < A.rows; ++r) {
    const int64_t next = static_cast<int64_t>(C.ptr[r]) + rowlen[r];
    if (next > INT32_MAX) throw std::runtime_error("spgemm: nnz exceeds int32.");
    C.ptr[r + 1] = static_cast<int32_t>(next);
  }
  C.col.resize(C.ptr[A.rows]);
  C.val.resize(C.ptr[A.rows]);
#pragma omp parallel
  {
    std::vector<int32_t> pos(B.cols, -1);
    std::vector<int32_t> cols;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < A.rows; ++r) {
      cols.clear();
      const int32_t base = C.ptr[r];
      for (int32_t ka = A.ptr[r]; ka < A.ptr[r + 1]; ++ka) {
        const int32_t j = A.col[ka];
        const double a = A.val[ka];
        for (int32_t kb = B.ptr[j]; kb < B.ptr[j + 1]; ++kb) {
          const int32_t c = B.col[kb];
          if (pos[c] < 0) {
            pos[c] = static_cast<int32_t>(cols.size());
            cols.push_back(c);
            C.val[base + pos[c]] = a * B.val[kb];
          } else {
            C.val[base + pos[c]] += a * B.val[kb];
          }
        }
      }
      // Sort the row by column, carrying values.
      std::vector<std::pair<int32_t, double>> tmp(cols.size());
      for (size_t i = 0; i < cols.size(); ++i)
        tmp[i] = {cols[i], C.val[base + i]};
      std::sort(tmp.begin(), tmp.end(),
                [](const std::pair<int32_t, double> &x,
                   const std::pair<int32_t, double> &y) { return x.first < y.first; });
      for (size_t i = 0; i < tmp.size(); ++i) {
        C.col[base + i] = tmp[i].first;
        C.val[base + i] = tmp[i].second;
        pos[tmp[i].first] = -1;
      }
    }
  }
  return C;
}

int32_t max_row_length(const Csr &A) {
  int32_t w = 0;
  for (int64_t r = 0; r < A.rows; ++r) w = std::max(w, A.ptr[r + 1] - A.ptr[r]);
  return w;
}

Ell ell_pack(const Csr &A, int32_t width) {
  if (A.rows == 0 || A.cols == 0)
    throw std::runtime_error("HostELLMatrix: matrix dimensions must be nonzero.");
  if (width <= 0) throw std::runtime_error("HostELLMatrix: ELL width must be nonzero.");
  Ell E;
  E.rows = A.rows;
  E.cols = A.cols;
  E.nnz = A.nnz();
  E.width = width;
  for (int64_t r = 0; r < A.rows; ++r) {
    const int32_t len = A.ptr[r + 1] - A.ptr[r];
    if (len < 0) throw std::runtime_error("HostELLMatrix: invalid CSR row pointer.");
    if (len > width) E.overflow_nnz += len - width;
  }
  const int64_t slots = static_cast<int64_t>(width) * A.rows;
  E.col.assign(slots, 0);
  E.val.assign(slots, 0.0);
  E.ovf_ptr.assign(A.rows + 1, 0);
  E.ovf_col.resize(E.overflow_nnz);
  E.ovf_val.resize(E.overflow_nnz);
  int64_t op = 0;
  for (int64_t r = 0; r < A.rows; ++r) {
    const int32_t b = A.ptr[r], e = A.ptr[r + 1];
    const int32_t inell = std::min(width, e - b);
    for (int32_t s = 0; s < inell; ++s) {
      E.col[static_cast<int64_t>(s) * A.rows + r] = A.col[b + s];
      E.val[static_cast<int64_t>(s) * A.rows + r] = A.val[b + s];
    }
    for (int32_t k = b + inell; k < e; ++k, ++op) {
      E.ovf_col[op] = A.col[k];
      E.ovf_val[op] = A.val[k];
    }
    E.ovf_ptr[r + 1] = static_cast<int32_t>(op);
  }
  if (op != E.overflow_nnz)
    throw std::runtime_error("HostELLMatrix: inconsistent overflow size.");
  return E;
}

bool bsr3_pack(const Csr &A, Bsr3Csr &out, const int32_t *key) {
  if (A.rows % 3 != 0 || A.cols % 3 != 0) return false;
  const int64_t nb = A.rows / 3, ncb = A.cols / 3;
  auto rank = [key](int32_t cb) { return key ? key[cb] : cb; };
  // Block columns of block row b: union over its three scalar rows, ordered
  // by rank, so that a row whose entries come in rank order is summed in
  // that same order by the block kernels.
  auto gather = [&](int64_t b, std::vector<int32_t> &bc) {
    bc.clear();
    for (int64_t r = 3 * b; r < 3 * b + 3; ++r)
      for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) bc.push_back(A.col[k] / 3);
    std::sort(bc.begin(), bc.end(), [&](int32_t u, int32_t v) {
      return rank(u) != rank(v) ? rank(u) < rank(v) : u < v;
    });
    bc.erase(std::unique(bc.begin(), bc.end()), bc.end());
  };
  check_columns(A, "BSR upload: block column out of range.");
  std::vector<int64_t> ptr(nb + 1, 0);
  int32_t width = 0;
#pragma omp parallel reduction(max : width)
  {
    std::vector<int32_t> bc;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t b = 0; b < nb; ++b) {
      gather(b, bc);
      ptr[b + 1] = static_cast<int64_t>(bc.size());
      width = std::max(width, static_cast<int32_t>(bc.size()));
    }
  }
  for (int64_t b = 0; b < nb; ++b) ptr[b + 1] += ptr[b];
  const int64_t nblocks = ptr[nb];
  std::vector<int32_t> bcol(nblocks);
  RawVec<double> val(9 * nblocks);
#pragma omp parallel
  {
    std::vector<int32_t> bc;
    // slot[c]: slot of block column c in the current block row (only the
    // row's own columns are written before they are read).
    std::vector<int32_t> slot(ncb);
#pragma omp for schedule(dynamic, 1024)
    for (int64_t b = 0; b < nb; ++b) {
      gather(b, bc);
      const int64_t k0 = ptr[b];
      for (size_t s = 0; s < bc.size(); ++s) {
        bcol[k0 + s] = bc[s];
        slot[bc[s]] = static_cast<int32_t>(s);
      }
      double *v = val.data() + 9 * k0;
      std::fill(v, v + 9 * bc.size(), 0.0);
      for (int l = 0; l < 3; ++l) {
        const int64_t r = 3 * b + l;
        for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k)
          v[9 * slot[A.col[k] / 3] + 3 * l + A.col[k] % 3] += A.val[k];
      }
    }
  }
  out.nb = nb;
  out.ncb = ncb;
  out.nblocks = nblocks;
  out.width = width;
  out.ptr = std::move(ptr);
  out.bcol = std::move(bcol);
  out.val = std::move(val);
  return true;
}

Bsr3Csr bsr3_take_rows(const Bsr3Csr &B, const int32_t *rows, int64_t nrows) {
  Bsr3Csr out;
  out.nb = nrows;
  out.ncb = B.ncb;
  out.ptr.assign(nrows + 1, 0);
  int64_t width = 0;
  for (int64_t i = 0; i < nrows; ++i) {
    if (rows[i] < 0 || rows[i] >= B.nb) throw std::runtime_error("bsr3_take_rows: row out of range.");
    const int64_t len = B.ptr[rows[i] + 1] - B.ptr[rows[i]];
    out.ptr[i + 1] = out.ptr[i] + len;
    width = std::max(width, len);
  }
  out.width = static_cast<int32_t>(width);
  out.nblocks = out.ptr[nrows];
  out.bcol.resize(out.nblocks);
  out.val.resize(9 * out.nblocks);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nrows; ++i) {
    const int64_t s0 = B.ptr[rows[i]], s1 = B.ptr[rows[i] + 1], d = out.ptr[i];
    std::copy(B.bcol.begin() + s0, B.bcol.begin() + s1, out.bcol.begin() + d);
    std::copy(B.val.begin() + 9 * s0, B.val.begin() + 9 * s1, out.val.begin() + 9 * d);
  }
  return out;
}

std::vector<int32_t> bsr3_sorted_order(const Bsr3Csr &B, int64_t sigma) {
  if (sigma < 1) throw std::runtime_error("SELL-C-sigma: window must be positive.");
  std::vector<int32_t> order(B.nb);
  for (int64_t i = 0; i < B.nb; ++i) order[i] = static_cast<int32_t>(i);
  for (int64_t w0 = 0; w0 < B.nb; w0 += sigma) {
    const int64_t w1 = std::min(B.nb, w0 + sigma);
    std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int32_t a, int32_t b) {
      return B.ptr[a + 1] - B.ptr[a] > B.ptr[b + 1] - B.ptr[b];
    });
  }
  return order;
}

void touch_pages(void *p, size_t bytes) {
  char *c = static_cast<char *>(p);
  const int64_t pages = static_cast<int64_t>((bytes + 4095) / 4096);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < pages; ++i) c[i * 4096] = 0;
}

void to_float(const double *src, size_t n, float *dst) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) dst[i] = static_cast<float>(src[i]);
}

template <class D>
static void gather_cast_t(const double *src, const int32_t *idx, int64_t n, D *dst) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) dst[i] = static_cast<D>(src[idx[i]]);
}
template <class S>
static void scatter_cast_t(const S *src, const int32_t *idx, int64_t n, double *dst) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) dst[idx[i]] = static_cast<double>(src[i]);
}
void gather_cast(const double *src, const int32_t *idx, int64_t n, double *dst) {
  gather_cast_t(src, idx, n, dst);
}
void gather_cast(const double *src, const int32_t *idx, int64_t n, float *dst) {
  gather_cast_t(src, idx, n, dst);
}
void scatter_cast(const double *src, const int32_t *idx, int64_t n, double *dst) {
  scatter_cast_t(src, idx, n, dst);
}
void scatter_cast(const float *src, const int32_t *idx, int64_t n, double *dst) {
  scatter_cast_t(src, idx, n, dst);
}

double sum_squares(const double *v, int64_t n) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) s += v[i] * v[i];
  return s;
}

Sell64 sell64_pack(const Csr &A) {
  constexpr int64_t S = 64;
  Sell64 out;
  out.rows = A.rows;
  out.cols = A.cols;
  out.nnz = A.nnz();
  const int64_t ns = (A.rows + S - 1) / S;
  out.slice_ptr.assign(ns + 1, 0);
  out.slice_w.assign(ns, 0);
  for (int64_t s = 0; s < ns; ++s) {
    int32_t w = 0;
    for (int64_t r = s * S; r < std::min(A.rows, (s + 1) * S); ++r)
      w = std::max(w, A.ptr[r + 1] - A.ptr[r]);
    out.slice_w[s] = w;
    out.slice_ptr[s + 1] = out.slice_ptr[s] + static_cast<int64_t>(w) * S;
  }
  out.slots = out.slice_ptr[ns];
  out.col.assign(out.slots, 0);
  out.val.assign(out.slots, 0.0);
  const bool square = A.rows == A.cols;
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < ns; ++s)
    for (int64_t l = 0; l < S; ++l) {
      const int64_t r = s * S + l;
      const int64_t base = out.slice_ptr[s] + l;
      int32_t len = 0;
      int32_t pad = 0;
      if (r < A.rows) {
        len = A.ptr[r + 1] - A.ptr[r];
        for (int32_t j = 0; j < len; ++j) {
          out.col[base + j * S] = A.col[A.ptr[r] + j];
          out.val[base + j * S] = A.val[A.ptr[r] + j];
        }
        // A column of the row other than the diagonal, else a valid
        // neighbouring column.
        pad = -1;
        for (int32_t j = len - 1; j >= 0 && pad < 0; --j)
          if (A.col[A.ptr[r] + j] != r) pad = A.col[A.ptr[r] + j];
        if (pad < 0) {
          if (A.cols == 1) pad = 0;
          else if (r + 1 < A.cols) pad = static_cast<int32_t>(r + 1);
          else if (r < A.cols) pad = static_cast<int32_t>(r - 1);
          else pad = 0;
        }
      }
      for (int32_t j = len; j < out.slice_w[s]; ++j) out.col[base + j * S] = pad;
    }
  // Validation of what the kernels assume.
  for (int64_t s = 0; s < ns; ++s)
    for (int64_t l = 0; l < S; ++l) {
      const int64_t r = s * S + l;
      if (r >= A.rows) break;
      const int32_t len = A.ptr[r + 1] - A.ptr[r];
      for (int32_t j = 0; j < out.slice_w[s]; ++j) {
        const int32_t c = out.col[out.slice_ptr[s] + l + j * S];
        if (c < 0 || c >= A.cols)
          throw std::runtime_error("SELL-64 packing: column index out of range.");
        if (square && j >= len && c == r && A.cols > 1)
          throw std::runtime_error("SELL-64 packing: padding on the diagonal.");
      }
    }
  return out;
}

void bsr3_pair_planes(Bsr3Sell64 &P) {
  const int64_t groups = P.slots / 64;
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < groups; ++g) {
    double t[576];
    double *v = P.val.data() + g * 576;
    for (int e = 0; e < 9; ++e)
      for (int l = 0; l < 64; ++l)
        t[e < 8 ? 128 * (e >> 1) + 2 * l + (e & 1) : 512 + l] = v[64 * e + l];
    for (int i = 0; i < 576; ++i) v[i] = t[i];
  }
}

Bsr3Sell64 bsr3_sell64_pack(const Bsr3Csr &B) {
  constexpr int64_t S = 64;
  Bsr3Sell64 out;
  out.nb = B.nb;
  out.nblocks = B.nblocks;
  const int64_t ns = (B.nb + S - 1) / S;
  out.slice_ptr.assign(ns + 1, 0);
  out.slice_w.assign(ns, 0);
  for (int64_t s = 0; s < ns; ++s) {
    int64_t w = 0;
    for (int64_t b = s * S; b < std::min(B.nb, (s + 1) * S); ++b)
      w = std::max(w, B.ptr[b + 1] - B.ptr[b]);
    out.slice_w[s] = static_cast<int32_t>(w);
    out.slice_ptr[s + 1] = out.slice_ptr[s] + w * S;
  }
  out.slots = out.slice_ptr[ns];
  out.bcol.resize(out.slots);
  out.val.resize(out.slots * 9);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t s = 0; s < ns; ++s)
    for (int32_t j = 0; j < out.slice_w[s]; ++j) {
      const int64_t slot = out.slice_ptr[s] + static_cast<int64_t>(j) * S;
      for (int64_t l = 0; l < S; ++l) {
        const int64_t b = s * S + l;
        double *v = out.val.data() + slot * 9 + l;
        if (b < B.nb && j < B.ptr[b + 1] - B.ptr[b]) {
          const int64_t k = B.ptr[b] + j;
          out.bcol[slot + l] = B.bcol[k];
          for (int e = 0; e < 9; ++e) v[e * S] = B.val[9 * k + e];
        } else {
          // Padding: a zero block on a valid column -- the row itself for
          // square operators (the Gauss-Seidel kernels skip that column),
          // else the row's last block column (0 for an empty row).
          int32_t pc = 0;
          if (B.nb == B.ncb) pc = static_cast<int32_t>(b < B.nb ? b : 0);
          else if (b < B.nb && B.ptr[b + 1] > B.ptr[b]) pc = B.bcol[B.ptr[b + 1] - 1];
          out.bcol[slot + l] = pc;
          for (int e = 0; e < 9; ++e) v[e * S] = 0.0;
        }
      }
    }
  return out;
}

Sell64 sell64_pack_sorted(const Csr &A, int64_t sigma) {
  if (A.rows == A.cols)
    throw std::runtime_error("SELL-C-sigma packing is for non-square operators.");
  if (sigma < 1) throw std::runtime_error("SELL-C-sigma: window must be positive.");
  std::vector<int32_t> order(A.rows);
  for (int64_t i = 0; i < A.rows; ++i) order[i] = static_cast<int32_t>(i);
  for (int64_t w0 = 0; w0 < A.rows; w0 += sigma) {
    const int64_t w1 = std::min(A.rows, w0 + sigma);
    std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int32_t a, int32_t b) {
      return A.ptr[a + 1] - A.ptr[a] > A.ptr[b + 1] - A.ptr[b];
    });
  }
  const Csr B = take_rows(A, order.data(), A.rows);
  Sell64 out = sell64_pack(B);
  out.row_map = std::move(order);
  return out;
}

static void inverse3x3_spd(const double a[9], double inv[9], bool &ok) {
  // Cholesky a = L L^T, then inv = L^-T L^-1.
  double l00 = a[0];
  if (!(l00 > 0.0)) { ok = false; return; }
  l00 = std::sqrt(l00);
  const double l10 = a[3] / l00, l20 = a[6] / l00;
  double d1 = a[4] - l10 * l10;
  if (!(d1 > 0.0)) { ok = false; return; }
  const double l11 = std::sqrt(d1);
  const double l21 = (a[7] - l20 * l10) / l11;
  double d2 = a[8] - l20 * l20 - l21 * l21;
  if (!(d2 > 0.0)) { ok = false; return; }
  const double l22 = std::sqrt(d2);
  // M = L^-1 (lower triangular)
  const double m00 = 1.0 / l00, m11 = 1.0 / l11, m22 = 1.0 / l22;
  const double m10 = -l10 * m00 / l11;
  const double m21 = -l21 * m11 / l22;
  const double m20 = -(l20 * m00 + l21 * m10) / l22;
  const double M[3][3] = {{m00, 0, 0}, {m10, m11, 0}, {m20, m21, m22}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += M[k][i] * M[k][j];
      inv[3 * i + j] = s;
    }
  ok = true;
}

std::vector<double> inverse_diag_blocks3(const Csr &A) {
  if (A.rows % 3 != 0)
    throw std::runtime_error(
        "3x3 block smoother requires number of DOFs divisible by 3.");
  const int64_t nb = A.rows / 3;
  std::vector<double> inv(9 * nb);
  bool all_ok = true;
#pragma omp parallel for schedule(static) reduction(&& : all_ok)
  for (int64_t b = 0; b < nb; ++b) {
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int l = 0; l < 3; ++l) {
      const int64_t r = 3 * b + l;
      for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
        const int64_t c = A.col[k];
        if (c >= 3 * b && c < 3 * b + 3) blk[3 * l + (c - 3 * b)] = A.val[k];
      }
    }
    bool ok = false;
    inverse3x3_spd(blk, &inv[9 * b], ok);
    all_ok = all_ok && ok;
  }
  if (!all_ok) throw std::runtime_error("Block diagonal is not SPD.");
  return inv;
}

std::vector<double> inverse_diagonal(const Csr &A) {
  std::vector<double> d(A.rows);
  for (int64_t r = 0; r < A.rows; ++r) {
    bool found = false;
    for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k)
      if (A.col[k] == r) {
        if (std::abs(A.val[k]) < 1e-12)
          throw std::runtime_error(
              "DampedJacobiPreconditioner::initialize: zero diagonal entry.");
        d[r] = 1.0 / A.val[k];
        found = true;
        break;
      }
    if (!found)
      throw std::runtime_error(
          "DampedJacobiPreconditioner::initialize: missing diagonal entry.");
  }
  return d;
}

void spd_inverse(int64_t n, std::vector<double> &a) {
  if (static_cast<int64_t>(a.size()) != n * n)
    throw std::runtime_error("spd_inverse: size mismatch.");
  // Right-looking Cholesky, lower triangle, a = L L^T.
  for (int64_t j = 0; j < n; ++j) {
    double d = a[j * n + j];
    for (int64_t k = 0; k < j; ++k) d -= a[j * n + k] * a[j * n + k];
    if (!(d > 0.0)) throw std::runtime_error("coarse level matrix is not SPD.");
    const double ljj = std::sqrt(d);
    a[j * n + j] = ljj;
#pragma omp parallel for schedule(static) if (n - j > 256)
    for (int64_t i = j + 1; i < n; ++i) {
      double s = a[i * n + j];
      const double *ri = &a[i * n];
      const double *rj = &a[j * n];
      for (int64_t k = 0; k < j; ++k) s -= ri[k] * rj[k];
      a[i * n + j] = s / ljj;
    }
  }
  // M = L^-1, column by column (forward substitution on unit vectors).
  std::vector<double> m(n * n, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t c = 0; c < n; ++c) {
    for (int64_t i = c; i < n; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int64_t k = c; k < i; ++k) s -= a[i * n + k] * m[k * n + c];
      m[i * n + c] = s / a[i * n + i];
    }
  }
  // inv = M^T M.
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j <= i; ++j) {
      double s = 0.0;
      for (int64_t k = i; k < n; ++k) s += m[k * n + i] * m[k * n + j];
      a[i * n + j] = s;
      a[j * n + i] = s;
    }
}

}  // namespace gsx
