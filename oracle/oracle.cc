// CPU oracle for the AMG-PCG solve phase -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain C++, the reference's CPU path
// (src/CPU/CPUBackend.cc), its PCG loop (src/CGSolver.cc), its V-cycle and
// smoothers (src/Smoothers/*.cc), its preconditioner wrappers
// (src/Preconditioners/*.cc) and its setup helpers (src/GraphColoring.cc,
// src/ReorderingStrategies/RCMReordering.cc,
// src/ReorderingStrategies/HierarchicalReordering.cc, src/HostELLMatrix.cc,
// src/HostSparseMatrix.cc, src/SparseMatrixUtils.cc, src/BlockUtils.cc),
// templated on the value type so the same code serves as the FP32
// "as-reference" path and the FP64 oracle. Only the test suite, the smoke
// check and the CPU-baseline leg of the benchmark load it, and only as the
// checker; the product library never links or calls it.
//
// Summation orders follow the reference exactly: row sums walk the CSR row
// in storage order starting from zero, Gauss-Seidel sums the off-diagonal
// products first and then forms (b - sum) / a_ii, block Gauss-Seidel sums
// the off-block products then applies the 3x3 inverse. Dot products are a
// sequential left-to-right sum (the reference's OpenMP reduction order
// depends on the thread count; `parallel_dot` switches to that form for
// timing runs). Build with -ffp-contract=off so no FMA is formed.
//
// Parity status of each function is recorded in the test files that pin
// it; orc_sa_* is absent on purpose: the AMG hierarchy is an input here.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

struct CsrView {
  int64_t rows, cols;
  const int32_t *ptr;
  const int32_t *col;
};

template <class T>
struct CsrT {
  int64_t rows = 0, cols = 0;
  std::vector<int32_t> ptr, col;
  std::vector<T> val;
};

// src/CPU/CPUBackend.cc:62-86 rowDotExcluding.
template <class T>
T row_dot_excluding(const CsrT<T> &A, const T *x, int64_t row, int64_t exb,
                    int64_t exe, T *diag_out) {
  T sum = T(0);
  if (diag_out) *diag_out = T(0);
  for (int32_t k = A.ptr[row]; k < A.ptr[row + 1]; ++k) {
    const int32_t c = A.col[k];
    if (c == row && diag_out) {
      *diag_out = A.val[k];
      continue;
    }
    if (c >= exb && c < exe) continue;
    sum += A.val[k] * x[c];
  }
  return sum;
}

// src/CPU/CPUBackend.cc:217-236.
template <class T>
void spmv(const CsrT<T> &A, const T *x, T *y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A.rows; ++r) y[r] = row_dot_excluding(A, x, r, 0, 0, (T *)nullptr);
}

// src/CPU/CPUBackend.cc:238-268. Returns false on a zero diagonal.
template <class T>
bool gs_color(const CsrT<T> &A, const int32_t *order, int64_t begin, int64_t count,
              const T *rhs, T *x, T omega) {
  bool zero = false;
#pragma omp parallel for schedule(static) reduction(|| : zero)
  for (int64_t k = begin; k < begin + count; ++k) {
    const int32_t row = order[k];
    T diag = T(0);
    const T off = row_dot_excluding(A, x, row, row, row + 1, &diag);
    if (std::abs(diag) < T(1e-12)) {
      zero = true;
      continue;
    }
    const T nv = (rhs[row] - off) / diag;
    x[row] = (T(1) - omega) * x[row] + omega * nv;
  }
  return !zero;
}

// src/CPU/CPUBackend.cc:350-387.
template <class T>
void block_gs_color(const CsrT<T> &A, const int32_t *order, const T *dinv,
                    int64_t begin, int64_t end, const T *rhs, T *x, T omega) {
#pragma omp parallel for schedule(static)
  for (int64_t k = begin; k < end; ++k) {
    const int32_t b = order[k];
    const int64_t r0 = 3 * static_cast<int64_t>(b);
    T lr[3];
    for (int l = 0; l < 3; ++l) {
      const T off = row_dot_excluding(A, x, r0 + l, r0, r0 + 3, (T *)nullptr);
      lr[l] = rhs[r0 + l] - off;
    }
    const T *bi = dinv + 9 * static_cast<int64_t>(b);
    for (int l = 0; l < 3; ++l) {
      T nv = T(0);
      for (int c = 0; c < 3; ++c) nv += bi[3 * l + c] * lr[c];
      x[r0 + l] = (T(1) - omega) * x[r0 + l] + omega * nv;
    }
  }
}

// src/CPU/CPUBackend.cc:323-348.
template <class T>
void apply_block_inverses(const T *inv, const T *res, T *corr, int64_t n) {
  const int64_t nb = n / 3;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    for (int r = 0; r < 3; ++r) {
      T s = T(0);
      for (int c = 0; c < 3; ++c) s += inv[9 * b + 3 * r + c] * res[3 * b + c];
      corr[3 * b + r] = s;
    }
  }
}

bool g_parallel_dot = false;

// src/CPU/CPUBackend.cc:269-284 (sequential order unless g_parallel_dot).
template <class T>
T dot(const T *x, const T *y, int64_t n) {
  T s = T(0);
  if (g_parallel_dot) {
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
  }
  for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

template <class T>
void axpy(T a, const T *x, T *y, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}

// src/GraphColoring.cc:9-72.
struct Coloring {
  int32_t ncolors = 0;
  std::vector<int32_t> colors, vertices, offsets;
};
Coloring color_graph(int64_t n, const int32_t *ptr, const int32_t *col) {
  Coloring c;
  c.colors.assign(n, -1);
  std::vector<int64_t> forbidden(n + 1, -1);
  for (int64_t i = 0; i < n; ++i) {
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int32_t j = col[k];
      if (j != i && c.colors[j] >= 0) forbidden[c.colors[j]] = i;
    }
    int32_t cc = 0;
    while (forbidden[cc] == i) ++cc;
    c.colors[i] = cc;
    c.ncolors = std::max(c.ncolors, cc + 1);
  }
  std::vector<int32_t> counts(c.ncolors, 0);
  for (int64_t v = 0; v < n; ++v) ++counts[c.colors[v]];
  c.offsets.assign(c.ncolors + 1, 0);
  for (int32_t k = 0; k < c.ncolors; ++k) c.offsets[k + 1] = c.offsets[k] + counts[k];
  c.vertices.assign(n, -1);
  std::vector<int32_t> next = c.offsets;
  for (int64_t v = 0; v < n; ++v) c.vertices[next[c.colors[v]]++] = static_cast<int32_t>(v);
  return c;
}

// src/BlockUtils.cc:126-216.
void block_graph(int64_t n, const int32_t *ptr, const int32_t *col, int32_t bs,
                 std::vector<int32_t> &gptr, std::vector<int32_t> &gcol) {
  const int64_t nb = n / bs;
  std::vector<std::vector<int32_t>> adj(nb);
  for (int64_t r = 0; r < n; ++r) {
    const int64_t br = r / bs;
    for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
      const int64_t bc = col[k] / bs;
      if (br == bc) continue;
      adj[br].push_back(static_cast<int32_t>(bc));
      adj[bc].push_back(static_cast<int32_t>(br));
    }
  }
  gptr.assign(nb + 1, 0);
  gcol.clear();
  for (int64_t b = 0; b < nb; ++b) {
    std::sort(adj[b].begin(), adj[b].end());
    adj[b].erase(std::unique(adj[b].begin(), adj[b].end()), adj[b].end());
    gptr[b + 1] = gptr[b] + static_cast<int32_t>(adj[b].size());
  }
  for (auto &a : adj) gcol.insert(gcol.end(), a.begin(), a.end());
}

// src/SparseMatrixUtils.cc:8-31 via Eigen::setFromTriplets: rows relabelled,
// columns relabelled, each row sorted by column, duplicates summed in the
// order the triplets were produced.
template <class T>
CsrT<T> permute_sorted(int64_t rows, int64_t cols, const int32_t *ptr,
                       const int32_t *col, const T *val, const int32_t *ro2n,
                       const int32_t *co2n) {
  std::vector<std::vector<std::pair<int32_t, T>>> R(rows);
  for (int64_t r = 0; r < rows; ++r)
    for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k)
      R[ro2n[r]].emplace_back(co2n[col[k]], val[k]);
  CsrT<T> A;
  A.rows = rows;
  A.cols = cols;
  A.ptr.assign(rows + 1, 0);
  for (int64_t r = 0; r < rows; ++r) {
    auto &v = R[r];
    std::stable_sort(v.begin(), v.end(), [](const std::pair<int32_t, T> &a,
                                            const std::pair<int32_t, T> &b) {
      return a.first < b.first;
    });
    for (size_t i = 0; i < v.size(); ++i) {
      if (!A.col.empty() && static_cast<int64_t>(A.col.size()) > A.ptr[r] &&
          A.col.back() == v[i].first)
        A.val.back() += v[i].second;
      else {
        A.col.push_back(v[i].first);
        A.val.push_back(v[i].second);
      }
    }
    A.ptr[r + 1] = static_cast<int32_t>(A.col.size());
  }
  return A;
}

// ---------------------------------------------------------------------------
// Multigrid (src/Smoothers/MultigridSmoother.cc:22-420) with the per-level
// smoothers of src/Smoothers/*.cc.
// ---------------------------------------------------------------------------
enum SmootherKind { kDampedJacobi = 0, kSGS = 1, kBlockJacobi = 2, kBlockGS = 3 };

template <class T>
struct Level {
  CsrT<T> A;
  int kind = kSGS;
  T omega = T(1);
  int pre = 1, post = 1;
  bool unsmoothed = false;  // a level below the finest given 0 sweeps
  Coloring col;            // SGS: scalar colouring; block GS: block colouring
  std::vector<T> dinv;     // Jacobi: 1/a_ii; block: 9 per block
  std::vector<T> rhs, res, corr, az, tmp;
  std::vector<T> jaz, jcorr;  // the Jacobi smoothers' own work vectors
};

template <class T>
struct MG {
  std::vector<Level<T>> L;
  std::vector<CsrT<T>> P, R;  // P[k]: k+1 -> k
  int coarse_steps = 2;
  bool direct_coarse = false;
  std::vector<double> coarse_inv;  // dense SPD inverse (row-major)
  int vcycles = 1;
  // Reference storage semantics: on every level below the finest, the
  // prolongated correction is written into the level's `correction` vector,
  // which is also the `solution` the level was called with
  // (src/Smoothers/MultigridSmoother.cc:394-404 with :475-476 of the
  // caller), so the level ends with solution = P e + P e. Off: the
  // textbook x += P e with a separate buffer.
  bool reference_aliasing = true;
};

template <class T>
void smooth(Level<T> &lv, const T *rhs, T *x, int iters) {
  const int64_t n = lv.A.rows;
  if (lv.kind == kSGS) {
    // src/Smoothers/SymmetricGaussSeidelSmoother.cc:89-92; a level below
    // the finest whose per-level sweep count is 0 (learned weights, no
    // counterpart in the reference) is left unsmoothed instead.
    if (iters == 0 && lv.unsmoothed) return;
    if (iters == 0)
      throw std::runtime_error("Preconditioner requires at least one smoothing step.");
    for (int it = 0; it < iters; ++it) {
      for (int c = 0; c < lv.col.ncolors; ++c)
        if (!gs_color(lv.A, lv.col.vertices.data(), lv.col.offsets[c],
                      lv.col.offsets[c + 1] - lv.col.offsets[c], rhs, x, lv.omega))
          throw std::runtime_error("CPUBackend::encodeGaussSeidelColor: zero diagonal entry.");
      for (int c = lv.col.ncolors; c > 0; --c)
        if (!gs_color(lv.A, lv.col.vertices.data(), lv.col.offsets[c - 1],
                      lv.col.offsets[c] - lv.col.offsets[c - 1], rhs, x, lv.omega))
          throw std::runtime_error("CPUBackend::encodeGaussSeidelColor: zero diagonal entry.");
    }
  } else if (lv.kind == kBlockGS) {
    // Each colour's full block range [offset_c, offset_c+1), as the Metal
    // backend sweeps it (src/metal/MetalBackend.cc:719-760, which reads the
    // pair the smoother passes as (start, count)). The smoother passes
    // (begin, count) (src/Smoothers/BlockGaussSeidelSmoother.cc:155-166,
    // 196-207) into the CPU backend's (begin, end) (src/CPU/CPUBackend.cc:
    // 364), whose literal loop would sweep colour 0 and a prefix of colour 1
    // and skip the rest; this oracle, like the GPU engine, follows the
    // intended full-colour semantics instead.
    for (int it = 0; it < iters; ++it) {
      for (int c = 0; c < lv.col.ncolors; ++c)
        if (lv.col.offsets[c + 1] > lv.col.offsets[c])
          block_gs_color(lv.A, lv.col.vertices.data(), lv.dinv.data(), lv.col.offsets[c],
                         lv.col.offsets[c + 1], rhs, x, lv.omega);
      for (int c = lv.col.ncolors; c > 0; --c)
        if (lv.col.offsets[c] > lv.col.offsets[c - 1])
          block_gs_color(lv.A, lv.col.vertices.data(), lv.dinv.data(), lv.col.offsets[c - 1],
                         lv.col.offsets[c], rhs, x, lv.omega);
    }
  } else {
    // src/Smoothers/DampedJacobiSmoother.cc:79-87 and
    // src/Smoothers/BlockJacobiSmoother.cc:17-87; the smoothers own their
    // work vectors (Az_, correction_), separate from the level's.
    for (int it = 0; it < iters; ++it) {
      spmv(lv.A, x, lv.jaz.data());
      std::vector<T> &c = lv.jcorr;
      for (int64_t i = 0; i < n; ++i) c[i] = rhs[i] + T(-1) * lv.jaz[i];
      if (lv.kind == kDampedJacobi) {
        for (int64_t i = 0; i < n; ++i) c[i] *= lv.dinv[i];
        axpy(lv.omega, c.data(), x, n);
      } else {
        std::vector<T> out(n);
        apply_block_inverses(lv.dinv.data(), c.data(), out.data(), n);
        axpy(lv.omega, out.data(), x, n);
      }
    }
  }
}

template <class T>
void vcycle(MG<T> &mg, size_t k, const T *rhs, T *x) {
  Level<T> &lv = mg.L[k];
  const int64_t n = lv.A.rows;
  if (k + 1 == mg.L.size()) {
    if (mg.direct_coarse) {
      // FP64 dense solve: x = A_c^-1 rhs (x is zero on entry); each row's
      // sum in column order, rows in parallel.
#pragma omp parallel for schedule(static) if (n > 512)
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) s += mg.coarse_inv[i * n + j] * static_cast<double>(rhs[j]);
        x[i] = static_cast<T>(s);
      }
    } else {
      smooth(lv, rhs, x, mg.coarse_steps);
    }
    return;
  }
  smooth(lv, rhs, x, lv.pre);
  spmv(lv.A, x, lv.az.data());
  for (int64_t i = 0; i < n; ++i) lv.res[i] = rhs[i] + T(-1) * lv.az[i];
  Level<T> &cl = mg.L[k + 1];
  spmv(mg.R[k], lv.res.data(), cl.rhs.data());
  std::fill(cl.corr.begin(), cl.corr.end(), T(0));
  vcycle(mg, k + 1, cl.rhs.data(), cl.corr.data());
  if (mg.reference_aliasing && k > 0) {
    // x is lv.corr here: spmv overwrites it, axpy doubles it.
    spmv(mg.P[k], cl.corr.data(), x);
    axpy(T(1), x, x, n);
  } else {
    spmv(mg.P[k], cl.corr.data(), lv.tmp.data());
    axpy(T(1), lv.tmp.data(), x, n);
  }
  smooth(lv, rhs, x, lv.post);
}

template <class T>
struct Oracle {
  MG<T> mg;
  int precond = 1;  // 0 identity, 1 multigrid
};

// 3x3 SPD inverse in FP64 through its Cholesky factor: a = L L^T,
// M = L^-1, inverse = M^T M (src/BlockUtils.cc:260-266 uses
// Eigen::LLT<Matrix3f>; the FP64 closed form here fixes the operation
// sequence so the device smoother can be compared bit for bit).
bool inverse3x3_cholesky(const double *a, double *inv) {
  const double d0 = a[0];
  if (!(d0 > 0.0)) return false;
  const double l00 = std::sqrt(d0);
  const double l10 = a[3] / l00, l20 = a[6] / l00;
  const double d1 = a[4] - l10 * l10;
  if (!(d1 > 0.0)) return false;
  const double l11 = std::sqrt(d1);
  const double l21 = (a[7] - l20 * l10) / l11;
  const double d2 = a[8] - l20 * l20 - l21 * l21;
  if (!(d2 > 0.0)) return false;
  const double l22 = std::sqrt(d2);
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
  return true;
}

// Dense Cholesky inverse in FP64 (restated independently of the product).
bool dense_spd_inverse(int64_t n, std::vector<double> &a) {
  std::vector<double> L(n * n, 0.0);
  for (int64_t j = 0; j < n; ++j) {
    double d = a[j * n + j];
    for (int64_t k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0)) return false;
    L[j * n + j] = std::sqrt(d);
#pragma omp parallel for schedule(static) if (n - j > 256)
    for (int64_t i = j + 1; i < n; ++i) {
      double s = a[i * n + j];
      for (int64_t k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / L[j * n + j];
    }
  }
  // Solve L L^T X = I column by column (columns are independent).
  std::vector<double> inv(n * n);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t c = 0; c < n; ++c) {
    std::vector<double> y(n);
    for (int64_t i = 0; i < n; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int64_t k = 0; k < i; ++k) s -= L[i * n + k] * y[k];
      y[i] = s / L[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int64_t k = i + 1; k < n; ++k) s -= L[k * n + i] * inv[k * n + c];
      inv[i * n + c] = s / L[i * n + i];
    }
  }
  a.swap(inv);
  return true;
}

}  // namespace

// ===========================================================================
// C entry points (ctypes). Every function returns 0 on success and a
// negative code on error; orc_last_error() returns the message.
// ===========================================================================
extern "C" const char *orc_last_error() { return g_err.c_str(); }

extern "C" void orc_set_parallel_dot(int on) { g_parallel_dot = on != 0; }

extern "C" int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

#define ORC_TRY(...)             \
  try {                          \
    __VA_ARGS__;                 \
    return 0;                    \
  } catch (const std::exception &e) { \
    g_err = e.what();            \
    return -1;                   \
  }

template <class T>
static CsrT<T> make_csr(int64_t rows, int64_t cols, const int32_t *ptr,
                        const int32_t *col, const T *val) {
  CsrT<T> A;
  A.rows = rows;
  A.cols = cols;
  A.ptr.assign(ptr, ptr + rows + 1);
  A.col.assign(col, col + ptr[rows]);
  A.val.assign(val, val + ptr[rows]);
  return A;
}

extern "C" int orc_dense_spd_inverse(int64_t n, double *a) {
  ORC_TRY(std::vector<double> m(a, a + n * n);
          if (!dense_spd_inverse(n, m)) throw std::runtime_error("matrix is not SPD");
          std::copy(m.begin(), m.end(), a));
}

extern "C" int orc_spmv_d(int64_t rows, int64_t cols, const int32_t *ptr, const int32_t *col,
               const double *val, const double *x, double *y) {
  ORC_TRY(spmv(make_csr(rows, cols, ptr, col, val), x, y));
}
extern "C" int orc_spmv_f(int64_t rows, int64_t cols, const int32_t *ptr, const int32_t *col,
               const float *val, const float *x, float *y) {
  ORC_TRY(spmv(make_csr(rows, cols, ptr, col, val), x, y));
}

extern "C" int orc_gs_color_d(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
                   const int32_t *order, int64_t begin, int64_t count, const double *rhs,
                   double *x, double omega) {
  ORC_TRY(if (!gs_color(make_csr(n, n, ptr, col, val), order, begin, count, rhs, x, omega))
              throw std::runtime_error("CPUBackend::encodeGaussSeidelColor: zero diagonal entry."));
}
extern "C" int orc_gs_color_f(int64_t n, const int32_t *ptr, const int32_t *col, const float *val,
                   const int32_t *order, int64_t begin, int64_t count, const float *rhs,
                   float *x, float omega) {
  ORC_TRY(if (!gs_color(make_csr(n, n, ptr, col, val), order, begin, count, rhs, x, omega))
              throw std::runtime_error("CPUBackend::encodeGaussSeidelColor: zero diagonal entry."));
}

extern "C" int orc_block_gs_color_d(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
                         const int32_t *order, const double *dinv, int64_t begin, int64_t end,
                         const double *rhs, double *x, double omega) {
  ORC_TRY(block_gs_color(make_csr(n, n, ptr, col, val), order, dinv, begin, end, rhs, x, omega));
}
extern "C" int orc_block_gs_color_f(int64_t n, const int32_t *ptr, const int32_t *col, const float *val,
                         const int32_t *order, const float *dinv, int64_t begin, int64_t end,
                         const float *rhs, float *x, float omega) {
  ORC_TRY(block_gs_color(make_csr(n, n, ptr, col, val), order, dinv, begin, end, rhs, x, omega));
}

extern "C" int orc_apply_block_inverses_d(const double *inv, const double *res, double *corr, int64_t n) {
  ORC_TRY(apply_block_inverses(inv, res, corr, n));
}
extern "C" int orc_apply_block_inverses_f(const float *inv, const float *res, float *corr, int64_t n) {
  ORC_TRY(apply_block_inverses(inv, res, corr, n));
}

extern "C" double orc_dot_d(const double *x, const double *y, int64_t n) { return dot(x, y, n); }
extern "C" float orc_dot_f(const float *x, const float *y, int64_t n) { return dot(x, y, n); }

extern "C" int orc_coloring(int64_t n, const int32_t *ptr, const int32_t *col, int32_t *colors,
                 int32_t *vertices, int32_t *offsets, int32_t *ncolors) {
  ORC_TRY({
    Coloring c = color_graph(n, ptr, col);
    std::memcpy(colors, c.colors.data(), n * sizeof(int32_t));
    std::memcpy(vertices, c.vertices.data(), n * sizeof(int32_t));
    std::memcpy(offsets, c.offsets.data(), (c.ncolors + 1) * sizeof(int32_t));
    *ncolors = c.ncolors;
  });
}

// src/ReorderingStrategies/RCMReordering.cc:11-110.
extern "C" int orc_rcm(int64_t n, const int32_t *ptr, const int32_t *col, int32_t *o2n, int32_t *n2o) {
  ORC_TRY({
    std::vector<int> deg(n);
    std::vector<bool> vis(n, false);
    for (int64_t r = 0; r < n; ++r) {
      int d = 0;
      for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) d += (col[k] != r);
      deg[r] = d;
    }
    std::queue<int> q;
    std::vector<int> nb, cm;
    const int *dp = deg.data();
    while (static_cast<int64_t>(cm.size()) < n) {
      int s = -1, md = std::numeric_limits<int>::max();
      for (int64_t r = 0; r < n; ++r)
        if (!vis[r] && deg[r] < md) { md = deg[r]; s = static_cast<int>(r); }
      q.push(s);
      vis[s] = true;
      while (!q.empty()) {
        int v = q.front();
        q.pop();
        cm.push_back(v);
        nb.clear();
        for (int32_t k = ptr[v]; k < ptr[v + 1]; ++k)
          if (col[k] != v) nb.push_back(col[k]);
        std::sort(nb.begin(), nb.end(), [dp](int a, int b) { return dp[a] < dp[b]; });
        for (int w : nb)
          if (!vis[w]) { vis[w] = true; q.push(w); }
      }
    }
    for (int64_t i = 0; i < n; ++i) n2o[i] = cm[n - 1 - i];
    for (int64_t i = 0; i < n; ++i) o2n[n2o[i]] = static_cast<int32_t>(i);
  });
}

// src/ReorderingStrategies/HierarchicalReordering.cc:13-52,74-200,217-278.
// Level k's matrix is a_ptr[k] / a_col[k] (n[k] rows, square), transfer k's
// restriction r_ptr[k] / r_col[k] / r_val[k] (n[k+1] x n[k]); the values are
// rounded to FP32 first, the reference's hierarchy type. o2n / n2o receive
// every level's permutation back to back, finest first; orphans[k] counts
// the DOFs of level k with no parent in R_k (the reference prints it).
extern "C" int orc_hierarchical_reordering(int32_t nlev, const int64_t *n, const int32_t *const *a_ptr,
                                           const int32_t *const *a_col, const int32_t *const *r_ptr,
                                           const int32_t *const *r_col, const double *const *r_val,
                                           int32_t *o2n, int32_t *n2o, int64_t *orphans) {
  ORC_TRY({
    if (nlev <= 0) throw std::runtime_error("HierarchicalReordering: empty hierarchy.");
    std::vector<int64_t> off(nlev + 1, 0);
    for (int32_t k = 0; k < nlev; ++k) off[k + 1] = off[k] + n[k];
    auto rcm_of = [&](int32_t k, std::vector<int32_t> &po, std::vector<int32_t> &pn) {
      po.resize(n[k]);
      pn.resize(n[k]);
      if (orc_rcm(n[k], a_ptr[k], a_col[k], po.data(), pn.data()) != 0) throw std::runtime_error(g_err);
    };
    // Coarsest level: plain RCM.
    std::vector<int32_t> co, cn;
    rcm_of(nlev - 1, co, cn);
    std::memcpy(o2n + off[nlev - 1], co.data(), n[nlev - 1] * sizeof(int32_t));
    std::memcpy(n2o + off[nlev - 1], cn.data(), n[nlev - 1] * sizeof(int32_t));
    for (int32_t k = nlev - 2; k >= 0; --k) {
      const int64_t nf = n[k], nc = n[k + 1];
      // Strongest parent: scan R row by row, keep the first largest |value|.
      std::vector<int> parent(nf, -1);
      std::vector<float> strongest(nf, -1.0f);
      for (int64_t c = 0; c < nc; ++c)
        for (int32_t e = r_ptr[k][c]; e < r_ptr[k][c + 1]; ++e) {
          const int32_t f = r_col[k][e];
          const float w = std::abs(static_cast<float>(r_val[k][e]));
          if (w > strongest[f]) {
            strongest[f] = w;
            parent[f] = static_cast<int>(c);
          }
        }
      int64_t orph = 0;
      for (int64_t f = 0; f < nf; ++f) orph += parent[f] < 0;
      orphans[k] = orph;
      std::vector<int32_t> fo, fn;
      rcm_of(k, fo, fn);
      const int32_t *cperm = o2n + off[k + 1];  // the coarse level's final permutation
      std::vector<int32_t> order(nf);
      for (int64_t i = 0; i < nf; ++i) order[i] = static_cast<int32_t>(i);
      std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const int64_t ra = parent[a] >= 0 ? cperm[parent[a]] : nc;
        const int64_t rb = parent[b] >= 0 ? cperm[parent[b]] : nc;
        if (ra != rb) return ra < rb;
        return fo[a] < fo[b];
      });
      for (int64_t i = 0; i < nf; ++i) {
        n2o[off[k] + i] = order[i];
        o2n[off[k] + order[i]] = static_cast<int32_t>(i);
      }
    }
  });
}

// src/HostELLMatrix.cc:10-206. ovf_col/ovf_val must hold the overflow count
// returned by a first call with null outputs.
extern "C" int orc_ell_pack_d(int64_t rows, const int32_t *ptr, const int32_t *col, const double *val,
                   int32_t width, int32_t *ell_col, double *ell_val, int32_t *ovf_ptr,
                   int32_t *ovf_col, double *ovf_val, int64_t *ovf_nnz) {
  ORC_TRY({
    if (width <= 0) throw std::runtime_error("HostELLMatrix: ELL width must be nonzero.");
    int64_t ov = 0;
    for (int64_t r = 0; r < rows; ++r) {
      const int len = ptr[r + 1] - ptr[r];
      if (len < 0) throw std::runtime_error("HostELLMatrix: invalid CSR row pointer.");
      if (len > width) ov += len - width;
    }
    *ovf_nnz = ov;
    if (!ell_col) return 0;
    for (int64_t i = 0; i < rows * width; ++i) { ell_col[i] = 0; ell_val[i] = 0.0; }
    ovf_ptr[0] = 0;
    int64_t pos = 0;
    for (int64_t r = 0; r < rows; ++r) {
      const int s = ptr[r], e = ptr[r + 1];
      const int m = std::min(width, e - s);
      for (int k = 0; k < m; ++k) {
        ell_col[static_cast<int64_t>(k) * rows + r] = col[s + k];
        ell_val[static_cast<int64_t>(k) * rows + r] = val[s + k];
      }
      for (int k = s + m; k < e; ++k, ++pos) { ovf_col[pos] = col[k]; ovf_val[pos] = val[k]; }
      ovf_ptr[r + 1] = static_cast<int32_t>(pos);
    }
  });
}

// src/HostSparseMatrix.cc:118-249 (row order kept, columns relabelled).
extern "C" int orc_permute_keep_order(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
                           const int32_t *o2n, const int32_t *n2o, int32_t *optr,
                           int32_t *ocol, double *oval) {
  ORC_TRY({
    optr[0] = 0;
    for (int64_t r = 0; r < n; ++r) optr[r + 1] = optr[r] + (ptr[n2o[r] + 1] - ptr[n2o[r]]);
    for (int64_t r = 0; r < n; ++r) {
      int d = optr[r];
      for (int k = ptr[n2o[r]]; k < ptr[n2o[r] + 1]; ++k, ++d) {
        ocol[d] = o2n[col[k]];
        oval[d] = val[k];
      }
    }
  });
}

// src/SparseMatrixUtils.cc:8-31. Output arrays sized for nnz (duplicates
// can only shrink it); *onnz receives the collapsed count.
extern "C" int orc_permute_sorted(int64_t rows, int64_t cols, const int32_t *ptr, const int32_t *col,
                       const double *val, const int32_t *ro2n, const int32_t *co2n,
                       int32_t *optr, int32_t *ocol, double *oval, int64_t *onnz) {
  ORC_TRY({
    CsrT<double> A = permute_sorted(rows, cols, ptr, col, val, ro2n, co2n);
    std::memcpy(optr, A.ptr.data(), (rows + 1) * sizeof(int32_t));
    std::memcpy(ocol, A.col.data(), A.col.size() * sizeof(int32_t));
    std::memcpy(oval, A.val.data(), A.val.size() * sizeof(double));
    *onnz = static_cast<int64_t>(A.col.size());
  });
}

extern "C" int orc_block_graph(int64_t n, const int32_t *ptr, const int32_t *col, int32_t bs,
                    int32_t *gptr, int32_t *gcol, int64_t *gnnz) {
  ORC_TRY({
    std::vector<int32_t> p, c;
    block_graph(n, ptr, col, bs, p, c);
    *gnnz = static_cast<int64_t>(c.size());
    if (gptr) {
      std::memcpy(gptr, p.data(), p.size() * sizeof(int32_t));
      std::memcpy(gcol, c.data(), c.size() * sizeof(int32_t));
    }
  });
}

// ---------------------------------------------------------------------------
// Full PCG with a multigrid or identity preconditioner.
//
// Hierarchy input (per level k < nlev): A_k CSR (FP64 values) in the
// hierarchy's ordering, the level permutation oldToNew perm[k] (or null for
// identity), P_k / R_k for k < nlev-1. Smoother kinds per level, omega,
// pre / post sweeps per level. The level matrices are permuted with the
// sorted flavour, as src/Smoothers/MultigridSmoother.cc:58-66,165-186; the
// CG matrix is the level-0 input permuted with the order-keeping flavour
// (src/CGSolver.cc:44-70) by perm[0].
//
// prec: 0 = identity preconditioner, 1 = multigrid.
// cg_double / mg_double select FP64 (1) or FP32 (0) for the CG vectors and
// the V-cycle. Output: x (original ordering, FP64), iterations in the
// reference convention, the final recursive relres^2, the relres^2 history
// (length maxit + 1, first entry before the loop) and the number of
// entries written.
// ---------------------------------------------------------------------------
struct OrcHier {
  int32_t nlev;
  const int64_t *n;
  const int32_t *const *aptr;
  const int32_t *const *acol;
  const double *const *aval;
  const int32_t *const *perm;
  const int64_t *pcols;  // unused placeholder for symmetry
  const int32_t *const *pptr;
  const int32_t *const *pcol;
  const double *const *pval;
  const int32_t *const *rptr;
  const int32_t *const *rcol;
  const double *const *rval;
};

template <class T>
static CsrT<T> csr_cast(const CsrT<double> &A) {
  CsrT<T> B;
  B.rows = A.rows;
  B.cols = A.cols;
  B.ptr = A.ptr;
  B.col = A.col;
  B.val.resize(A.val.size());
  for (size_t i = 0; i < A.val.size(); ++i) B.val[i] = static_cast<T>(A.val[i]);
  return B;
}

template <class TM>
static void build_mg(MG<TM> &mg, const OrcHier &h, const int32_t *kinds, const double *omegas,
                     const int32_t *pres, const int32_t *posts, int32_t coarse_steps,
                     int32_t direct_coarse, int32_t vcycles, int32_t ref_alias,
                     const double *coarse_inv) {
  const int nl = h.nlev;
  mg.L.resize(nl);
  mg.reference_aliasing = ref_alias != 0;
  mg.coarse_steps = coarse_steps;
  mg.direct_coarse = direct_coarse != 0;
  mg.vcycles = vcycles;
  std::vector<std::vector<int32_t>> ident(nl);
  auto o2n = [&](int k) -> const int32_t * {
    if (h.perm && h.perm[k]) return h.perm[k];
    if (ident[k].empty()) {
      ident[k].resize(h.n[k]);
      for (int64_t i = 0; i < h.n[k]; ++i) ident[k][i] = static_cast<int32_t>(i);
    }
    return ident[k].data();
  };
  for (int k = 0; k < nl; ++k) {
    const int64_t n = h.n[k];
    CsrT<double> Ad = permute_sorted(n, n, h.aptr[k], h.acol[k], h.aval[k], o2n(k), o2n(k));
    Level<TM> &lv = mg.L[k];
    lv.A = csr_cast<TM>(Ad);
    lv.kind = kinds[k];
    lv.omega = static_cast<TM>(omegas[k]);
    lv.pre = pres[k];
    lv.post = posts[k];
    lv.unsmoothed = k > 0 && k + 1 < nl && pres[k] == 0 && posts[k] == 0;
    if (lv.kind == kSGS) {
      if (!(lv.omega > TM(0) && lv.omega < TM(2)))
        throw std::runtime_error("SymmetricGaussSeidelSmoother: invalid omega.");
      lv.col = color_graph(n, lv.A.ptr.data(), lv.A.col.data());
    } else if (lv.kind == kBlockGS || lv.kind == kBlockJacobi) {
      if (n % 3 != 0)
        throw std::runtime_error("BlockGaussSeidelSmoother: number of DOFs must be divisible by 3.");
      // Inverse diagonal blocks in FP64, rounded to the V-cycle type.
      const int64_t nb = n / 3;
      lv.dinv.resize(9 * nb);
      for (int64_t b = 0; b < nb; ++b) {
        std::vector<double> blk(9, 0.0);
        for (int l = 0; l < 3; ++l)
          for (int32_t e = Ad.ptr[3 * b + l]; e < Ad.ptr[3 * b + l + 1]; ++e) {
            const int64_t c = Ad.col[e];
            if (c >= 3 * b && c < 3 * b + 3) blk[3 * l + (c - 3 * b)] = Ad.val[e];
          }
        double inv3[9];
        if (!inverse3x3_cholesky(blk.data(), inv3)) throw std::runtime_error("Block diagonal is not SPD.");
        for (int i = 0; i < 9; ++i) lv.dinv[9 * b + i] = static_cast<TM>(inv3[i]);
      }
      if (lv.kind == kBlockGS) {
        if (!(lv.omega > TM(0) && lv.omega < TM(2)))
          throw std::runtime_error("BlockGaussSeidelSmoother: omega must be in (0, 2).");
        std::vector<int32_t> gp, gc;
        block_graph(n, lv.A.ptr.data(), lv.A.col.data(), 3, gp, gc);
        lv.col = color_graph(n / 3, gp.data(), gc.data());
      }
    } else {
      lv.dinv.resize(n);
      for (int64_t r = 0; r < n; ++r) {
        bool found = false;
        for (int32_t e = Ad.ptr[r]; e < Ad.ptr[r + 1]; ++e)
          if (Ad.col[e] == r) {
            if (std::abs(Ad.val[e]) < 1e-12)
              throw std::runtime_error("DampedJacobiPreconditioner::initialize: zero diagonal entry.");
            lv.dinv[r] = TM(1) / static_cast<TM>(Ad.val[e]);
            found = true;
            break;
          }
        if (!found)
          throw std::runtime_error("DampedJacobiPreconditioner::initialize: missing diagonal entry.");
      }
    }
    lv.rhs.assign(n, TM(0));
    lv.res.assign(n, TM(0));
    lv.corr.assign(n, TM(0));
    lv.az.assign(n, TM(0));
    lv.tmp.assign(n, TM(0));
    lv.jaz.assign(n, TM(0));
    lv.jcorr.assign(n, TM(0));
    if (k > 0) {
      const int64_t nf = h.n[k - 1];
      CsrT<double> Pd = permute_sorted(nf, n, h.pptr[k - 1], h.pcol[k - 1], h.pval[k - 1],
                                       o2n(k - 1), o2n(k));
      CsrT<double> Rd = permute_sorted(n, nf, h.rptr[k - 1], h.rcol[k - 1], h.rval[k - 1],
                                       o2n(k), o2n(k - 1));
      mg.P.push_back(csr_cast<TM>(Pd));
      mg.R.push_back(csr_cast<TM>(Rd));
    }
    if (k == nl - 1 && mg.direct_coarse && coarse_inv) {
      // A caller-supplied FP64 inverse of the coarsest matrix (computed by
      // LAPACK for levels too large for the scalar Cholesky below).
      mg.coarse_inv.assign(coarse_inv, coarse_inv + n * n);
    } else if (k == nl - 1 && mg.direct_coarse) {
      mg.coarse_inv.assign(n * n, 0.0);
      for (int64_t r = 0; r < n; ++r)
        for (int32_t e = Ad.ptr[r]; e < Ad.ptr[r + 1]; ++e) mg.coarse_inv[r * n + Ad.col[e]] += Ad.val[e];
      if (!dense_spd_inverse(n, mg.coarse_inv))
        throw std::runtime_error("coarse level matrix is not SPD.");
    }
  }
}

template <class TC, class TM>
static void pcg(const OrcHier &h, int prec, const int32_t *kinds, const double *omegas,
                const int32_t *pres, const int32_t *posts, int32_t coarse_steps,
                int32_t direct_coarse, int32_t vcycles, int32_t ref_alias,
                const double *coarse_inv, const double *b_in,
                double *x_io, double tol, int64_t maxit, int64_t *iters, double *relres2,
                double *history, int64_t *hist_len) {
  const int64_t n = h.n[0];
  MG<TM> mg;
  if (prec == 1)
    build_mg(mg, h, kinds, omegas, pres, posts, coarse_steps, direct_coarse, vcycles, ref_alias,
             coarse_inv);
  // CG matrix: level-0 input with the order-keeping permutation.
  std::vector<int32_t> ident;
  const int32_t *o2n = nullptr, *n2o = nullptr;
  std::vector<int32_t> n2o_v;
  if (h.perm && h.perm[0]) {
    o2n = h.perm[0];
    n2o_v.resize(n);
    for (int64_t i = 0; i < n; ++i) n2o_v[o2n[i]] = static_cast<int32_t>(i);
    n2o = n2o_v.data();
  }
  CsrT<TC> A;
  {
    CsrT<double> Ad;
    Ad.rows = Ad.cols = n;
    Ad.ptr.assign(n + 1, 0);
    const int32_t *p = h.aptr[0];
    for (int64_t r = 0; r < n; ++r) {
      const int64_t o = n2o ? n2o[r] : r;
      Ad.ptr[r + 1] = Ad.ptr[r] + (p[o + 1] - p[o]);
    }
    Ad.col.resize(Ad.ptr[n]);
    Ad.val.resize(Ad.ptr[n]);
    for (int64_t r = 0; r < n; ++r) {
      const int64_t o = n2o ? n2o[r] : r;
      int d = Ad.ptr[r];
      for (int32_t k = p[o]; k < p[o + 1]; ++k, ++d) {
        Ad.col[d] = o2n ? o2n[h.acol[0][k]] : h.acol[0][k];
        Ad.val[d] = h.aval[0][k];
      }
    }
    A = csr_cast<TC>(Ad);
  }
  std::vector<TC> b(n), x(n);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t o = n2o ? n2o[i] : i;
    b[i] = static_cast<TC>(b_in[o]);
    x[i] = static_cast<TC>(x_io[o]);
  }
  *hist_len = 0;
  // src/CGSolver.cc:113-118: ||b|| below 1e-12 returns x = 0.
  {
    double nb = 0.0;
    for (int64_t i = 0; i < n; ++i) nb += b_in[i] * b_in[i];
    if (std::sqrt(nb) < 1e-12) {
      for (int64_t i = 0; i < n; ++i) x_io[i] = 0.0;
      *iters = 0;
      *relres2 = 0.0;
      return;
    }
  }
  std::vector<TC> r(n), rn(n), z(n), zn(n), p(n), q(n);
  std::vector<TM> rm(n), zm(n);
  auto apply_prec = [&](const std::vector<TC> &rr, std::vector<TC> &zz) {
    if (prec == 0) {
      zz = rr;
      return;
    }
    for (int64_t i = 0; i < n; ++i) rm[i] = static_cast<TM>(rr[i]);
    std::fill(zm.begin(), zm.end(), TM(0));
    for (int v = 0; v < mg.vcycles; ++v) vcycle(mg, 0, rm.data(), zm.data());
    for (int64_t i = 0; i < n; ++i) zz[i] = static_cast<TC>(zm[i]);
  };
  spmv(A, x.data(), q.data());
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] + TC(-1) * q[i];
  apply_prec(r, z);
  p = z;
  const TC bb = dot(b.data(), b.data(), n);
  TC rho = dot(r.data(), z.data(), n);
  TC rr = dot(r.data(), r.data(), n);
  TC rel = rr / bb;
  history[(*hist_len)++] = static_cast<double>(rel);
  if (!std::isfinite(static_cast<double>(rel)))
    throw std::runtime_error("CGSolver: initial residual is NaN/Inf - check matrix and RHS.");
  const TC tol2 = static_cast<TC>(tol) * static_cast<TC>(tol);
  if (rel < tol2) {
    *iters = 0;
    *relres2 = static_cast<double>(rel);
    for (int64_t i = 0; i < n; ++i) x_io[n2o ? n2o[i] : i] = static_cast<double>(x[i]);
    return;
  }
  *iters = maxit;
  for (int64_t it = 0; it < maxit; ++it) {
    spmv(A, p.data(), q.data());
    const TC pap = dot(p.data(), q.data(), n);
    const TC alpha = rho / pap;
    axpy(alpha, p.data(), x.data(), n);
    const TC na = -alpha;
    for (int64_t i = 0; i < n; ++i) rn[i] = r[i] + na * q[i];
    rr = dot(rn.data(), rn.data(), n);
    rel = rr / bb;
    apply_prec(rn, zn);
    const TC rhon = dot(rn.data(), zn.data(), n);
    const TC beta = rhon / rho;
    for (int64_t i = 0; i < n; ++i) p[i] *= beta;
    axpy(TC(1), zn.data(), p.data(), n);
    history[(*hist_len)++] = static_cast<double>(rel);
    if (!std::isfinite(static_cast<double>(rel)))
      throw std::runtime_error("CGSolver: solver diverged (NaN/Inf residual).");
    if (rel < tol2) {
      *iters = it;
      break;
    }
    rho = rhon;
    r.swap(rn);
    z.swap(zn);
  }
  *relres2 = static_cast<double>(rel);
  for (int64_t i = 0; i < n; ++i) x_io[n2o ? n2o[i] : i] = static_cast<double>(x[i]);
}

extern "C" int orc_pcg(const OrcHier *h, int32_t prec, int32_t cg_double, int32_t mg_double,
            const int32_t *kinds, const double *omegas, const int32_t *pres,
            const int32_t *posts, int32_t coarse_steps, int32_t direct_coarse,
            int32_t vcycles, int32_t ref_alias, const double *coarse_inv, const double *b,
            double *x, double tol, int64_t maxit, int64_t *iters, double *relres2,
            double *history, int64_t *hist_len) {
  ORC_TRY({
    if (cg_double && mg_double)
      pcg<double, double>(*h, prec, kinds, omegas, pres, posts, coarse_steps, direct_coarse,
                          vcycles, ref_alias, coarse_inv, b, x, tol, maxit, iters, relres2, history,
                          hist_len);
    else if (cg_double && !mg_double)
      pcg<double, float>(*h, prec, kinds, omegas, pres, posts, coarse_steps, direct_coarse,
                         vcycles, ref_alias, coarse_inv, b, x, tol, maxit, iters, relres2, history,
                          hist_len);
    else if (!cg_double && !mg_double)
      pcg<float, float>(*h, prec, kinds, omegas, pres, posts, coarse_steps, direct_coarse,
                        vcycles, ref_alias, coarse_inv, b, x, tol, maxit, iters, relres2, history,
                          hist_len);
    else
      throw std::runtime_error("orc_pcg: FP32 CG with an FP64 V-cycle is not a supported mode.");
  });
}

// One preconditioner application z = M r on level-0 ordering (after the
// level permutation), for V-cycle parity tests.
extern "C" int orc_vcycle(const OrcHier *h, int32_t mg_double, const int32_t *kinds,
               const double *omegas, const int32_t *pres, const int32_t *posts,
               int32_t coarse_steps, int32_t direct_coarse, int32_t vcycles,
               int32_t ref_alias, const double *coarse_inv, const double *r, double *z) {
  ORC_TRY({
    const int64_t n = h->n[0];
    if (mg_double) {
      MG<double> mg;
      build_mg(mg, *h, kinds, omegas, pres, posts, coarse_steps, direct_coarse, vcycles, ref_alias,
               coarse_inv);
      std::vector<double> zz(n, 0.0);
      for (int v = 0; v < vcycles; ++v) vcycle(mg, 0, r, zz.data());
      std::memcpy(z, zz.data(), n * sizeof(double));
    } else {
      MG<float> mg;
      build_mg(mg, *h, kinds, omegas, pres, posts, coarse_steps, direct_coarse, vcycles, ref_alias,
               coarse_inv);
      std::vector<float> rr(n), zz(n, 0.0f);
      for (int64_t i = 0; i < n; ++i) rr[i] = static_cast<float>(r[i]);
      for (int v = 0; v < vcycles; ++v) vcycle(mg, 0, rr.data(), zz.data());
      for (int64_t i = 0; i < n; ++i) z[i] = zz[i];
    }
  });
}

