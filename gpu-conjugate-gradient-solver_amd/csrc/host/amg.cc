// Smoothed-aggregation hierarchy. The reference obtains its hierarchy from
// the third-party amgcl library (smoothed_aggregation<builtin<float>>, with
// aggr.block_size and aggr.eps_strong set, called from
// examples/amgSolver/src/SmoothedAggregationCoarsener.cc:64-108, levels
// driven by examples/amgSolver/src/AMGHierarchyBuilder.cc:58-148). amgcl is
// not vendored in the reference, so this is a restatement of amgcl's
// published smoothed-aggregation algorithm:
//   1. strength: off-diagonal a_ij is strong when
//      eps^2 * |a_ii * a_jj| < a_ij^2 (on the block-max "pointwise" matrix
//      when block_size > 1);
//   2. plain aggregation in index order: rows without strong connections are
//      left out; an unassigned row seeds a new aggregate, takes its strong
//      neighbours, then their still-undefined strong neighbours; emptied
//      aggregates are compacted away;
//   3. tentative prolongation: piecewise constant, or the per-aggregate QR
//      of the near-nullspace (Q becomes P's rows, R the coarse nullspace);
//   4. prolongation smoothing P = (I - w D_f^-1 A_f) P_tent with the
//      filtered matrix A_f (weak connections lumped into the diagonal) and
//      w = relax * 2/3;
//   5. R = P^T and the Galerkin product A_c = R A P.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace gsx {
namespace {

// No strong connection anywhere on a level: nothing to aggregate. The
// hierarchy stops there (amgcl's amg builder ends coarsening on its
// empty_level error the same way).
struct EmptyLevel : std::runtime_error {
  EmptyLevel() : std::runtime_error("smoothed aggregation: empty level.") {}
};

constexpr int32_t kUndefined = -1;
constexpr int32_t kRemoved = -2;

struct Aggregates {
  int64_t count = 0;               // number of aggregates (scalar sense)
  std::vector<int32_t> id;         // per row of the aggregated matrix
  std::vector<char> strong;        // per nonzero of the aggregated matrix
};

std::vector<double> diagonal_of(const Csr &A) {
  std::vector<double> d(A.rows, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A.rows; ++r)
    for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k)
      if (A.col[k] == r) d[r] += A.val[k];
  return d;
}

Aggregates plain_aggregates(const Csr &A, double eps) {
  Aggregates ag;
  const int64_t n = A.rows;
  const double eps2 = eps * eps;
  const std::vector<double> dia = diagonal_of(A);
  ag.strong.assign(A.nnz(), 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double ei = eps2 * dia[i];
    for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t c = A.col[k];
      const double v = A.val[k];
      ag.strong[k] = (c != i) && (ei * dia[c] < v * v);
    }
  }
  ag.id.assign(n, kRemoved);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
      if (ag.strong[k]) { ag.id[i] = kUndefined; break; }
  std::vector<int32_t> neib;
  int64_t count = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (ag.id[i] != kUndefined) continue;
    const int32_t cur = static_cast<int32_t>(count++);
    ag.id[i] = cur;
    neib.clear();
    for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t c = A.col[k];
      if (ag.strong[k] && ag.id[c] != kRemoved) {
        ag.id[c] = cur;
        neib.push_back(c);
      }
    }
    for (int32_t c : neib)
      for (int32_t k = A.ptr[c]; k < A.ptr[c + 1]; ++k) {
        const int32_t cc = A.col[k];
        if (ag.strong[k] && ag.id[cc] == kUndefined) ag.id[cc] = cur;
      }
  }
  if (count == 0) throw EmptyLevel();
  std::vector<int32_t> keep(count, 0);
  for (int32_t v : ag.id)
    if (v >= 0) keep[v] = 1;
  std::partial_sum(keep.begin(), keep.end(), keep.begin());
  if (keep.back() < count) {
    for (auto &v : ag.id)
      if (v >= 0) v = keep[v] - 1;
    count = keep.back();
  }
  ag.count = count;
  return ag;
}

// Merge aggregates with fewer than min_points members into the aggregate of
// their strongest already-aggregated neighbour (or drop them when none), so
// the per-aggregate QR always has at least as many rows as nullspace
// columns.
void merge_small_aggregates(const Csr &A, Aggregates &ag, int64_t min_points) {
  if (min_points <= 1) return;
  std::vector<int64_t> size(ag.count, 0);
  for (int32_t v : ag.id)
    if (v >= 0) size[v]++;
  bool any = false;
  for (int64_t s : size)
    if (s > 0 && s < min_points) { any = true; break; }
  if (!any) return;
  for (int64_t i = 0; i < A.rows; ++i) {
    const int32_t a = ag.id[i];
    if (a < 0 || size[a] >= min_points) continue;
    int32_t best = -1;
    double bestv = -1.0;
    for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t c = A.col[k];
      const int32_t ac = ag.id[c];
      if (c == i || ac < 0 || ac == a || size[ac] < min_points) continue;
      const double v = std::abs(A.val[k]);
      if (v > bestv) { bestv = v; best = ac; }
    }
    size[a]--;
    if (best >= 0) { ag.id[i] = best; size[best]++; }
    else ag.id[i] = kRemoved;
  }
  std::vector<int32_t> keep(ag.count, 0);
  for (int32_t v : ag.id)
    if (v >= 0) keep[v] = 1;
  std::partial_sum(keep.begin(), keep.end(), keep.begin());
  for (auto &v : ag.id)
    if (v >= 0) v = keep[v] - 1;
  ag.count = keep.empty() ? 0 : keep.back();
}

// Block-max "pointwise" matrix of a block_size-blocked scalar matrix: row
// ip holds the distinct column blocks of rows ip*bs .. ip*bs+bs-1 in
// ascending order, each with the largest |a| of its entries. Two parallel
// passes (distinct count, then fill and sort) straight into the CSR arrays.
Csr pointwise_matrix(const Csr &A, int32_t bs) {
  const int64_t np = A.rows / bs;
  Csr P;
  P.rows = P.cols = np;
  P.ptr.assign(np + 1, 0);
#pragma omp parallel
  {
    std::vector<int32_t> mark(np, -1);
#pragma omp for schedule(static)
    for (int64_t ip = 0; ip < np; ++ip) {
      int32_t cnt = 0;
      for (int64_t r = ip * bs; r < (ip + 1) * bs; ++r)
        for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
          const int32_t cp = A.col[k] / bs;
          if (mark[cp] != static_cast<int32_t>(ip)) {
            mark[cp] = static_cast<int32_t>(ip);
            ++cnt;
          }
        }
      P.ptr[ip + 1] = cnt;
    }
  }
  for (int64_t ip = 0; ip < np; ++ip) P.ptr[ip + 1] += P.ptr[ip];
  P.col.resize(P.ptr[np]);
  P.val.resize(P.ptr[np]);
#pragma omp parallel
  {
    std::vector<int32_t> slot(np, -1);
    std::vector<std::pair<int32_t, double>> row;
#pragma omp for schedule(static)
    for (int64_t ip = 0; ip < np; ++ip) {
      row.clear();
      for (int64_t r = ip * bs; r < (ip + 1) * bs; ++r)
        for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
          const int32_t cp = A.col[k] / bs;
          const double v = std::abs(A.val[k]);
          if (slot[cp] < 0) {
            slot[cp] = static_cast<int32_t>(row.size());
            row.emplace_back(cp, v);
          } else {
            row[slot[cp]].second = std::max(row[slot[cp]].second, v);
          }
        }
      for (auto &e : row) slot[e.first] = -1;
      std::sort(row.begin(), row.end());
      int64_t d = P.ptr[ip];
      for (auto &e : row) {
        P.col[d] = e.first;
        P.val[d] = e.second;
        ++d;
      }
    }
  }
  return P;
}

// Householder QR of an m x k column-major matrix b (m >= k): b = Q R with
// Q m x k orthonormal, R k x k upper triangular.
void householder_qr(int m, int k, std::vector<double> &b, std::vector<double> &q,
                    std::vector<double> &r) {
  std::vector<double> a = b;  // m x k col-major, becomes R / reflectors
  std::vector<double> tau(k, 0.0);
  for (int j = 0; j < k; ++j) {
    double norm = 0.0;
    for (int i = j; i < m; ++i) norm += a[j * m + i] * a[j * m + i];
    norm = std::sqrt(norm);
    if (norm == 0.0) { tau[j] = 0.0; continue; }
    const double alpha = a[j * m + j] > 0 ? -norm : norm;
    const double v0 = a[j * m + j] - alpha;
    // v = (1, a[j+1..]/v0), tau = (alpha - a_jj)/(-alpha) standard form
    for (int i = j + 1; i < m; ++i) a[j * m + i] /= v0;
    tau[j] = (alpha - a[j * m + j]) / alpha;
    a[j * m + j] = alpha;
    for (int c = j + 1; c < k; ++c) {
      double s = a[c * m + j];
      for (int i = j + 1; i < m; ++i) s += a[j * m + i] * a[c * m + i];
      s *= tau[j];
      a[c * m + j] -= s;
      for (int i = j + 1; i < m; ++i) a[c * m + i] -= s * a[j * m + i];
    }
  }
  r.assign(k * k, 0.0);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i <= j; ++i) r[i * k + j] = a[j * m + i];  // row-major R
  // Q = H_0 ... H_{k-1} applied to the first k unit columns.
  q.assign(static_cast<size_t>(m) * k, 0.0);  // col-major m x k
  for (int j = 0; j < k; ++j) q[j * m + j] = 1.0;
  for (int j = k - 1; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    for (int c = 0; c < k; ++c) {
      double s = q[c * m + j];
      for (int i = j + 1; i < m; ++i) s += a[j * m + i] * q[c * m + i];
      s *= tau[j];
      q[c * m + j] -= s;
      for (int i = j + 1; i < m; ++i) q[c * m + i] -= s * a[j * m + i];
    }
  }
}

// A copy of A whose entry arrays are copied by all threads (a large level-0
// matrix is over a gigabyte).
Csr parallel_copy(const Csr &A) {
  Csr C;
  C.rows = A.rows;
  C.cols = A.cols;
  C.ptr = A.ptr;
  const int64_t nnz = A.nnz();
  C.col.resize(nnz);
  C.val.resize(nnz);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nnz; ++k) {
    C.col[k] = A.col[k];
    C.val[k] = A.val[k];
  }
  return C;
}

struct CoarsenResult {
  Csr P, R, Ac;
  std::vector<int32_t> agg;      // per fine DOF
  std::vector<double> nullspace;  // coarse nullspace (nc x ncols row-major)
};

// The prolongation smoother M = I - w D_f^-1 A_f as an explicit matrix: row
// i keeps, in A's entry order, its diagonal entry (value 1 - w) and its
// strong off-diagonal entries (value -w a_ij / d_i), d_i being the filtered
// diagonal (diagonal plus the weak entries lumped into it). The products
// of M P_tent then run in the same order as a fused loop over A's entries.
Csr smoothing_operator(const Csr &A, const std::vector<char> &strong, double omega) {
  const int64_t n = A.rows;
  Csr M;
  M.rows = n;
  M.cols = A.cols;
  M.ptr.assign(n + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int32_t cnt = 0;
    for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) cnt += (A.col[k] == i || strong[k]) ? 1 : 0;
    M.ptr[i + 1] = cnt;
  }
  for (int64_t i = 0; i < n; ++i) M.ptr[i + 1] += M.ptr[i];
  M.col.resize(M.ptr[n]);
  M.val.resize(M.ptr[n]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    double dia = 0.0;
    for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
      if (A.col[k] == i || !strong[k]) dia += A.val[k];
    const double scale = -omega / dia;
    int32_t d = M.ptr[i];
    for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t ca = A.col[k];
      if (ca != i && !strong[k]) continue;
      M.col[d] = ca;
      M.val[d] = (ca == i) ? (1.0 - omega) : scale * A.val[k];
      ++d;
    }
  }
  return M;
}

CoarsenResult coarsen(const Csr &A, const SAParams &prm,
                      const std::vector<double> &B, int32_t ncols, const GalerkinOps &ops) {
  const int32_t bs = prm.block_size;
  if (A.rows % bs != 0)
    throw std::runtime_error(
        "SmoothedAggregationCoarsener: matrix size is not divisible by block size.");
  const int64_t n = A.rows;
  // ---- aggregates ---------------------------------------------------------
  std::vector<int32_t> point_id;  // per point
  int64_t point_count = 0;
  std::vector<char> strong(A.nnz(), 0);
  if (bs == 1) {
    Aggregates ag = plain_aggregates(A, prm.eps_strong);
    if (ncols > 0) merge_small_aggregates(A, ag, ncols);
    point_id = ag.id;
    point_count = ag.count;
    strong = ag.strong;
  } else {
    Csr Ap = pointwise_matrix(A, bs);
    Aggregates ag = plain_aggregates(Ap, prm.eps_strong);
    if (ncols > 0) merge_small_aggregates(Ap, ag, (ncols + bs - 1) / bs);
    point_id = ag.id;
    point_count = ag.count;
    // Expand the pointwise strength to scalar entries: couplings inside the
    // point's own block are strong, others inherit the pointwise flag.
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < Ap.rows; ++ip) {
      for (int32_t l = 0; l < bs; ++l) {
        const int64_t r = ip * bs + l;
        for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
          const int32_t cp = A.col[k] / bs;
          bool s = (cp == ip);
          if (!s) {
            const int32_t *b = &Ap.col[Ap.ptr[ip]];
            const int32_t *e = &Ap.col[Ap.ptr[ip + 1]];
            const int32_t *f = std::lower_bound(b, e, cp);
            s = ag.strong[Ap.ptr[ip] + (f - b)] != 0;
          }
          strong[k] = s && (A.col[k] != r);
        }
      }
    }
  }
  if (point_count == 0) throw EmptyLevel();

  CoarsenResult res;
  res.agg.assign(n, -1);
  // ---- tentative prolongation -----------------------------------------------
  Csr Pt;
  Pt.rows = n;
  Pt.ptr.assign(n + 1, 0);
  if (ncols == 0) {
    const int64_t nc = point_count * bs;
    Pt.cols = nc;
    for (int64_t i = 0; i < n; ++i) {
      const int32_t p = point_id[i / bs];
      res.agg[i] = p >= 0 ? static_cast<int32_t>(p * bs + i % bs) : -1;
      Pt.ptr[i + 1] = Pt.ptr[i] + (p >= 0 ? 1 : 0);
    }
    Pt.col.resize(Pt.ptr[n]);
    Pt.val.assign(Pt.ptr[n], 1.0);
    for (int64_t i = 0; i < n; ++i)
      if (res.agg[i] >= 0) Pt.col[Pt.ptr[i]] = res.agg[i];
  } else {
    const int64_t nc = point_count * ncols;
    Pt.cols = nc;
    for (int64_t i = 0; i < n; ++i) {
      const int32_t p = point_id[i / bs];
      res.agg[i] = p;
      Pt.ptr[i + 1] = Pt.ptr[i] + (p >= 0 ? ncols : 0);
    }
    Pt.col.resize(Pt.ptr[n]);
    Pt.val.resize(Pt.ptr[n]);
    // DOFs of each aggregate in ascending order.
    std::vector<int32_t> aptr(point_count + 1, 0);
    for (int64_t i = 0; i < n; ++i)
      if (res.agg[i] >= 0) aptr[res.agg[i] + 1]++;
    for (int64_t a = 0; a < point_count; ++a) aptr[a + 1] += aptr[a];
    std::vector<int32_t> members(aptr[point_count]);
    {
      std::vector<int32_t> next(aptr.begin(), aptr.end() - 1);
      for (int64_t i = 0; i < n; ++i)
        if (res.agg[i] >= 0) members[next[res.agg[i]]++] = static_cast<int32_t>(i);
    }
    res.nullspace.assign(nc * ncols, 0.0);
    for (int64_t a = 0; a < point_count; ++a)
      if (aptr[a + 1] - aptr[a] < ncols)
        throw std::runtime_error("smoothed aggregation: aggregate too small.");
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t a = 0; a < point_count; ++a) {
      const int m = aptr[a + 1] - aptr[a];
      std::vector<double> bpart(static_cast<size_t>(m) * ncols), q, r;
      for (int jj = 0; jj < m; ++jj) {
        const int32_t i = members[aptr[a] + jj];
        for (int c = 0; c < ncols; ++c) bpart[c * m + jj] = B[static_cast<size_t>(i) * ncols + c];
      }
      householder_qr(m, ncols, bpart, q, r);
      for (int ii = 0; ii < ncols; ++ii)
        for (int jj = 0; jj < ncols; ++jj)
          res.nullspace[(a * ncols + ii) * ncols + jj] = r[ii * ncols + jj];
      for (int jj = 0; jj < m; ++jj) {
        const int32_t i = members[aptr[a] + jj];
        for (int c = 0; c < ncols; ++c) {
          Pt.col[Pt.ptr[i] + c] = static_cast<int32_t>(a * ncols + c);
          Pt.val[Pt.ptr[i] + c] = q[c * m + jj];
        }
      }
    }
  }
  // ---- prolongation smoothing and Galerkin product ----------------------------
  // P = M P_tent with M = I - w D_f^-1 A_f, then R = P^T and A_c = R (A P).
  const double omega = prm.relax * (2.0 / 3.0);
  const Csr M = smoothing_operator(A, strong, omega);
  ops.run(A, M, Pt, res.P, res.R, res.Ac);
  return res;
}

}  // namespace

void GalerkinOps::run(const Csr &A, const Csr &M, const Csr &Pt, Csr &P, Csr &R,
                      Csr &Ac) const {
  P = spgemm(M, Pt);
  R = transpose(P);
  const Csr AP = spgemm(A, P);
  Ac = spgemm(R, AP);
}

Hierarchy build_sa_hierarchy(const Csr &A, const SAParams &prm,
                             const std::vector<double> &nullspace,
                             int32_t ncols, const GalerkinOps *ops) {
  const GalerkinOps host_ops;
  const GalerkinOps &gops = ops ? *ops : host_ops;
  if (A.rows != A.cols) throw std::runtime_error("AMGHierarchyBuilder: matrix must be square.");
  if (prm.max_levels <= 0) throw std::runtime_error("AMGHierarchyBuilder: maxLevels must be positive.");
  if (prm.min_dofs <= 0) throw std::runtime_error("AMGHierarchyBuilder: minDofs must be positive.");
  if (ncols > 0 && static_cast<int64_t>(nullspace.size()) != A.rows * ncols)
    throw std::runtime_error("AMGHierarchyBuilder: near-nullspace row count mismatch.");
  Hierarchy H;
  H.A.push_back(parallel_copy(A));
  std::vector<double> B = nullspace;
  SAParams lp = prm;
  for (int32_t level = 0; level < prm.max_levels; ++level) {
    const Csr &cur = H.A.back();
    if (cur.rows <= prm.min_dofs) break;
    CoarsenResult r;
    try {
      r = coarsen(cur, lp, B, ncols, gops);
    } catch (const EmptyLevel &) {
      break;
    }
    // With a near-nullspace every aggregate becomes one coarse node of
    // ncols unknowns, so coarser levels aggregate ncols-blocks (amgcl's
    // smoothed_aggregation switches aggr.block_size to nullspace.cols the
    // same way).
    if (ncols > 0) lp.block_size = ncols;
    // amgcl halves the strength threshold after every level
    // (aggr.eps_strong *= 0.5 in transfer_operators, carried across levels
    // by its multilevel driver): Galerkin operators spread their couplings
    // over more neighbours, each weaker than on the level above.
    lp.eps_strong *= 0.5;
    if (r.Ac.rows == 0 || r.Ac.rows >= cur.rows) break;  // no progress
    H.P.push_back(std::move(r.P));
    H.R.push_back(std::move(r.R));
    H.aggregates.push_back(std::move(r.agg));
    H.A.push_back(std::move(r.Ac));
    if (ncols > 0) B.swap(r.nullspace);
  }
  return H;
}

std::vector<double> level_features(const Hierarchy &H, int32_t block_size) {
  std::vector<double> f;
  for (size_t k = 0; k < H.A.size(); ++k) {
    const Csr &A = H.A[k];
    double dd = 0.0, dmax = 0.0, dmin = 1e300;
    for (int64_t r = 0; r < A.rows; ++r) {
      double d = 0.0, off = 0.0;
      for (int32_t e = A.ptr[r]; e < A.ptr[r + 1]; ++e) {
        if (A.col[e] == r) d += std::abs(A.val[e]);
        else off += std::abs(A.val[e]);
      }
      dd += off > 0.0 ? std::min(10.0, d / off) : 10.0;
      dmax = std::max(dmax, d);
      if (d > 0.0) dmin = std::min(dmin, d);
    }
    int32_t colors = 0;
    if (block_size == 3 && A.rows % 3 == 0) {
      Graph g = block_graph(A, 3);
      colors = greedy_coloring(A.rows / 3, g.ptr.data(), g.col.data()).ncolors;
    } else {
      colors = greedy_coloring(A.rows, A.ptr.data(), A.col.data()).ncolors;
    }
    f.push_back(static_cast<double>(A.nnz()) / static_cast<double>(A.rows));
    f.push_back(dd / static_cast<double>(A.rows));
    f.push_back(k == 0 ? 1.0 : static_cast<double>(A.rows) / static_cast<double>(H.A[k - 1].rows));
    f.push_back(dmax > 0.0 && dmin < 1e300 ? std::log10(dmax / dmin) : 0.0);
    f.push_back(static_cast<double>(colors));
  }
  return f;
}

}  // namespace gsx
