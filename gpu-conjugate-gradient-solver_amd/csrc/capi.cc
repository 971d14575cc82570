// C ABI over the host setup library and the solve engine
// (declarations and reference citations in include/gpusolver.h).
#include "../../include/gpusolver.h"

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "dense.hpp"
#include "engine.hpp"
#include "host/setup.h"
#include "setup_device.hpp"

struct gs_csr_s {
  gsx::Csr A;
};
struct gs_hier_s {
  gsx::Hierarchy H;
};
struct gs_solver_s {
  int precision = GS_PREC_DOUBLE;
  gse::Options opt;
  bool coarse_kind_set = false;
  std::unique_ptr<gse::EngineBase> eng;
  gse::SolveResult last;
  // Learned-weight guard: the plain solver restarts the solve when the
  // learned one breaks down (p.Ap <= 0 or r.z <= 0), produces a non-finite
  // residual, or lags the plain rate (checked at every host check of the
  // stop word, and at most lag * the plain iteration count in total).
  gs_solver_s *fallback = nullptr;
  double lag = 1.5;
  bool fell_back = false;
  // Why the last solve fell back (0 none, 1 lagging the plain rate, 2 CG
  // breakdown, 3 non-finite residual, 4 budget spent) and how many PCG
  // iterations the guarded attempt ran.
  int fallback_reason = 0;
  int64_t first_iterations = 0;
  // Convergence rate of this solver's last converged solve: mean log
  // reduction of relres^2 per PCG iteration (0: none yet).
  double log_rate = 0.0;
};

namespace gsc {
thread_local std::string g_error;
void set_error(const char *msg) { g_error = msg; }
}  // namespace gsc
using gsc::g_error;

namespace {

template <class F>
int guard(F &&f) {
  try {
    f();
    return GS_OK;
  } catch (const std::exception &e) {
    g_error = e.what();
    return GS_ERR;
  } catch (...) {
    g_error = "unknown error";
    return GS_ERR;
  }
}

void need(const void *p, const char *what) {
  if (!p) throw std::runtime_error(std::string(what) + " must not be null.");
}

gsx::Csr make_csr(int64_t rows, int64_t cols, int64_t nnz, const int32_t *ptr,
                  const int32_t *col, const double *val) {
  if (rows <= 0 || cols <= 0)
    throw std::runtime_error("HostCSRMatrix: matrix dimensions must be nonzero.");
  if (!ptr || (nnz > 0 && (!col || !val)))
    throw std::runtime_error("HostCSRMatrix: input pointers must not be null.");
  gsx::Csr A;
  A.rows = rows;
  A.cols = cols;
  A.ptr.assign(ptr, ptr + rows + 1);
  A.col.assign(col, col + nnz);
  A.val.assign(val, val + nnz);
  A.check();
  return A;
}

gsx::Perm perm_from_o2n(int64_t n, const int32_t *o2n) {
  gsx::Perm p;
  p.old_to_new.assign(o2n, o2n + n);
  p.new_to_old.assign(n, -1);
  for (int64_t i = 0; i < n; ++i) {
    if (o2n[i] < 0 || o2n[i] >= n)
      throw std::runtime_error("Permutation: oldToNew contains invalid index.");
    p.new_to_old[o2n[i]] = static_cast<int32_t>(i);
  }
  p.check();
  return p;
}
}  // namespace

extern "C" {

const char *gs_last_error(void) { return g_error.c_str(); }
int gs_version(void) { return 1; }

int gs_csr_create(int64_t rows, int64_t cols, int64_t nnz, const int32_t *row_ptr,
                  const int32_t *col_idx, const double *values, gs_csr_h *out) {
  return guard([&] {
    need(out, "out");
    auto h = std::make_unique<gs_csr_s>();
    h->A = make_csr(rows, cols, nnz, row_ptr, col_idx, values);
    *out = h.release();
  });
}

int gs_csr_info(gs_csr_h A, int64_t *rows, int64_t *cols, int64_t *nnz) {
  return guard([&] {
    need(A, "matrix");
    if (rows) *rows = A->A.rows;
    if (cols) *cols = A->A.cols;
    if (nnz) *nnz = A->A.nnz();
  });
}

int gs_csr_export(gs_csr_h A, int32_t *row_ptr, int32_t *col_idx, double *values) {
  return guard([&] {
    need(A, "matrix");
    if (row_ptr) std::memcpy(row_ptr, A->A.ptr.data(), A->A.ptr.size() * sizeof(int32_t));
    if (col_idx) std::memcpy(col_idx, A->A.col.data(), A->A.col.size() * sizeof(int32_t));
    if (values) std::memcpy(values, A->A.val.data(), A->A.val.size() * sizeof(double));
  });
}

int gs_csr_destroy(gs_csr_h A) {
  delete A;
  return GS_OK;
}

int gs_csr_permute_keep_order(gs_csr_h A, const int32_t *o2n, gs_csr_h *out) {
  return guard([&] {
    need(A, "matrix");
    need(o2n, "permutation");
    need(out, "out");
    auto h = std::make_unique<gs_csr_s>();
    h->A = gsx::permute_keep_order(A->A, perm_from_o2n(A->A.rows, o2n));
    *out = h.release();
  });
}

int gs_csr_permute_sorted(gs_csr_h A, const int32_t *ro2n, const int32_t *co2n, gs_csr_h *out) {
  return guard([&] {
    need(A, "matrix");
    need(ro2n, "row permutation");
    need(co2n, "column permutation");
    need(out, "out");
    auto h = std::make_unique<gs_csr_s>();
    h->A = gsx::permute_sorted(A->A, ro2n, co2n);
    *out = h.release();
  });
}

int gs_ell_pack(gs_csr_h A, int32_t width, int32_t *ell_col, double *ell_val, int32_t *ovf_ptr,
                int32_t *ovf_col, double *ovf_val, int64_t *overflow_nnz) {
  return guard([&] {
    need(A, "matrix");
    gsx::Ell E = gsx::ell_pack(A->A, width);
    if (overflow_nnz) *overflow_nnz = E.overflow_nnz;
    if (!ell_col) return;
    std::memcpy(ell_col, E.col.data(), E.col.size() * sizeof(int32_t));
    std::memcpy(ell_val, E.val.data(), E.val.size() * sizeof(double));
    std::memcpy(ovf_ptr, E.ovf_ptr.data(), E.ovf_ptr.size() * sizeof(int32_t));
    if (E.overflow_nnz) {
      std::memcpy(ovf_col, E.ovf_col.data(), E.ovf_col.size() * sizeof(int32_t));
      std::memcpy(ovf_val, E.ovf_val.data(), E.ovf_val.size() * sizeof(double));
    }
  });
}

int gs_csr_sell_layout(gs_csr_h A, int64_t *slots, int32_t *min_col, int32_t *max_col) {
  return guard([&] {
    need(A, "matrix");
    const gsx::Sell64 P = gsx::sell64_pack(A->A);
    if (slots) *slots = P.slots;
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int32_t c : P.col) {
      lo = std::min(lo, c);
      hi = std::max(hi, c);
    }
    if (min_col) *min_col = P.col.empty() ? 0 : lo;
    if (max_col) *max_col = P.col.empty() ? 0 : hi;
  });
}

int gs_graph_coloring(int64_t n, const int32_t *ptr, const int32_t *col, int32_t *colors,
                      int32_t *vertices, int32_t *offsets, int32_t *ncolors) {
  return guard([&] {
    need(ptr, "row pointer");
    gsx::Coloring c = gsx::greedy_coloring(n, ptr, col);
    if (colors) std::memcpy(colors, c.colors.data(), n * sizeof(int32_t));
    if (vertices) std::memcpy(vertices, c.vertices.data(), n * sizeof(int32_t));
    if (offsets) std::memcpy(offsets, c.offsets.data(), c.offsets.size() * sizeof(int32_t));
    if (ncolors) *ncolors = c.ncolors;
  });
}

int gs_rcm(int64_t n, const int32_t *ptr, const int32_t *col, int32_t *o2n, int32_t *n2o) {
  return guard([&] {
    need(ptr, "row pointer");
    gsx::Perm p = gsx::rcm(n, ptr, col);
    std::memcpy(o2n, p.old_to_new.data(), n * sizeof(int32_t));
    std::memcpy(n2o, p.new_to_old.data(), n * sizeof(int32_t));
  });
}

int gs_permutation_check(int64_t n, const int32_t *o2n, const int32_t *n2o) {
  return guard([&] {
    if (n == 0) throw std::runtime_error("Permutation: size must be nonzero.");
    if (!o2n || !n2o) throw std::runtime_error("Permutation: input pointers must not be null.");
    gsx::Perm p;
    p.old_to_new.assign(o2n, o2n + n);
    p.new_to_old.assign(n2o, n2o + n);
    p.check();
  });
}

int gs_block_graph(gs_csr_h A, int32_t bs, int32_t *gptr, int32_t *gcol, int64_t *gnnz) {
  return guard([&] {
    need(A, "matrix");
    gsx::Graph g = gsx::block_graph(A->A, bs);
    if (gnnz) *gnnz = static_cast<int64_t>(g.col.size());
    if (gptr) std::memcpy(gptr, g.ptr.data(), g.ptr.size() * sizeof(int32_t));
    if (gcol) std::memcpy(gcol, g.col.data(), g.col.size() * sizeof(int32_t));
  });
}

int gs_expand_block_permutation(int64_t nodes, int32_t bs, const int32_t *node_o2n, int32_t *o2n,
                                int32_t *n2o) {
  return guard([&] {
    gsx::Perm np = perm_from_o2n(nodes, node_o2n);
    gsx::Perm p = gsx::expand_block_perm(np, nodes, bs);
    std::memcpy(o2n, p.old_to_new.data(), p.old_to_new.size() * sizeof(int32_t));
    std::memcpy(n2o, p.new_to_old.data(), p.new_to_old.size() * sizeof(int32_t));
  });
}

int gs_inverse_diag_blocks3(gs_csr_h A, double *inv) {
  return guard([&] {
    need(A, "matrix");
    std::vector<double> v = gsx::inverse_diag_blocks3(A->A);
    std::memcpy(inv, v.data(), v.size() * sizeof(double));
  });
}

int gs_dense_spd_inverse(int64_t n, double *a) {
  return guard([&] {
    if (n < 0) throw std::runtime_error("dense SPD inverse: negative size.");
    if (n > 0) need(a, "matrix");
    std::vector<double> m(a, a + n * n);
    gsx::spd_inverse(n, m);
    std::memcpy(a, m.data(), m.size() * sizeof(double));
  });
}

int gs_dense_spd_inverse_device(int64_t n, double *a) {
  return guard([&] {
    if (n < 0) throw std::runtime_error("dense SPD inverse: negative size.");
    if (n == 0) return;
    need(a, "matrix");
    std::vector<double> m(a, a + n * n);
    const int64_t N = gsdense::padded_size(n);
    hipStream_t st = nullptr;
    GS_HIP(hipStreamCreate(&st));
    gsd::DevBuf inv(static_cast<size_t>(N) * N * sizeof(double));
    try {
      gsdense::spd_inverse_device(n, m, inv.as<double>(), st);
      GS_HIP(hipMemcpy2DAsync(a, n * sizeof(double), inv.get(), N * sizeof(double),
                              n * sizeof(double), n, hipMemcpyDeviceToHost, st));
      GS_HIP(hipStreamSynchronize(st));
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    GS_HIP(hipStreamDestroy(st));
  });
}

int gs_problem_poisson2d(int32_t nx, int32_t ny, gs_csr_h *out) {
  return guard([&] {
    auto h = std::make_unique<gs_csr_s>();
    h->A = gsx::poisson2d(nx, ny);
    *out = h.release();
  });
}

int gs_problem_poisson3d(int32_t nx, int32_t ny, int32_t nz, gs_csr_h *out) {
  return guard([&] {
    auto h = std::make_unique<gs_csr_s>();
    h->A = gsx::poisson3d(nx, ny, nz);
    *out = h.release();
  });
}

int gs_problem_varcoef3d(int32_t n, uint64_t seed, double sigma, double scale, gs_csr_h *out) {
  return guard([&] {
    auto h = std::make_unique<gs_csr_s>();
    h->A = gsx::varcoef3d(n, seed, sigma, scale, nullptr);
    *out = h.release();
  });
}

int gs_problem_elasticity(int32_t nv, double E, double nu, double hs, double hscale, uint64_t seed,
                          gs_csr_h *out, double *rhs, double *nullspace) {
  return guard([&] {
    gsx::ElasticityProblem P = gsx::elasticity(nv, E, nu, hs, hscale, seed);
    if (rhs) std::memcpy(rhs, P.rhs.data(), P.rhs.size() * sizeof(double));
    if (nullspace) std::memcpy(nullspace, P.nullspace.data(), P.nullspace.size() * sizeof(double));
    auto h = std::make_unique<gs_csr_s>();
    h->A = std::move(P.A);
    *out = h.release();
  });
}

int gs_problem_elasticity_mesh(int64_t nverts, const double *V, int64_t ntets, const int32_t *T,
                               double young, double poisson, const double *young_per_tet,
                               double clamp_rel, double density, double gravity, gs_csr_h *out,
                               double *rhs, double *nullspace) {
  return guard([&] {
    need(V, "vertices");
    need(T, "tetrahedra");
    need(out, "out");
    gsx::ElasticityProblem P = gsx::elasticity_mesh(nverts, V, ntets, T, young, poisson,
                                                    young_per_tet, clamp_rel, density, gravity);
    if (rhs) std::memcpy(rhs, P.rhs.data(), P.rhs.size() * sizeof(double));
    if (nullspace) std::memcpy(nullspace, P.nullspace.data(), P.nullspace.size() * sizeof(double));
    auto h = std::make_unique<gs_csr_s>();
    h->A = std::move(P.A);
    *out = h.release();
  });
}

int gs_uniform_pm1(uint64_t seed, int64_t n, double *out) {
  return guard([&] { gsx::uniform_pm1(seed, n, out); });
}

int gs_gaussian_field(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double sigma,
                      double *out) {
  return guard([&] {
    std::vector<double> g = gsx::gaussian_random_field(nx, ny, nz, seed, sigma);
    std::memcpy(out, g.data(), g.size() * sizeof(double));
  });
}

namespace {
gsx::Hierarchy build_sa(gs_csr_h A, int32_t bs, double eps, int32_t max_levels, int64_t min_dofs,
                        const double *ns, int32_t ncols, const gsx::GalerkinOps *ops) {
  need(A, "matrix");
  gsx::SAParams prm;
  prm.block_size = bs;
  prm.eps_strong = eps;
  prm.max_levels = max_levels;
  prm.min_dofs = min_dofs;
  std::vector<double> B;
  if (ncols > 0) {
    need(ns, "nullspace");
    B.assign(ns, ns + A->A.rows * ncols);
  }
  return gsx::build_sa_hierarchy(A->A, prm, B, ncols, ops);
}
}  // namespace

int gs_hierarchy_build_sa(gs_csr_h A, int32_t bs, double eps, int32_t max_levels, int64_t min_dofs,
                          const double *ns, int32_t ncols, gs_hier_h *out) {
  return guard([&] {
    need(out, "out");
    auto h = std::make_unique<gs_hier_s>();
    h->H = build_sa(A, bs, eps, max_levels, min_dofs, ns, ncols, nullptr);
    *out = h.release();
  });
}

int gs_hierarchy_build_sa_device(gs_csr_h A, int32_t bs, double eps, int32_t max_levels,
                                 int64_t min_dofs, const double *ns, int32_t ncols,
                                 gs_hier_h *out) {
  return guard([&] {
    need(out, "out");
    const gsdev::DeviceGalerkinOps ops;
    auto h = std::make_unique<gs_hier_s>();
    h->H = build_sa(A, bs, eps, max_levels, min_dofs, ns, ncols, &ops);
    *out = h.release();
  });
}

int gs_csr_spgemm(gs_csr_h A, gs_csr_h B, int32_t on_device, gs_csr_h *out) {
  return guard([&] {
    need(A, "matrix");
    need(B, "matrix");
    need(out, "out");
    auto h = std::make_unique<gs_csr_s>();
    h->A = on_device ? gsdev::device_spgemm(A->A, B->A) : gsx::spgemm(A->A, B->A);
    *out = h.release();
  });
}

int gs_csr_transpose(gs_csr_h A, int32_t on_device, gs_csr_h *out) {
  return guard([&] {
    need(A, "matrix");
    need(out, "out");
    auto h = std::make_unique<gs_csr_s>();
    h->A = on_device ? gsdev::device_transpose(A->A) : gsx::transpose(A->A);
    *out = h.release();
  });
}

int gs_csr_save(gs_csr_h A, const char *path) {
  return guard([&] {
    need(A, "matrix");
    need(path, "path");
    gsx::save_csr(A->A, path);
  });
}

int gs_csr_load(const char *path, gs_csr_h *out) {
  return guard([&] {
    need(path, "path");
    need(out, "out");
    auto h = std::make_unique<gs_csr_s>();
    h->A = gsx::load_csr(path);
    *out = h.release();
  });
}

int gs_hierarchy_save(gs_hier_h h, const char *path) {
  return guard([&] {
    need(h, "hierarchy");
    need(path, "path");
    gsx::save_hierarchy(h->H, path);
  });
}

int gs_hierarchy_load(const char *path, gs_hier_h *out) {
  return guard([&] {
    need(path, "path");
    need(out, "out");
    auto h = std::make_unique<gs_hier_s>();
    h->H = gsx::load_hierarchy(path);
    *out = h.release();
  });
}

int gs_hierarchy_create(gs_hier_h *out) {
  return guard([&] { *out = new gs_hier_s(); });
}

int gs_hierarchy_push_level(gs_hier_h h, gs_csr_h A) {
  return guard([&] {
    need(h, "hierarchy");
    need(A, "matrix");
    h->H.A.push_back(A->A);
  });
}

int gs_hierarchy_push_transfer(gs_hier_h h, gs_csr_h P, gs_csr_h R) {
  return guard([&] {
    need(h, "hierarchy");
    need(P, "prolongation");
    need(R, "restriction");
    h->H.P.push_back(P->A);
    h->H.R.push_back(R->A);
  });
}

int gs_hierarchy_levels(gs_hier_h h, int32_t *nlev) {
  return guard([&] {
    need(h, "hierarchy");
    *nlev = static_cast<int32_t>(h->H.A.size());
  });
}

int gs_hierarchy_get(gs_hier_h h, int32_t which, int32_t k, gs_csr_h *out) {
  return guard([&] {
    need(h, "hierarchy");
    const std::vector<gsx::Csr> &v = which == 0 ? h->H.A : which == 1 ? h->H.P : h->H.R;
    if (k < 0 || k >= static_cast<int32_t>(v.size()))
      throw std::runtime_error("hierarchy index out of range.");
    auto c = std::make_unique<gs_csr_s>();
    c->A = v[k];
    *out = c.release();
  });
}

int gs_hierarchy_aggregates(gs_hier_h h, int32_t k, int32_t *agg) {
  return guard([&] {
    need(h, "hierarchy");
    if (k < 0 || k >= static_cast<int32_t>(h->H.aggregates.size()))
      throw std::runtime_error("hierarchy index out of range.");
    std::memcpy(agg, h->H.aggregates[k].data(), h->H.aggregates[k].size() * sizeof(int32_t));
  });
}

int gs_hierarchy_level_rows(gs_hier_h h, int32_t k, int64_t *rows) {
  return guard([&] {
    need(h, "hierarchy");
    need(rows, "rows");
    if (k < 0 || k >= static_cast<int32_t>(h->H.A.size()))
      throw std::runtime_error("hierarchy index out of range.");
    *rows = h->H.A[k].rows;
  });
}

int gs_hierarchical_reordering(gs_hier_h h, int32_t *o2n, int32_t *n2o, int64_t *orphans) {
  return guard([&] {
    need(h, "hierarchy");
    need(o2n, "old_to_new");
    need(n2o, "new_to_old");
    std::vector<int64_t> orph;
    const std::vector<gsx::Perm> perms = gsx::hierarchical_reordering(h->H, &orph);
    int64_t off = 0;
    for (const gsx::Perm &p : perms) {
      const size_t n = p.old_to_new.size();
      std::memcpy(o2n + off, p.old_to_new.data(), n * sizeof(int32_t));
      std::memcpy(n2o + off, p.new_to_old.data(), n * sizeof(int32_t));
      off += static_cast<int64_t>(n);
    }
    if (orphans && !orph.empty()) std::memcpy(orphans, orph.data(), orph.size() * sizeof(int64_t));
  });
}

int gs_hierarchy_features(gs_hier_h h, int32_t block_size, double *features) {
  return guard([&] {
    need(h, "hierarchy");
    need(features, "features");
    std::vector<double> f = gsx::level_features(h->H, block_size);
    std::memcpy(features, f.data(), f.size() * sizeof(double));
  });
}

int gs_hierarchy_destroy(gs_hier_h h) {
  delete h;
  return GS_OK;
}

int gs_solver_create(int32_t precision, gs_solver_h *out) {
  return guard([&] {
    if (precision < GS_PREC_DOUBLE || precision > GS_PREC_SINGLE)
      throw std::runtime_error("unknown precision mode.");
    auto s = std::make_unique<gs_solver_s>();
    s->precision = precision;
    *out = s.release();
  });
}

int gs_solver_set_int(gs_solver_h s, const char *key, int64_t v) {
  return guard([&] {
    need(s, "solver");
    need(key, "key");
    const std::string k(key);
    gse::Options &o = s->opt;
    if (k == "preconditioner") o.precond = static_cast<int>(v);
    else if (k == "smoother") {
      o.fine_kind = static_cast<int>(v);
      if (!s->coarse_kind_set) o.coarse_kind = o.fine_kind;
    } else if (k == "coarse_smoother") {
      o.coarse_kind = static_cast<int>(v);
      s->coarse_kind_set = true;
    } else if (k == "pre") o.pre = static_cast<int>(v);
    else if (k == "post") o.post = static_cast<int>(v);
    else if (k == "coarse_steps") o.coarse_steps = static_cast<int>(v);
    else if (k == "vcycles") o.vcycles = static_cast<int>(v);
    else if (k == "coarse_direct") o.coarse_direct = v != 0;
    else if (k == "graph_iters") o.graph_iters = static_cast<int>(v);
    else if (k == "graph_tail") o.graph_tail = v != 0;
    else if (k == "cg_block3") o.cg_block3 = v != 0;
    else if (k == "reference_vcycle") o.reference_vcycle = v != 0;
    else if (k == "max_direct") o.max_direct = v;
    else if (k == "direct_truncate") o.direct_truncate = v != 0;
    else if (k == "transfer_level_order") o.transfer_level_order = v != 0;
    else if (k == "block_transfers") o.block_transfers = v != 0;
    else if (k == "cg_level_order") o.cg_level_order = v != 0;
    else if (k == "cg_vcycle_operator") o.cg_vcycle_operator = v != 0;
    else if (k == "cg_pair_layout") o.cg_pair_layout = v != 0;
    else if (k == "device_pack") o.device_pack = v != 0;
    else if (k == "persist_grid") {
      if (v < 8 || v % 8 != 0 || v > (int64_t(1) << 20))
        throw std::runtime_error("persist_grid must be a positive multiple of 8.");
      o.persist_grid = static_cast<int>(v);
    }
    else if (k == "cg_replace_every") {
      if (v < 1) throw std::runtime_error("cg_replace_every must be at least 1.");
      o.cg_replace_every = static_cast<int>(v);
    }
    else if (k == "cg_true_residual_every") {
      if (v < 0 || v > (int64_t(1) << 30))
        throw std::runtime_error("cg_true_residual_every must be non-negative.");
      o.cg_true_residual_every = static_cast<int>(v);
    }
    else if (k == "tail_rows") {
      if (v < 0) throw std::runtime_error("tail_rows must be non-negative.");
      o.tail_rows = v;
    }
    else if (k == "zero_start_lower") {
      if (v < 0) throw std::runtime_error("zero_start_lower must be non-negative.");
      o.zero_start_lower = static_cast<int>(v);
    }
    else if (k == "transfer_sigma") {
      if (v < 0) throw std::runtime_error("transfer_sigma must be non-negative.");
      o.transfer_sigma = v;
    }
    else if (k == "group_max_transfer") {
      if (v != 16 && v != 32 && v != 64)
        throw std::runtime_error("group_max_transfer must be 16, 32 or 64.");
      o.group_max_transfer = static_cast<int>(v);
    }
    else throw std::runtime_error("unknown integer option: " + k);
    if (o.pre < 0 || o.post < 0 || o.coarse_steps < 0 || o.vcycles < 0 || o.graph_iters < 0)
      throw std::runtime_error("option values must be non-negative.");
    if (o.fine_kind < 0 || o.fine_kind > 3 || o.coarse_kind < 0 || o.coarse_kind > 3)
      throw std::runtime_error("MultigridSmoother: unsupported smoother type.");
  });
}

int gs_solver_set_real(gs_solver_h s, const char *key, double v) {
  return guard([&] {
    need(s, "solver");
    need(key, "key");
    const std::string k(key);
    if (k == "omega") s->opt.omega = v;
    else if (k == "group_min_scalar") s->opt.group_min_scalar = v;
    else if (k == "group_min_block") s->opt.group_min_block = v;
    else if (k == "full_group_min_scalar") s->opt.full_group_min_scalar = v;
    else if (k == "full_group_min_block") s->opt.full_group_min_block = v;
    else throw std::runtime_error(std::string("unknown real option: ") + key);
  });
}

int gs_solver_set_level_weights(gs_solver_h s, int32_t nlev, const double *omega,
                                const int32_t *sweeps) {
  return guard([&] {
    need(s, "solver");
    s->opt.level_omega.assign(nlev, 0.0);
    s->opt.level_sweeps.assign(nlev, -1);
    for (int32_t k = 0; k < nlev; ++k) {
      if (omega) s->opt.level_omega[k] = omega[k];
      if (sweeps) s->opt.level_sweeps[k] = sweeps[k];
    }
    if (s->eng) s->eng->set_level_weights(s->opt.level_omega, s->opt.level_sweeps);
  });
}

int gs_solver_set_fallback(gs_solver_h s, gs_solver_h plain, double lag) {
  return guard([&] {
    need(s, "solver");
    if (plain == s) throw std::runtime_error("fallback solver must be a different solver.");
    if (plain && !(lag >= 1.0)) throw std::runtime_error("lag factor must be >= 1.");
    s->fallback = plain;
    s->lag = lag;
  });
}

int gs_solver_setup(gs_solver_h s, gs_hier_h h, const int32_t *const *o2n) {
  return guard([&] {
    need(s, "solver");
    need(h, "hierarchy");
    std::vector<gsx::Perm> perms;
    for (size_t k = 0; k < h->H.A.size(); ++k) {
      const int64_t n = h->H.A[k].rows;
      if (o2n && o2n[k]) perms.push_back(perm_from_o2n(n, o2n[k]));
      else perms.push_back(gsx::Perm::identity(n));
    }
    s->eng.reset();
    s->last = gse::SolveResult{};
    s->log_rate = 0.0;
    s->eng = gse::make_engine(s->precision, s->opt);
    s->eng->setup(h->H, perms);
  });
}

namespace {
// Rate of a converged solve from its relres^2 history (first entry: the
// initial residual).
double solve_log_rate(const gse::EngineBase &e, const gse::SolveResult &r) {
  const auto &h = e.history();
  if (!r.converged || r.loop_iterations <= 0 || h.size() < 2 || !(h.front() > 0.0) ||
      !(h.back() > 0.0))
    return 0.0;
  return std::log(h.back() / h.front()) / static_cast<double>(r.loop_iterations);
}
}  // namespace

int gs_solver_solve(gs_solver_h s, int64_t b_len, const double *b, int64_t x_len, double *x,
                    double tol, int64_t maxit, int64_t *iters, double *relres2,
                    double *true_relres, double *solve_ms) {
  return guard([&] {
    need(s, "solver");
    if (!s->eng) throw std::runtime_error("CGSolver: solver is not set up.");
    // src/CGSolver.cc:90-98.
    if (b_len != s->eng->size())
      throw std::runtime_error("CGSolver::solve: RHS size does not match matrix.");
    if (x_len != s->eng->size())
      throw std::runtime_error("CGSolver::solve: initial guess size does not match matrix.");
    need(b, "b");
    need(x, "x");
    if (maxit < 0) throw std::runtime_error("maxIterations must be non-negative.");
    s->fell_back = false;
    s->fallback_reason = 0;
    s->first_iterations = 0;
    s->last = gse::SolveResult{};
    if (s->fallback) {
      gs_solver_s *pl = s->fallback;
      if (!pl->eng) throw std::runtime_error("fallback solver is not set up.");
      if (pl->eng->size() != s->eng->size())
        throw std::runtime_error("fallback solver was set up for another matrix size.");
      if (pl->last.loop_iterations <= 0 || !(pl->log_rate < 0.0))
        throw std::runtime_error(
            "fallback solver has no converged solve to take the plain rate from; solve with it "
            "first.");
      // Budget from the plain solver's last iteration count, and a rate
      // check at every host check of the stop word.
      const int64_t budget = std::min<int64_t>(
          maxit, static_cast<int64_t>(std::ceil(s->lag * pl->last.loop_iterations)));
      gse::LagGuard lg;
      lg.log_rate = pl->log_rate;
      lg.lag = s->lag;
      const std::vector<double> x0(x, x + s->eng->size());
      int reason = 0;
      double first_ms = 0.0;
      try {
        s->last = s->eng->solve(b, x, tol, budget, &lg);
        first_ms = s->last.solve_ms;
        s->first_iterations = s->last.loop_iterations;
        if (!s->last.zero_rhs) {
          if (s->last.breakdown != 0) reason = 2;
          else if (s->last.lagging) reason = 1;
          else if (!s->last.converged) reason = 4;
        }
      } catch (const gse::NonFiniteResidual &) {
        reason = 3;  // non-finite residual in the learned solve
      }
      if (reason != 0) {
        s->fallback_reason = reason;
        std::copy(x0.begin(), x0.end(), x);
        s->last = pl->eng->solve(b, x, tol, maxit);
        s->last.solve_ms += first_ms;
        s->fell_back = true;
      }
    } else {
      s->last = s->eng->solve(b, x, tol, maxit);
      const double r = solve_log_rate(*s->eng, s->last);
      if (r < 0.0) s->log_rate = r;
    }
    if (iters) *iters = s->last.iterations;
    if (relres2) *relres2 = s->last.relres2;
    if (true_relres) *true_relres = s->last.true_relres;
    if (solve_ms) *solve_ms = s->last.solve_ms;
  });
}

int gs_solver_result(gs_solver_h s, const char *key, double *value) {
  return guard([&] {
    need(s, "solver");
    const std::string k(key);
    const gse::SolveResult &r = s->last;
    if (k == "converged") *value = r.converged ? 1.0 : 0.0;
    else if (k == "loop_iterations") *value = static_cast<double>(r.loop_iterations);
    else if (k == "launched_iterations") *value = static_cast<double>(r.launched_iterations);
    else if (k == "breakdown") *value = r.breakdown;
    else if (k == "wall_ms") *value = r.wall_ms;
    else if (k == "zero_rhs") *value = r.zero_rhs ? 1.0 : 0.0;
    else if (k == "fell_back") *value = s->fell_back ? 1.0 : 0.0;
    else if (k == "lagging") *value = r.lagging ? 1.0 : 0.0;
    else if (k == "fallback_reason") *value = s->fallback_reason;
    else if (k == "first_iterations") *value = static_cast<double>(s->first_iterations);
    else if (k == "log_rate") *value = s->log_rate;
    else throw std::runtime_error("unknown result key: " + k);
  });
}

int gs_solver_history(gs_solver_h s, double *out, int64_t cap, int64_t *len) {
  return guard([&] {
    need(s, "solver");
    if (!s->eng) throw std::runtime_error("solver is not set up.");
    const auto &h = s->eng->history();
    const int64_t n = std::min<int64_t>(cap, static_cast<int64_t>(h.size()));
    if (out) std::memcpy(out, h.data(), n * sizeof(double));
    if (len) *len = static_cast<int64_t>(h.size());
  });
}

int gs_solver_vcycle(gs_solver_h s, int64_t n, const double *r, double *z) {
  return guard([&] {
    need(s, "solver");
    if (!s->eng) throw std::runtime_error("solver is not set up.");
    if (n != s->eng->size())
      throw std::runtime_error("MultigridSmoother: vector size does not match the finest level.");
    need(r, "r");
    need(z, "z");
    s->eng->vcycle_level_order(r, z);
  });
}

int gs_solver_spmv(gs_solver_h s, int64_t n, const double *x, double *y) {
  return guard([&] {
    need(s, "solver");
    if (!s->eng) throw std::runtime_error("solver is not set up.");
    // src/CPU/CPUBackend.cc:220-222.
    if (n != s->eng->size()) throw std::runtime_error("CPUBackend::encodeSpmv: size mismatch.");
    need(x, "x");
    need(y, "y");
    s->eng->spmv_level_order(x, y);
  });
}

int gs_solver_size(gs_solver_h s, int64_t *n) {
  return guard([&] {
    need(s, "solver");
    need(n, "n");
    if (!s->eng) throw std::runtime_error("solver is not set up.");
    *n = s->eng->size();
  });
}

int gs_solver_time_kernel(gs_solver_h s, int32_t which, int32_t reps, double *ms, double *bytes) {
  return guard([&] {
    need(s, "solver");
    if (!s->eng) throw std::runtime_error("solver is not set up.");
    if (reps <= 0) throw std::runtime_error("reps must be positive.");
    double b = 0.0;
    *ms = s->eng->time_kernel(which, reps, &b);
    if (bytes) *bytes = b;
  });
}

int gs_solver_describe(gs_solver_h s, char *buf, int64_t cap) {
  return guard([&] {
    need(s, "solver");
    if (!s->eng) throw std::runtime_error("solver is not set up.");
    const std::string d = s->eng->describe();
    if (cap > 0) {
      const size_t m = std::min<size_t>(cap - 1, d.size());
      std::memcpy(buf, d.data(), m);
      buf[m] = '\0';
    }
  });
}

int gs_solver_destroy(gs_solver_h s) {
  delete s;
  return GS_OK;
}

}  // extern "C"
