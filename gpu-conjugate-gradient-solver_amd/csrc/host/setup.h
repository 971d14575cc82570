// Host-side setup for the AMG-PCG solve phase: CSR containers, the two
// permutation flavours, ELL / BSR3-ELL packing, greedy multicolouring, RCM,
// block utilities, smoothed-aggregation coarsening and the problem
// generators that stand in for the reference's meshes.
//
// Everything here runs once per matrix on the CPU. The per-iteration work
// lives in the HIP sources next to this directory.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

namespace gsx {

// Vector whose elements are default-initialised (left uninitialised for
// arithmetic types): large packing buffers are written in parallel by their
// packers, so a serial zero fill beforehand would only cost time.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <class U>
  void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void *>(p)) U;
  }
  template <class U, class... Args>
  void construct(U *p, Args &&...args) {
    ::new (static_cast<void *>(p)) U(std::forward<Args>(args)...);
  }
};
template <class T>
using RawVec = std::vector<T, NoInitAlloc<T>>;

// Compressed sparse row matrix with int32 indices (the reference's index
// type, include/GPUSolver/HostSparseMatrix.h:285-291) and FP64 values.
// Values are held in FP64 on the host; FP32 device copies are rounded from
// them once at upload.
struct Csr {
  int64_t rows = 0;
  int64_t cols = 0;
  std::vector<int32_t> ptr;  // rows + 1
  RawVec<int32_t> col;      // nnz (default-initialised)
  RawVec<double> val;       // nnz
  int64_t nnz() const { return static_cast<int64_t>(col.size()); }
  void check() const;        // structural validation, throws on error
};

// ---------------------------------------------------------------------------
// Permutations (src/Permutation.cc:8-35).
// ---------------------------------------------------------------------------
struct Perm {
  std::vector<int32_t> old_to_new;
  std::vector<int32_t> new_to_old;
  static Perm identity(int64_t n);
  static Perm from_new_to_old(const std::vector<int32_t> &n2o);
  void check() const;  // bijectivity, throws like the reference ctor
};

// A' = P A P^T keeping every row's entry order (columns relabelled only),
// as HostCSRMatrix::applyPermutation (src/HostSparseMatrix.cc:118-249).
Csr permute_keep_order(const Csr &A, const Perm &p);

// A' with rows relabelled by rowOldToNew and columns by colOldToNew, columns
// sorted ascending within each row and duplicates summed, as permuteMatrix
// (src/SparseMatrixUtils.cc:8-31) through Eigen::setFromTriplets.
Csr permute_sorted(const Csr &A, const int32_t *row_old_to_new,
                   const int32_t *col_old_to_new);

// Keep-order relabelling: row i of the result is row rq[i] of A, column c
// becomes cmap[c]; entry order inside each row is unchanged.
Csr relabel(const Csr &A, const int32_t *rq, const int32_t *cmap);
// Row p of the result is row rows[p] of A (entries unchanged), p < nrows.
Csr take_rows(const Csr &A, const int32_t *rows, int64_t nrows);
// Throws std::runtime_error(what) when a column index lies outside
// [0, A.cols).
void check_columns(const Csr &A, const char *what);
Csr transpose(const Csr &A);                // sorted output
Csr spgemm(const Csr &A, const Csr &B);     // sorted output, OpenMP
void sort_rows(Csr &A);                     // sort + sum duplicates
// True when every row's columns are strictly increasing (sort_rows would
// leave A unchanged).
bool rows_sorted_unique(const Csr &A);

// ---------------------------------------------------------------------------
// ELL packing (src/HostELLMatrix.cc:10-206): slot-major
// index = slot * rows + row, padding (col 0, value 0), rows longer than the
// width spill into an overflow CSR in row order.
// ---------------------------------------------------------------------------
struct Ell {
  int64_t rows = 0, cols = 0, nnz = 0;
  int32_t width = 0;
  int64_t overflow_nnz = 0;
  std::vector<int32_t> col;   // width * rows
  std::vector<double> val;    // width * rows
  std::vector<int32_t> ovf_ptr;  // rows + 1
  std::vector<int32_t> ovf_col;
  std::vector<double> ovf_val;
};
Ell ell_pack(const Csr &A, int32_t width);
int32_t max_row_length(const Csr &A);

// Row-contiguous 3x3 block CSR: block k of block row b
// (ptr[b] <= k < ptr[b + 1]) has block column bcol[k] and values
// val[9 k + 3 r + c]. A block row stores the union of the block columns of
// its three scalar rows; blocks missing a scalar entry hold exact zeros.
struct Bsr3Csr {
  int64_t nb = 0;            // block rows
  int64_t ncb = 0;           // block columns
  int64_t nblocks = 0;       // stored blocks
  int32_t width = 0;         // max blocks per block row
  std::vector<int64_t> ptr;  // nb + 1
  std::vector<int32_t> bcol;
  RawVec<double> val;        // 9 * nblocks
};
// Returns false (and leaves out untouched) unless A's rows and columns are
// multiples of 3 (rectangular operators, the transfers, are allowed).
// Blocks of a block row are stored in increasing key[block column] (block
// column itself when key is null); duplicate entries are summed in their
// CSR order.
bool bsr3_pack(const Csr &A, Bsr3Csr &out, const int32_t *key = nullptr);
// Block row p of the result is block row rows[p] of B (blocks unchanged).
Bsr3Csr bsr3_take_rows(const Bsr3Csr &B, const int32_t *rows, int64_t nrows);
// Stable order of block rows by decreasing block count within windows of
// `sigma` rows (SELL-C-sigma): position p holds block row result[p].
std::vector<int32_t> bsr3_sorted_order(const Bsr3Csr &B, int64_t sigma);
// Writes one byte of every 4 KiB page of [p, p + bytes) from all threads:
// arrays allocated without initialisation map their pages in parallel
// before a copy into them (which would take the page faults one by one).
void touch_pages(void *p, size_t bytes);
// dst[i] = (float)src[i], in parallel.
void to_float(const double *src, size_t n, float *dst);
// Parallel permuted copies with a cast: gather dst[i] = src[idx[i]],
// scatter dst[idx[i]] = src[i], for i < n.
void gather_cast(const double *src, const int32_t *idx, int64_t n, double *dst);
void gather_cast(const double *src, const int32_t *idx, int64_t n, float *dst);
void scatter_cast(const double *src, const int32_t *idx, int64_t n, double *dst);
void scatter_cast(const float *src, const int32_t *idx, int64_t n, double *dst);
// sum of v[i]^2 (parallel; the order of the partial sums is fixed for a
// given thread count).
double sum_squares(const double *v, int64_t n);

// SELL-64: the device layout of every scalar operator. Rows are cut into
// slices of 64; slice s stores its rows slot-major with width slice_w[s]
// starting at slice_ptr[s]; entry j of row r sits at
// slice_ptr[r / 64] + 64 * j + r % 64. Each row keeps its CSR order. Padding
// slots hold value 0 and a column inside [0, cols) that, for square
// operators, is never the row itself (a Gauss-Seidel row kernel takes the
// entry on column == row as the diagonal). sell64_pack validates all of
// this and throws otherwise.
struct Sell64 {
  int64_t rows = 0, cols = 0, nnz = 0, slots = 0;
  std::vector<int64_t> slice_ptr;  // slices + 1
  std::vector<int32_t> slice_w;    // slices
  std::vector<int32_t> col;        // slots
  std::vector<double> val;         // slots
  // Stored position -> row of A (SELL-C-sigma); empty: position == row.
  std::vector<int32_t> row_map;
};
Sell64 sell64_pack(const Csr &A);
// SELL-C-sigma for non-square operators (transfers): rows are stably
// sorted by decreasing length inside windows of `sigma` rows before the
// SELL-64 packing, so a slice holds rows of similar length and little
// padding; row_map gives each stored position's row. Entry order inside a
// row is unchanged. Throws for square operators.
Sell64 sell64_pack_sorted(const Csr &A, int64_t sigma);

// BSR3-SELL-64: the device layout of 3x3-block operators. Block rows in
// slices of 64; block slot j of block row b: bcol at slice_ptr[s] + 64 j +
// lane, value e (= 3 r + c) at (slice_ptr[s] + 64 j) * 9 + 64 e + lane.
// Padding blocks are zero with bcol = own block row.
struct Bsr3Sell64 {
  int64_t nb = 0, nblocks = 0, slots = 0;
  std::vector<int64_t> slice_ptr;
  std::vector<int32_t> slice_w;
  std::vector<int32_t> bcol;
  RawVec<double> val;
};
Bsr3Sell64 bsr3_sell64_pack(const Bsr3Csr &B);
// Regroups each 64-slot group's nine value planes into four planes of value
// pairs (lane l holds values 2q and 2q + 1 of its block side by side, 16
// bytes in FP64) followed by the ninth plane: offsets 128 q + 2 l + (e & 1)
// for e = 2q, 2q + 1, and 512 + l for e = 8.
void bsr3_pair_planes(Bsr3Sell64 &P);

// ---------------------------------------------------------------------------
// Colouring and orderings.
// ---------------------------------------------------------------------------
// Sequential greedy first-fit colouring in vertex order plus the counting
// sort into colour groups (src/GraphColoring.cc:9-72).
struct Coloring {
  int32_t ncolors = 0;
  std::vector<int32_t> colors;    // n
  std::vector<int32_t> vertices;  // n, grouped by colour, ascending inside
  std::vector<int32_t> offsets;   // ncolors + 1
};
Coloring greedy_coloring(int64_t n, const int32_t *ptr, const int32_t *col);

// Reverse Cuthill-McKee with the reference's start-vertex rule, degree
// definition and std::sort comparator
// (src/ReorderingStrategies/RCMReordering.cc:11-110).
Perm rcm(int64_t n, const int32_t *ptr, const int32_t *col);

// Block graph of a scalar matrix (src/BlockUtils.cc:126-216).
struct Graph {
  std::vector<int32_t> ptr, col;
};
Graph block_graph(const Csr &A, int32_t bs);
// Node permutation expanded to DOFs (src/BlockUtils.cc:7-55).
Perm expand_block_perm(const Perm &node_perm, int64_t nodes, int32_t bs);

// Inverse of every 3x3 diagonal block through a Cholesky factorisation in
// FP64; throws when a block is not SPD (src/BlockUtils.cc:218-277).
std::vector<double> inverse_diag_blocks3(const Csr &A);
// Inverse diagonal, throws on zero / missing diagonal
// (src/Smoothers/DampedJacobiSmoother.cc:14-56).
std::vector<double> inverse_diagonal(const Csr &A);

// ---------------------------------------------------------------------------
// Smoothed aggregation (amgcl smoothed_aggregation as called from
// examples/amgSolver/src/SmoothedAggregationCoarsener.cc:64-108).
// ---------------------------------------------------------------------------
struct SAParams {
  int32_t block_size = 1;
  double eps_strong = 0.1;
  double relax = 1.0;          // prolongation smoother weight factor (x 2/3)
  int32_t max_levels = 10;
  int64_t min_dofs = 1000;
};
struct Hierarchy {
  std::vector<Csr> A;  // levels
  std::vector<Csr> P;  // levels - 1, P[k]: level k+1 -> level k
  std::vector<Csr> R;  // levels - 1
  std::vector<std::vector<int32_t>> aggregates;  // per transfer, fine DOF -> aggregate (or -1)
};
// nullspace: n x ncols row-major near-nullspace (may be empty).
// The sparse products of one coarsening step: P = M P_tent (M the
// prolongation smoother as a matrix), R = P^T and A_c = R (A P). The base
// class runs them on the host (spgemm, transpose); a GPU implementation
// (gsdev::DeviceGalerkinOps) overrides run() with the same results bit for
// bit.
struct GalerkinOps {
  virtual ~GalerkinOps() = default;
  virtual void run(const Csr &A, const Csr &M, const Csr &Pt, Csr &P, Csr &R, Csr &Ac) const;
};

// ops: where the products of each coarsening step run (null: the host).
Hierarchy build_sa_hierarchy(const Csr &A, const SAParams &prm,
                             const std::vector<double> &nullspace,
                             int32_t ncols, const GalerkinOps *ops = nullptr);

// Binary files (layout in io.cc): one CSR matrix, or a whole hierarchy
// (A, P, R and the aggregate maps). Loading validates every record.
void save_csr(const Csr &A, const std::string &path);
Csr load_csr(const std::string &path);
void save_hierarchy(const Hierarchy &H, const std::string &path);
Hierarchy load_hierarchy(const std::string &path);

// HierarchicalReordering::compute
// (src/ReorderingStrategies/HierarchicalReordering.cc:217-278): one
// permutation per level, finest first. orphans (may be null) receives, per
// transfer, the number of fine DOFs that R gives no parent.
std::vector<Perm> hierarchical_reordering(const Hierarchy &H, std::vector<int64_t> *orphans);

// Per-level features for the learned smoother weights, 5 per level:
// mean row nnz, mean diagonal-dominance ratio |a_ii| / sum_{j!=i} |a_ij|
// (capped at 10), coarsening ratio n_k / n_{k-1} (1 on level 0),
// coefficient contrast log10(max |a_ii| / min |a_ii|), and the colour count
// of the level's smoother colouring (block graph when block_size is 3).
std::vector<double> level_features(const Hierarchy &H, int32_t block_size);

// ---------------------------------------------------------------------------
// Problem generators (counter-based SplitMix64 streams, reproducible in
// numpy with uint64 arithmetic).
// ---------------------------------------------------------------------------
uint64_t splitmix_at(uint64_t seed, uint64_t i);
// U[-1,1) stream: 2 * ((splitmix_at(seed, i) >> 11) * 2^-53) - 1.
void uniform_pm1(uint64_t seed, int64_t n, double *out);
// Smoothed Gaussian random field on an nx*ny*nz cell grid: white noise by
// Box-Muller, separable Gaussian filter of width sigma (cells, truncated at
// 4 sigma, zero padded), renormalised to zero mean and unit variance.
std::vector<double> gaussian_random_field(int nx, int ny, int nz,
                                          uint64_t seed, double sigma);

Csr poisson2d(int nx, int ny);          // 5-point (4,-1), Dirichlet eliminated
Csr poisson3d(int nx, int ny, int nz);  // 7-point (6,-1)
// -div(k grad u) on n^3 cells, k = exp(scale * g), harmonic face means,
// Dirichlet faces with coefficient 2 k.
Csr varcoef3d(int n, uint64_t seed, double sigma, double scale,
              std::vector<double> *kfield);

struct ElasticityProblem {
  Csr A;
  std::vector<double> rhs;        // 3 * free vertices
  std::vector<double> nullspace;  // (3 * free) x 6 row-major rigid body modes
  int64_t free_vertices = 0;
  int64_t tets = 0;
};
// Unit cube with nv^3 vertices, 6 Kuhn tets per hex (split along the
// (1,1,1) diagonal), linear elasticity element matrices as
// examples/elasticity/LinearElasticity.ipp:239-266, duplicates summed
// (:346-348), the x = 0 face clamped and eliminated. hetero_sigma > 0 draws
// E = E0 * exp(hetero_scale * g) per hex from a Gaussian random field.
ElasticityProblem elasticity(int nv, double E0, double nu, double hetero_sigma,
                             double hetero_scale, uint64_t seed);
// The reference's cantilever driver on a caller's tet mesh (V: nverts x 3,
// T: ntets x 4, both row-major, 0-based): clamp x <= x_min + clamp_rel *
// length, element matrices with E (or young_per_tet[t]) and nu, duplicates
// summed in element order, rhs = density * lumped mass * gravity in z.
ElasticityProblem elasticity_mesh(int64_t nverts, const double *V, int64_t ntets,
                                  const int32_t *T, double young, double nu,
                                  const double *young_per_tet, double clamp_rel,
                                  double density, double gravity);

// Dense SPD inverse through Cholesky in FP64 (coarsest-level direct solve).
// a: n x n row-major, overwritten with the inverse. Throws if not SPD.
void spd_inverse(int64_t n, std::vector<double> &a);

}  // namespace gsx
