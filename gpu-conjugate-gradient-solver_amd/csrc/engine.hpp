// Solve-engine interface shared by the engine and the C ABI.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.hpp"
#include "host/setup.h"

namespace gse {

// Smoother kinds in the order of MultigridSmoother::SmootherType
// (include/GPUSolver/Smoothers/MultigridSmoother.h:18-23).
enum { kDampedJacobi = 0, kSGS = 1, kBlockJacobi = 2, kBlockGS = 3 };
enum { kPrecDouble = 0, kPrecMixed = 1, kPrecSingle = 2 };
enum { kPrecIdentity = 0, kPrecMultigrid = 1 };

struct Options {
  int fine_kind = kSGS;
  int coarse_kind = kSGS;
  int pre = 1, post = 1, coarse_steps = 2, vcycles = 1;
  double omega = 1.0;
  bool coarse_direct = false;
  int precond = kPrecMultigrid;
  int graph_iters = 8;
  // Near convergence, replay a one-iteration graph as many times as the
  // last chunk's rate says are left, instead of a whole chunk whose
  // iterations past convergence are no-op launches.
  bool graph_tail = true;
  bool cg_block3 = false;  // store the CG operator as 3x3 blocks when possible
  // Reproduce the reference's coarse-level correction aliasing
  // (solution = 2 P e below the finest level); off: x += P e.
  bool reference_vcycle = false;
  // Largest level accepted for the dense FP64 inverse (built on the device).
  int64_t max_direct = 16384;
  // With coarse_direct: solve directly at the first level of at most
  // max_direct unknowns and leave the levels below it unused.
  bool direct_truncate = false;
  // Operators whose average row length reaches these (scalar entries per
  // row; 3x3 blocks per block row) run on lane-group kernels (same
  // arithmetic, same order); negative: never.
  double group_min_scalar = 24.0;
  double group_min_block = 20.0;
  // The same thresholds for operators read only by whole-level kernels:
  // transfers and the operators of Jacobi-smoothed levels. A Gauss-Seidel
  // level's colour launches cover a fraction of its rows each, so lane
  // groups pay off there at shorter rows; a whole-level launch of 25-40
  // entry scalar rows runs faster one thread per row (measured: the
  // hybrid 3D Poisson solve 5.7 -> 4.6 ms, 32..64 all equal).
  double full_group_min_scalar = 40.0;
  double full_group_min_block = 20.0;
  // Largest lane group for the transfer operators' rows (16, 32 or 64);
  // 32 measured best for the ~190-entry rows of R = P^T (two rows per
  // wavefront, two load rounds per row).
  int group_max_transfer = 32;
  // Thread-per-row transfer operators: rows sorted by length within
  // windows of this many rows (SELL-C-sigma), cutting slice padding; 0 keeps
  // the natural row order.
  int64_t transfer_sigma = 1024;
  // Execute the transfer operators' rows in the level's own (pre-colouring)
  // order instead of the colour order, so that rows running together are
  // spatial neighbours and share their gathers in cache; outputs are
  // written through the row map, sums unchanged.
  bool transfer_level_order = true;
  // Transfers between two levels whose orderings keep node triples (block
  // smoothers) are stored as 3x3 blocks, one column index per 9 values;
  // the sums are those of the scalar rows.
  bool block_transfers = true;
  // The 3x3-block CG operator's rows run in level-0 order (row map).
  bool cg_level_order = true;
  // Gauss-Seidel levels below this index keep a second operator with the
  // earlier-colour entries only, read by the forward colours of a sweep
  // from x = 0 (0: none).
  int zero_start_lower = 1;
  // Scalar Gauss-Seidel levels below the finest with at most this many
  // unknowns, down to the coarsest, run their part of the V-cycle as one
  // single-workgroup kernel (same results); 0: one launch per colour.
  int64_t tail_rows = 4096;
  // Mixed precision: the CG SpMV reads level 0's V-cycle-type operator
  // (FP32 values, products and sums in FP64) instead of the FP64 copy; the
  // FP64 operator recomputes r = b - A x every cg_replace_every iterations
  // and before every stop on convergence.
  bool cg_vcycle_operator = false;
  // The FP64 3x3-block CG operator's values as planes of value pairs
  // (16-byte loads) and its SpMV as a persistent launch walking slices.
  bool cg_pair_layout = true;
  // Workgroups of the persistent CG SpMV (a multiple of 8: one share per
  // XCD); each wavefront walks one SELL slice per pass.
  int persist_grid = static_cast<int>(gsk::kPersistGrid);
  int cg_replace_every = 8;
  // Mixed precision with the FP64 CG operator: the true residual b - A x is
  // recomputed in FP64 every cg_true_residual_every iterations and before
  // every stop on convergence, and replaces the recursive residual
  // (0: only the final true residual after the solve).
  int cg_true_residual_every = 32;
  // Block-smoother levels (and a 3x3-block CG operator) are relabelled and
  // packed on the GPU from one upload of the level matrix, with the same
  // layouts as the host packers; off: packed on the host.
  bool device_pack = true;
  std::vector<double> level_omega;  // per-level overrides (> 0 applies)
  std::vector<int> level_sweeps;    // per-level pre = post sweeps (>= 0 applies)
};

// A non-finite residual (NaN / Inf) in the PCG loop: CGSolver::solve's
// "initial residual is NaN/Inf" / "solver diverged" errors
// (src/CGSolver.cc:182,255), as a type of its own so that a caller
// can tell it from a device or setup failure.
struct NonFiniteResidual : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Convergence-rate guard of a solve: at each host check of the stop word
// after min_iter iterations, the solve ends (result "lagging") when its
// residual reduction log(rel_k / rel_0) is above log_rate * k / lag, i.e.
// it needs more than lag times the iterations of a solver whose relres^2
// falls by exp(log_rate) per iteration.
struct LagGuard {
  double log_rate = 0.0;  // < 0: mean log relres^2 reduction per iteration
  double lag = 1.5;
  int64_t min_iter = 8;
};

struct SolveResult {
  int64_t iterations = 0;       // reference convention (src/CGSolver.cc:257-268)
  int64_t loop_iterations = 0;  // PCG updates actually applied
  int64_t launched_iterations = 0;  // iterations launched (graph iterations past convergence are no-op launches)
  double relres2 = 0.0;         // final recursive ||r||^2 / ||b||^2
  double true_relres = 0.0;     // ||b - A x|| / ||b|| recomputed in FP64
  double solve_ms = 0.0;        // HIP-event time of the solve on the device
  double wall_ms = 0.0;         // host wall time of the same region
  bool converged = false;
  bool zero_rhs = false;
  int breakdown = 0;            // 1: p.Ap <= 0 seen, 2: r.z <= 0 seen
  bool lagging = false;         // ended by a LagGuard
};

struct LevelInfo {
  int64_t n = 0, nnz = 0, blocks3 = 0;
  int ncolors = 0, kind = 0;
  int64_t max_color_rows = 0;  // rows (block rows) of the largest colour
  bool direct = false;
  double omega = 1.0;
  int pre = 1, post = 1;
  int group = 0;    // lanes per row of the level operator (0: one thread)
  int r_group = 0;  // same for the restriction to the next level
  int p_group = 0;  // and the prolongation from it
  bool transfer_b3 = false;  // the transfers to the next level are 3x3 blocks
};

class EngineBase {
 public:
  virtual ~EngineBase() = default;
  virtual void setup(const gsx::Hierarchy &H, const std::vector<gsx::Perm> &perms) = 0;
  // guard (may be null): ends the solve early when it lags a given rate.
  virtual SolveResult solve(const double *b, double *x, double tol, int64_t maxit,
                            const LagGuard *guard = nullptr) = 0;
  virtual void vcycle_level_order(const double *r, double *z) = 0;
  virtual void spmv_level_order(const double *x, double *y) = 0;
  virtual double time_kernel(int which, int reps, double *bytes) = 0;
  virtual std::string describe() const = 0;
  virtual int64_t size() const = 0;
  virtual const std::vector<double> &history() const = 0;
  virtual void level_info(std::vector<LevelInfo> &out) const = 0;
  // Per-level damping / pre = post sweeps after setup (entries <= 0 keep
  // the current value); every entry is validated before any is applied, and
  // the captured graph is rebuilt on the next solve.
  virtual void set_level_weights(const std::vector<double> &omega,
                                 const std::vector<int> &sweeps) = 0;
};

std::unique_ptr<EngineBase> make_engine(int precision, const Options &o);

}  // namespace gse
