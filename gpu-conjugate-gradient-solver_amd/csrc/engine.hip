// AMG-PCG solve engine: the reference's CGSolver::solve
// (src/CGSolver.cc:87-306) driving MultigridSmoother::vCycle
// (src/Smoothers/MultigridSmoother.cc:337-420), rebuilt for gfx950.
//
// Setup (host, once): every level matrix is permuted as the reference does
// (sorted flavour for V-cycle operators, order-keeping flavour for the CG
// operator), coloured with the reference's greedy colouring, and then laid
// out on the device in colour-permuted ("internal") order, so each colour
// of a Gauss-Seidel sweep is a contiguous range of rows and every launch
// streams unit-stride SELL-64 / BSR3-SELL-64 data. The relabelling keeps
// each row's entry order, so every row sum is the reference's.
//
// Solve (device): the PCG loop runs as fused kernels (SpMV + p.Ap with
// alpha, x/r update + r.r with the stop test, V-cycle, r.z with beta,
// p update), each iteration recorded into a hipGraph that is replayed
// `graph_iters` times between host checks of the stop word; iterations
// after convergence are no-ops. Precision is a template pair: CG type TC
// and V-cycle type TM (double/double plain, double/float mixed,
// float/float as-reference).
#include "dense.hpp"
#include "engine.hpp"
#include "kernels_tail.hpp"
#include "setup_device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>

namespace gse {

using gsd::DBsr3;
using gsd::DevBuf;
using gsd::DMat;
using gsd::DOp;
using gsd::DSell;
using gsk::grid_for;
using gsk::kBlock;

namespace {

// Keep-order relabelling (row i <- row rq[i], column c -> cinv[c]).
gsx::Csr relabel(const gsx::Csr &A, const std::vector<int32_t> &rq,
                 const std::vector<int32_t> &cinv) {
  return gsx::relabel(A, rq.data(), cinv.data());
}

std::vector<int32_t> inverse_of(const std::vector<int32_t> &q) {
  std::vector<int32_t> inv(q.size());
  for (size_t i = 0; i < q.size(); ++i) inv[q[i]] = static_cast<int32_t>(i);
  return inv;
}

// True when the ordering q moves unknowns in node triples: q[3j + l] =
// q[3j] + l with q[3j] a multiple of 3 (block smoothers' orders, identity).
bool keeps_triples(const std::vector<int32_t> &q) {
  if (q.size() % 3 != 0) return false;
  for (size_t j = 0; j < q.size(); j += 3)
    if (q[j] % 3 != 0 || q[j + 1] != q[j] + 1 || q[j + 2] != q[j] + 2) return false;
  return true;
}

std::vector<int32_t> iota_vec(int64_t n) {
  std::vector<int32_t> v(n);
  std::iota(v.begin(), v.end(), 0);
  return v;
}

template <class T>
void launch_fill(T *p, T v, int64_t n, hipStream_t s, const int32_t *stop) {
  if (n <= 0) return;
  gsk::k_fill<T><<<grid_for(n), kBlock, 0, s>>>(p, v, n, stop);
}

// Calls f with the transfer rows' entries per lane and round (2, 4, 8) as a
// compile-time constant.
template <class F>
void with_k(int K, F &&f) {
  switch (K) {
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: throw std::runtime_error("entries per lane must be 2, 4 or 8.");
  }
}

// Calls f with the 3x3-block transfer rows' blocks per lane and round (2, 3,
// 4) as a compile-time constant.
template <class F>
void with_bk(int K, F &&f) {
  switch (K) {
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    default: throw std::runtime_error("blocks per lane must be 2, 3 or 4.");
  }
}

// Calls f with the lane-group size G as a compile-time constant.
template <class F>
void with_group(int G, F &&f) {
  switch (G) {
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    default: throw std::runtime_error("lane-group size must be 8, 16, 32 or 64.");
  }
}

// 3x3-block operator from a block CSR packed on the device: the fields and
// storage DOp::build gives the same blocks on the host (lane groups from
// group_min blocks per block row, else BSR3-SELL-64 with rows in `order`).
template <class T>
void build_b3_device(DOp<T> &op, const gsdev::DBsr3Csr &B, double group_min,
                     const int32_t *order) {
  op.n = 3 * B.nb;
  op.b3 = true;
  op.nblocks = B.nblocks;
  op.nnz = 9 * B.nblocks;
  op.bptr = B.host_ptr;
  op.G = gsd::pick_block_group(B.nb ? static_cast<double>(B.nblocks) / B.nb : 0.0, group_min);
  if (op.G) gsdev::bsr3_rows<T>(B, op.bw);
  else gsdev::bsr3_sell_pack<T>(B, order, op.b);
}

bool is_identity(const std::vector<int32_t> &o2n) {
  for (size_t i = 0; i < o2n.size(); ++i)
    if (o2n[i] != static_cast<int32_t>(i)) return false;
  return true;
}

// permute_sorted(A, rows, cols), or A itself when both permutations are
// the identity and A's rows are already sorted and unique (the permuted
// copy would equal A).
const gsx::Csr &permuted(const gsx::Csr &A, const gsx::Perm &rows, const gsx::Perm &cols,
                         gsx::Csr &store) {
  if (is_identity(rows.old_to_new) && is_identity(cols.old_to_new) &&
      gsx::rows_sorted_unique(A))
    return A;
  store = gsx::permute_sorted(A, rows.old_to_new.data(), cols.old_to_new.data());
  return store;
}

}  // namespace

// ---------------------------------------------------------------------------
// Level and engine types.
// ---------------------------------------------------------------------------
template <class TM>
struct Level {
  int64_t n = 0;
  int kind = kSGS;
  bool direct = false;  // coarsest level solved with the dense inverse
  TM omega = TM(1);
  int pre = 1, post = 1;
  DOp<TM> A;
  // Gauss-Seidel levels: the entries of earlier colours only (column below
  // the row's colour range), for the forward colours of a sweep that
  // starts from x = 0, whose later-colour neighbours are still exactly 0.
  DOp<TM> Alow;
  bool has_low = false;
  std::vector<int64_t> coff;  // colour offsets (rows, or blocks for block GS)
  std::vector<int64_t> cunits;  // stored entries (blocks for block GS) per colour
  int ncolors = 0;
  DevBuf diag, dinv;
  // Direct level: lower 64 x 64 tiles of the symmetric FP64 inverse, the
  // strip table (tile column, first and last tile row) and the partial sums
  // of the tiled product.
  DevBuf inv, sym_strips, sym_prow, sym_pcol;
  int64_t sym_nt = 0, sym_kmax = 0, sym_nstrips = 0;
  DevBuf rhs_buf, x_buf, res, tmp;
  TM *rhs = nullptr, *x = nullptr;  // bound to own buffers or the CG vectors
  // Where the level's smoothed x lies after its last encoded smoothing: x,
  // or (Jacobi levels below the finest) the second buffer tmp, which the
  // prolongation then reads instead of copying it back.
  TM *xout = nullptr;
  DMat<TM> R, P;                     // transfer to level k + 1
  std::vector<int32_t> q;            // internal index -> level index
  // Single-workgroup tail (small scalar Gauss-Seidel levels): the level
  // operator, the transfers to level k + 1 and the colour offsets as
  // row-contiguous CSR in the internal order.
  bool tail_ok = false;
  gsd::DRows<TM> tA, tR, tP;
  DevBuf tcoff;
};

template <class TC, class TM>
class Engine final : public EngineBase {
 public:
  explicit Engine(const Options &o) : opt_(o) {
    GS_HIP(hipStreamCreate(&stream_));
  }
  ~Engine() override {
    drop_graphs();
    if (host_ctl_) (void)hipHostFree(host_ctl_);
    if (stage_) (void)hipHostFree(stage_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    (void)hipStreamDestroy(stream_);
  }

  void setup(const gsx::Hierarchy &H, const std::vector<gsx::Perm> &perms) override;
  SolveResult solve(const double *b, double *x, double tol, int64_t maxit,
                    const LagGuard *guard) override;
  void vcycle_level_order(const double *r, double *z) override;
  void spmv_level_order(const double *x, double *y) override;
  double time_kernel(int which, int reps, double *bytes) override;
  std::string describe() const override;
  int64_t size() const override { return n_; }
  const std::vector<double> &history() const override { return hist_host_; }
  void level_info(std::vector<LevelInfo> &out) const override;
  void set_level_weights(const std::vector<double> &omega, const std::vector<int> &sweeps) override {
    const int nl = std::min<int>(nlev_, static_cast<int>(L_.size()));
    // Validate every entry first, so a rejected entry leaves the weights
    // (and the captured graph) as they were.
    for (int k = 0; k < nl; ++k) {
      const Level<TM> &lv = L_[k];
      if (k < static_cast<int>(omega.size()) && omega[k] > 0.0 &&
          (lv.kind == kSGS || lv.kind == kBlockGS) && !(omega[k] < 2.0))
        throw std::runtime_error("Gauss-Seidel smoother: omega must be in (0, 2).");
      // A per-level weight may leave a coarse level unsmoothed (the
      // V-cycle stays symmetric); the finest level keeps the reference's
      // rule for symmetric Gauss-Seidel.
      if (k < static_cast<int>(sweeps.size()) && sweeps[k] == 0 && lv.kind == kSGS &&
          !lv.direct && k == 0 && k + 1 < nlev_)
        throw std::runtime_error("Preconditioner requires at least one smoothing step.");
    }
    drop_graphs();
    for (int k = 0; k < nl; ++k) {
      Level<TM> &lv = L_[k];
      if (k < static_cast<int>(omega.size()) && omega[k] > 0.0)
        lv.omega = static_cast<TM>(omega[k]);
      if (k < static_cast<int>(sweeps.size()) && sweeps[k] >= 0) lv.pre = lv.post = sweeps[k];
    }
  }

 private:
  void build_level(int k, const gsx::Csr &Ak_level, bool coarsest);
  // zero_x: the level's x starts at zero; x_zeroed: and that zero is
  // already in memory (written by the restriction).
  // first_done: the restriction into level k already wrote its first
  // Jacobi sweep from x = 0 (closed form).
  void encode_vcycle(int k, const int32_t *stop, bool zero_x, bool x_zeroed = false,
                     bool first_done = false);
  // may_leave: a Jacobi result may stay in lv.tmp (lv.xout says where it
  // is); only for smoothing whose result the prolongation reads next.
  void encode_smooth(Level<TM> &lv, int iters, bool zero_x, const int32_t *stop,
                     bool x_zeroed = false, bool may_leave = false, bool first_done = false);
  void encode_iteration(const int32_t *stop, bool with_replace = true);
  void encode_residual(Level<TM> &lv, const int32_t *stop);
  // dinv_c: also write the coarse level's closed-form first Jacobi sweep
  // into corr_c (the k_restrict kernels' dinv_c).
  void encode_restrict(int k, const TM *res, TM *corr_c, const int32_t *stop,
                       const TM *dinv_c = nullptr, TM omega_c = TM(0));
  void encode_prolong(int k, bool alias, const int32_t *stop);
  void encode_tail(const int32_t *stop);
  void capture_graph();
  void apply_preconditioner(const int32_t *stop);

  Options opt_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  int64_t n_ = 0;
  int nlev_ = 0;
  std::vector<Level<TM>> L_;
  DOp<TC> Acg_;
  std::vector<int32_t> int2orig_;  // internal index -> user (original) index
  std::vector<int32_t> int2lvl0_;  // internal index -> level-0 (permuted) index
  // CG vectors
  DevBuf b_, x_, r_, p_, q_, z_, rm_, partials_, ctl_, hist_, tr_;
  TM *rvc_ = nullptr;  // V-cycle view of r (r itself or its rounded copy)
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  // One PCG iteration (with the replacement pair when the mode has one),
  // replayed for the last iterations of a solve (Options::graph_tail).
  hipGraph_t graph1_ = nullptr;
  hipGraphExec_t graph1_exec_ = nullptr;
  // Loop iterations of the last converged solve (0: none since the last
  // setup or weight change): the first launch of the next solve runs at
  // most that many iterations, so a repeated solve shorter than a chunk
  // launches no idle iterations either.
  int64_t last_loop_iters_ = 0;
  void drop_graphs() {
    last_loop_iters_ = 0;
    if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    if (graph1_exec_) { (void)hipGraphExecDestroy(graph1_exec_); graph1_exec_ = nullptr; }
    if (graph1_) { (void)hipGraphDestroy(graph1_); graph1_ = nullptr; }
  }
  gsk::Ctl<TC> *host_ctl_ = nullptr;
  TC *stage_ = nullptr;  // pinned staging for b and x (2 n values)
  int64_t stage_n_ = 0;
  int64_t hist_cap_ = 0;
  std::vector<double> hist_host_;
  bool setup_done_ = false;
  // The CG SpMV reads level 0's V-cycle-type operator (cg_vcycle_operator,
  // when the types differ and level 0 is stored one row or block row per
  // thread); r is then replaced from the FP64 operator.
  bool cg_vop_ = false;
  // The CG operator is stored as planes of value pairs (FP64, block SELL).
  bool cg_pairs_ = false;
  // Kernel nodes per PCG iteration in the captured graph (0: none yet).
  double graph_nodes_per_iter_ = 0.0;
  // First level of the single-workgroup tail (-1: none) and its LDS bytes.
  int tail_k0_ = -1;
  int64_t tail_lds_ = 0;
  // Residual replacement period: mixed precision only (the FP64 operator
  // recomputes r = b - A x and confirms every stop on convergence).
  int64_t replace_every() const {
    return std::is_same<TC, TM>::value ? 0 : (cg_vop_ ? opt_.cg_replace_every : opt_.cg_true_residual_every);
  }
  // Level 0's block CSR packed on the device, kept until the CG operator
  // (the same blocks when level 0 is not permuted) has been built.
  std::unique_ptr<gsdev::DBsr3Csr> dev_l0_;
  // Host wall time of the setup phases (ms), in order, for describe().
  std::vector<std::pair<std::string, double>> setup_ms_;
  std::chrono::steady_clock::time_point tmark_;
  void mark(const std::string &phase) {
    const auto t = std::chrono::steady_clock::now();
    setup_ms_.emplace_back(phase, std::chrono::duration<double, std::milli>(t - tmark_).count());
    tmark_ = t;
  }
};

template <class TC, class TM>
void Engine<TC, TM>::build_level(int k, const gsx::Csr &Ak, bool coarsest) {
  Level<TM> &lv = L_[k];
  lv.n = Ak.rows;
  lv.kind = (k == 0) ? opt_.fine_kind : opt_.coarse_kind;
  lv.direct = coarsest && opt_.coarse_direct;
  double om = opt_.omega;
  int pre = opt_.pre, post = opt_.post;
  if (k < static_cast<int>(opt_.level_omega.size()) && opt_.level_omega[k] > 0.0)
    om = opt_.level_omega[k];
  // Per-level sweeps (learned weights): 0 leaves a level below the finest
  // unsmoothed (no symmetric Gauss-Seidel sweep; the V-cycle stays
  // symmetric). Without them, the reference's rule holds (src/Smoothers/
  // SymmetricGaussSeidelSmoother.cc:89-92).
  bool level_off = false;
  if (k < static_cast<int>(opt_.level_sweeps.size()) && opt_.level_sweeps[k] >= 0) {
    pre = post = opt_.level_sweeps[k];
    level_off = pre == 0 && k > 0 && !coarsest;
  }
  lv.omega = static_cast<TM>(om);
  lv.pre = pre;
  lv.post = post;
  const bool blocky = lv.kind == kBlockGS || lv.kind == kBlockJacobi;
  if (!lv.direct) {
    if (lv.kind == kSGS) {
      if (!std::isfinite(om) || om <= 0.0 || om >= 2.0)
        throw std::runtime_error("SymmetricGaussSeidelSmoother: invalid omega.");
      const int steps_needed = coarsest ? opt_.coarse_steps : std::min(pre, post);
      if (steps_needed == 0 && !level_off)
        throw std::runtime_error("Preconditioner requires at least one smoothing step.");
    }
    if (lv.kind == kBlockGS && (om <= 0.0 || om >= 2.0))
      throw std::runtime_error("BlockGaussSeidelSmoother: omega must be in (0, 2).");
    if (blocky && lv.n % 3 != 0)
      throw std::runtime_error("BlockGaussSeidelSmoother: number of DOFs must be divisible by 3.");
  }
  mark("level " + std::to_string(k) + " checks");
  // Internal ordering.
  lv.coff.clear();
  if (!lv.direct && lv.kind == kSGS) {
    gsx::Coloring c = gsx::greedy_coloring(Ak.rows, Ak.ptr.data(), Ak.col.data());
    lv.q = c.vertices;
    lv.coff.assign(c.offsets.begin(), c.offsets.end());
    lv.ncolors = c.ncolors;
  } else if (!lv.direct && lv.kind == kBlockGS) {
    gsx::Graph g = gsx::block_graph(Ak, 3);
    gsx::Coloring c = gsx::greedy_coloring(Ak.rows / 3, g.ptr.data(), g.col.data());
    lv.q.resize(Ak.rows);
    for (int64_t i = 0; i < Ak.rows / 3; ++i)
      for (int l = 0; l < 3; ++l) lv.q[3 * i + l] = 3 * c.vertices[i] + l;
    lv.coff.assign(c.offsets.begin(), c.offsets.end());
    lv.ncolors = c.ncolors;
  } else {
    lv.q = iota_vec(Ak.rows);
    lv.ncolors = 0;
  }
  mark("level " + std::to_string(k) + " colouring");
  const std::vector<int32_t> qinv = inverse_of(lv.q);
  // Block-smoother levels: relabelled and packed on the device from one
  // upload of Ak (the host copy Ai is then not formed).
  const bool dev_pack = opt_.device_pack && blocky && !lv.direct && Ak.rows % 3 == 0 &&
                        keeps_triples(lv.q);
  gsx::Csr Ai;
  if (!dev_pack) Ai = relabel(Ak, lv.q, qinv);
  mark("level " + std::to_string(k) + " relabel");
  if (lv.direct) {
    if (lv.n > opt_.max_direct)
      throw std::runtime_error("coarsest level has " + std::to_string(lv.n) +
                               " unknowns, above the dense direct-solve limit of " +
                               std::to_string(opt_.max_direct) + "; coarsen further.");
    const int64_t N = gsdense::padded_size(lv.n);
    {
      DevBuf full(static_cast<size_t>(N) * N * sizeof(double));
      gsdense::spd_inverse_device_csr(lv.n, Ai.ptr.data(), Ai.col.data(), Ai.val.data(),
                                      full.as<double>(), stream_);
      lv.inv.alloc(static_cast<size_t>(gsdense::lower_tiles_size(lv.n)) * sizeof(double));
      gsdense::pack_lower_tiles(lv.n, full.as<double>(), lv.inv.template as<double>(), stream_);
      GS_HIP(hipStreamSynchronize(stream_));
    }
    const int64_t nt = N / 64, S = gsk::kSymStrip;
    std::vector<int32_t> strips;
    for (int64_t J = 0; J < nt; ++J)
      for (int64_t I0 = J; I0 < nt; I0 += S) {
        strips.push_back(static_cast<int32_t>(J));
        strips.push_back(static_cast<int32_t>(I0));
        strips.push_back(static_cast<int32_t>(std::min(I0 + S - 1, nt - 1)));
      }
    lv.sym_nt = nt;
    lv.sym_kmax = (nt + S - 1) / S;
    lv.sym_nstrips = static_cast<int64_t>(strips.size() / 3);
    lv.sym_strips.upload(strips);
    lv.sym_prow.alloc(static_cast<size_t>(nt * (nt + 1) / 2) * 64 * sizeof(double));
    lv.sym_pcol.alloc(static_cast<size_t>(nt * lv.sym_kmax) * 64 * sizeof(double));
  }
  if (lv.direct) mark("level " + std::to_string(k) + " dense inverse");
  // Blocks summed in the level's own (pre-relabelling) column order, the
  // order of the CSR rows the reference walks.
  std::vector<int32_t> key;
  if (blocky && lv.n % 3 == 0) {
    key.resize(lv.n / 3);
    for (int64_t j = 0; j < lv.n / 3; ++j) key[j] = lv.q[3 * j] / 3;
  }
  // A Jacobi level's operator is read only by whole-level kernels (sweep,
  // residual), which take its own lane-group threshold.
  const bool jacobi_level = lv.kind == kDampedJacobi || lv.kind == kBlockJacobi;
  std::unique_ptr<gsdev::DBsr3Csr> DB;
  if (dev_pack) {
    DB = std::make_unique<gsdev::DBsr3Csr>();
    {
      const gsdev::DCsr dAk = gsdev::upload(Ak);
      const gsdev::DCsr dAi = gsdev::relabel(dAk, lv.q, qinv);
      gsdev::bsr3_pack(dAi, key.data(), *DB);
    }
    build_b3_device(lv.A, *DB, jacobi_level ? opt_.full_group_min_block : opt_.group_min_block,
                    nullptr);
  } else {
    lv.A.build(Ai, blocky, key.empty() ? nullptr : key.data(),
               jacobi_level ? opt_.full_group_min_scalar : opt_.group_min_scalar,
               jacobi_level ? opt_.full_group_min_block : opt_.group_min_block);
  }
  lv.has_low = false;
  if (dev_pack && lv.kind == kBlockGS && k < opt_.zero_start_lower && lv.ncolors > 1) {
    // The earlier-colour blocks: block column below the row's colour range.
    std::vector<int32_t> start(lv.n / 3);
    for (int c = 0; c < lv.ncolors; ++c)
      for (int64_t u = lv.coff[c]; u < lv.coff[c + 1]; ++u)
        start[u] = static_cast<int32_t>(lv.coff[c]);
    const gsdev::DBsr3Csr L = gsdev::bsr3_below(*DB, start);
    build_b3_device(lv.Alow, L, lv.A.G ? 0.0 : -1.0, nullptr);
    lv.has_low = true;
  } else if (!dev_pack && !lv.direct && (lv.kind == kSGS || lv.kind == kBlockGS) &&
             k < opt_.zero_start_lower && lv.ncolors > 1) {
    // Colours are contiguous internal ranges: an entry belongs to an earlier
    // colour exactly when its column (block column) lies below the row's
    // colour range. Row order and entry order are kept.
    const int unit = lv.kind == kBlockGS ? 3 : 1;
    gsx::Csr Low;
    Low.rows = Ai.rows;
    Low.cols = Ai.cols;
    Low.ptr.assign(Ai.rows + 1, 0);
    std::vector<int64_t> lo_of(Ai.rows);
    for (int c = 0; c < lv.ncolors; ++c)
      for (int64_t u = lv.coff[c]; u < lv.coff[c + 1]; ++u)
        for (int l = 0; l < unit; ++l) lo_of[unit * u + l] = unit * lv.coff[c];
    for (int64_t r = 0; r < Ai.rows; ++r) {
      int32_t cnt = 0;
      for (int32_t e = Ai.ptr[r]; e < Ai.ptr[r + 1]; ++e) cnt += Ai.col[e] < lo_of[r];
      Low.ptr[r + 1] = Low.ptr[r] + cnt;
    }
    Low.col.resize(Low.ptr[Ai.rows]);
    Low.val.resize(Low.ptr[Ai.rows]);
    for (int64_t r = 0; r < Ai.rows; ++r) {
      int64_t d = Low.ptr[r];
      for (int32_t e = Ai.ptr[r]; e < Ai.ptr[r + 1]; ++e)
        if (Ai.col[e] < lo_of[r]) {
          Low.col[d] = Ai.col[e];
          Low.val[d] = Ai.val[e];
          ++d;
        }
    }
    // Same kernel family as the full operator (lane groups when it uses
    // them; its half-length rows would otherwise fall back to one latency-
    // bound thread per row on the coarse levels); the group size follows
    // the lower rows' own length.
    const double gmin = lv.A.G ? 0.0 : -1.0;
    lv.Alow.build(Low, blocky, key.empty() ? nullptr : key.data(), gmin, gmin);
    lv.has_low = lv.Alow.b3 == lv.A.b3;
  }
  // Every colour of a tail level fits one round of the workgroup's row
  // groups (at most kTailNG rows): a colour step in the tail costs about
  // what one launch of it does, so a level with a larger colour, which
  // would take several rounds in one workgroup, stays on launches.
  int64_t widest = 0;
  for (int c = 0; c < lv.ncolors; ++c) widest = std::max<int64_t>(widest, lv.coff[c + 1] - lv.coff[c]);
  lv.tail_ok = k > 0 && !lv.direct && lv.kind == kSGS && opt_.tail_rows > 0 &&
               lv.n <= opt_.tail_rows && lv.ncolors > 0 && widest <= gsk::kTailNG;
  if (lv.tail_ok) {
    lv.tA.build(Ai);
    lv.tcoff.upload(lv.coff);
  }
  mark("level " + std::to_string(k) + " pack + upload");
  // Stored entries (3x3 blocks for block GS) per colour, for the bytes of a
  // colour launch.
  lv.cunits.assign(lv.ncolors, 0);
  for (int c = 0; c < lv.ncolors; ++c) {
    if (lv.kind == kBlockGS) {
      if (!lv.A.b3) throw std::runtime_error("block Gauss-Seidel level without block storage.");
      lv.cunits[c] = lv.A.bptr[lv.coff[c + 1]] - lv.A.bptr[lv.coff[c]];
    } else {
      lv.cunits[c] = Ai.ptr[lv.coff[c + 1]] - Ai.ptr[lv.coff[c]];
    }
  }
  if (!lv.direct) {
    if (lv.kind == kSGS || lv.kind == kDampedJacobi) {
      std::vector<TM> d(lv.n);
      std::vector<TM> di(lv.n);
      for (int64_t r = 0; r < lv.n; ++r) {
        double dv = 0.0;
        bool found = false;
        for (int32_t e = Ai.ptr[r]; e < Ai.ptr[r + 1]; ++e)
          if (Ai.col[e] == r) { dv = Ai.val[e]; found = true; }
        if (!found) {
          if (lv.kind == kSGS)
            throw std::runtime_error("CPUBackend::encodeGaussSeidelColor: zero diagonal entry.");
          throw std::runtime_error("DampedJacobiPreconditioner::initialize: missing diagonal entry.");
        }
        if (std::abs(static_cast<TM>(dv)) < TM(1e-12)) {
          if (lv.kind == kSGS)
            throw std::runtime_error("CPUBackend::encodeGaussSeidelColor: zero diagonal entry.");
          throw std::runtime_error("DampedJacobiPreconditioner::initialize: zero diagonal entry.");
        }
        d[r] = static_cast<TM>(dv);
        di[r] = TM(1) / static_cast<TM>(dv);
      }
      if (lv.kind == kSGS) lv.diag.upload(d);
      else lv.dinv.upload(di);
    } else if (dev_pack) {
      // The inverse diagonal blocks of Ak moved to the internal block
      // order: internal block row b is Ak's block row q[3b] / 3, with the
      // same diagonal block.
      const std::vector<double> dk = gsx::inverse_diag_blocks3(Ak);
      std::vector<double> dinv(dk.size());
      for (int64_t b = 0; b < lv.n / 3; ++b)
        std::copy(&dk[9 * (lv.q[3 * b] / 3)], &dk[9 * (lv.q[3 * b] / 3)] + 9, &dinv[9 * b]);
      gsd::upload_values<TM>(lv.dinv, dinv.data(), dinv.size());
    } else {
      const std::vector<double> dinv = gsx::inverse_diag_blocks3(Ai);
      gsd::upload_values<TM>(lv.dinv, dinv.data(), dinv.size());
    }
  }
  if (dev_pack && k == 0) dev_l0_ = std::move(DB);
  if (k > 0) {
    lv.rhs_buf.alloc(lv.n * sizeof(TM));
    lv.x_buf.alloc(lv.n * sizeof(TM));
    lv.rhs = lv.rhs_buf.template as<TM>();
    lv.x = lv.x_buf.template as<TM>();
  }
  lv.res.alloc(lv.n * sizeof(TM));
  lv.tmp.alloc(lv.n * sizeof(TM));
  mark("level " + std::to_string(k) + " diagonal + vectors");
}

template <class TC, class TM>
void Engine<TC, TM>::setup(const gsx::Hierarchy &H, const std::vector<gsx::Perm> &perms) {
  setup_ms_.clear();
  tmark_ = std::chrono::steady_clock::now();
  nlev_ = static_cast<int>(H.A.size());
  if (nlev_ == 0) throw std::runtime_error("MultigridSmoother: empty hierarchy.");
  if (static_cast<int>(perms.size()) != nlev_)
    throw std::runtime_error("MultigridSmoother: invalid number of permutations.");
  if (static_cast<int>(H.P.size()) + 1 != nlev_ || static_cast<int>(H.R.size()) + 1 != nlev_)
    throw std::runtime_error("MultigridSmoother: invalid number of transfers.");
  const int hier_levels = nlev_;
  for (int k = 0; k < hier_levels; ++k) {
    H.A[k].check();
    if (H.A[k].rows != H.A[k].cols) throw std::runtime_error("CGSolver: matrix must be square.");
    if (static_cast<int64_t>(perms[k].old_to_new.size()) != H.A[k].rows)
      throw std::runtime_error("MultigridSmoother: permutation size does not match level size.");
    perms[k].check();
  }
  // Transfer shapes before any permutation is applied to them.
  for (int k = 0; k + 1 < hier_levels; ++k) {
    H.P[k].check();
    H.R[k].check();
    if (H.P[k].rows != H.A[k].rows || H.P[k].cols != H.A[k + 1].rows ||
        H.R[k].rows != H.A[k + 1].rows || H.R[k].cols != H.A[k].rows)
      throw std::runtime_error("MultigridSmoother: transfer operator shape mismatch.");
  }
  if (opt_.precond == kPrecIdentity && !std::is_same<TC, TM>::value)
    throw std::runtime_error("identity preconditioner requires one precision for CG and V-cycle.");
  // Direct solve at the first level small enough for the dense inverse; the
  // levels below it are not used.
  if (opt_.coarse_direct && opt_.direct_truncate) {
    for (int k = 0; k < hier_levels; ++k)
      if (H.A[k].rows <= opt_.max_direct) {
        nlev_ = k + 1;
        break;
      }
  }
  n_ = H.A[0].rows;
  L_.clear();
  L_.resize(nlev_);
  const int used_levels = opt_.precond == kPrecMultigrid ? nlev_ : 1;
  mark("validate");
  for (int k = 0; k < used_levels; ++k) {
    gsx::Csr store;
    const gsx::Csr &Ak = permuted(H.A[k], perms[k], perms[k], store);
    mark("level " + std::to_string(k) + " permute");
    if (opt_.precond == kPrecMultigrid) build_level(k, Ak, k == nlev_ - 1);
    else {
      L_[0].n = Ak.rows;
      L_[0].q = iota_vec(Ak.rows);
    }
  }
  if (opt_.precond == kPrecMultigrid) {
    for (int k = 0; k + 1 < nlev_; ++k) {
      gsx::Csr p_store, r_store;
      const gsx::Csr &Pk = permuted(H.P[k], perms[k], perms[k + 1], p_store);
      const gsx::Csr &Rk = permuted(H.R[k], perms[k + 1], perms[k], r_store);
      if (Pk.rows != L_[k].n || Pk.cols != L_[k + 1].n || Rk.rows != L_[k + 1].n ||
          Rk.cols != L_[k].n)
        throw std::runtime_error("MultigridSmoother: transfer operator shape mismatch.");
      const std::vector<int32_t> qf_inv = inverse_of(L_[k].q);
      const std::vector<int32_t> qc_inv = inverse_of(L_[k + 1].q);
      // Level order: position p holds the internal row of level index p.
      const bool lo = opt_.transfer_level_order;
      const gsx::Csr Pi = relabel(Pk, L_[k].q, qc_inv);
      const gsx::Csr Ri = relabel(Rk, L_[k + 1].q, qf_inv);
      if (opt_.block_transfers && keeps_triples(L_[k].q) && keeps_triples(L_[k + 1].q)) {
        // 3x3 blocks: each block row's blocks in the CSR rows' column order
        // (ascending level index), block rows in level order.
        const int64_t nfb = L_[k].n / 3, ncb = L_[k + 1].n / 3;
        std::vector<int32_t> key_c(ncb), key_f(nfb), ord_f(nfb), ord_c(ncb);
        for (int64_t j = 0; j < ncb; ++j) {
          key_c[j] = L_[k + 1].q[3 * j] / 3;
          ord_c[j] = qc_inv[3 * j] / 3;
        }
        for (int64_t j = 0; j < nfb; ++j) {
          key_f[j] = L_[k].q[3 * j] / 3;
          ord_f[j] = qf_inv[3 * j] / 3;
        }
        L_[k].P.build_blocks(Pi, key_c.data(), opt_.full_group_min_block, opt_.transfer_sigma,
                             lo ? ord_f.data() : nullptr);
        L_[k].R.build_blocks(Ri, key_f.data(), opt_.full_group_min_block, opt_.transfer_sigma,
                             lo ? ord_c.data() : nullptr);
      } else {
        L_[k].P.build(Pi, opt_.full_group_min_scalar, opt_.group_max_transfer,
                      opt_.transfer_sigma, lo ? qf_inv.data() : nullptr);
        L_[k].R.build(Ri, opt_.full_group_min_scalar, opt_.group_max_transfer,
                      opt_.transfer_sigma, lo ? qc_inv.data() : nullptr);
      }
      if (L_[k].tail_ok && L_[k + 1].tail_ok) {
        L_[k].tR.build(Ri);
        L_[k].tP.build(Pi);
      }
      mark("transfers " + std::to_string(k));
    }
  }
  // The tail: the longest run of tail levels ending at the coarsest whose
  // transfers were kept, within the kernel's level limit and one
  // workgroup's LDS (every tail level's vectors and row offsets, and the
  // staging slots).
  tail_k0_ = -1;
  tail_lds_ = 0;
  if (opt_.precond == kPrecMultigrid) {
    int k0 = nlev_;
    int64_t bytes = int64_t(gsk::kTailBlock) * gsk::kTailKS * sizeof(TM);
    while (k0 - 1 >= 1 && L_[k0 - 1].tail_ok && (k0 == nlev_ || L_[k0 - 1].tR.rows > 0) &&
           nlev_ - (k0 - 1) <= gsk::kTailMaxLevels &&
           bytes + gsk::tail_level_bytes<TM>(L_[k0 - 1].n) <= gsk::kTailLdsMax) {
      bytes += gsk::tail_level_bytes<TM>(L_[k0 - 1].n);
      --k0;
    }
    if (k0 < nlev_) {
      tail_k0_ = k0;
      tail_lds_ = bytes;
      GS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&gsk::k_vcycle_tail<TM>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(bytes)));
    }
  }
  // CG operator: level-0 input with the order-keeping permutation, then the
  // internal relabelling of level 0.
  {
    // One keep-order relabelling of the input: internal row i is input row
    // new_to_old[q[i]], input column c becomes q^-1[old_to_new[c]].
    const std::vector<int32_t> q0inv = inverse_of(L_[0].q);
    std::vector<int32_t> rq(n_), cmap(n_);
    for (int64_t i = 0; i < n_; ++i) {
      rq[i] = perms[0].new_to_old[L_[0].q[i]];
      cmap[i] = q0inv[perms[0].old_to_new[i]];
    }
    const bool blocky = opt_.fine_kind == kBlockGS || opt_.fine_kind == kBlockJacobi ||
                        opt_.cg_block3;
    const bool dev_cg = opt_.device_pack && blocky && n_ % 3 == 0;
    gsx::Csr Ai;
    if (!dev_cg) Ai = relabel(H.A[0], rq, cmap);
    int2lvl0_ = L_[0].q;
    int2orig_.resize(n_);
    for (int64_t i = 0; i < n_; ++i) int2orig_[i] = perms[0].new_to_old[L_[0].q[i]];
    // The keep-order permutation leaves each row in the input's column
    // order: rank block columns by their input index.
    std::vector<int32_t> key;
    if (blocky && n_ % 3 == 0) {
      key.resize(n_ / 3);
      for (int64_t j = 0; j < n_ / 3; ++j) key[j] = int2orig_[3 * j] / 3;
    }
    // Block rows executed in level-0 (spatial) order when level 0 moves
    // node triples together: neighbouring rows run together and share
    // their x gathers in cache.
    std::vector<int32_t> ord;
    if (blocky && n_ % 3 == 0 && opt_.cg_level_order && keeps_triples(L_[0].q)) {
      ord.resize(n_ / 3);
      for (int64_t p = 0; p < n_ / 3; ++p) ord[p] = q0inv[3 * p] / 3;
    }
    Acg_.b.pairs = opt_.cg_pair_layout && std::is_same<TC, double>::value;
    if (dev_cg) {
      // With level 0 unpermuted and its input already sorted, this is the
      // block CSR level 0 was packed from (same relabelling, same block
      // ranks); otherwise the input is uploaded and packed here.
      const bool reuse = dev_l0_ && is_identity(perms[0].old_to_new) &&
                         gsx::rows_sorted_unique(H.A[0]);
      gsdev::DBsr3Csr own;
      if (!reuse) {
        const gsdev::DCsr dA = gsdev::upload(H.A[0]);
        const gsdev::DCsr dAi = gsdev::relabel(dA, rq, cmap);
        gsdev::bsr3_pack(dAi, key.data(), own);
      }
      build_b3_device(Acg_, reuse ? *dev_l0_ : own, -1.0, ord.empty() ? nullptr : ord.data());
    } else {
      Acg_.build(Ai, blocky && n_ % 3 == 0, key.empty() ? nullptr : key.data(), -1.0, -1.0,
                 ord.empty() ? nullptr : ord.data());
    }
    dev_l0_.reset();
    cg_pairs_ = Acg_.b3 && Acg_.b.pairs;
    mark("CG operator");
  }
  // CG vectors and control block.
  b_.alloc(n_ * sizeof(TC));
  x_.alloc(n_ * sizeof(TC));
  r_.alloc(n_ * sizeof(TC));
  p_.alloc(n_ * sizeof(TC));
  q_.alloc(n_ * sizeof(TC));
  z_.alloc(n_ * sizeof(TM));
  if (!std::is_same<TC, TM>::value) {
    rm_.alloc(n_ * sizeof(TM));
    rvc_ = rm_.template as<TM>();
  } else {
    rvc_ = reinterpret_cast<TM *>(r_.template as<TC>());
  }
  L_[0].rhs = rvc_;
  L_[0].x = z_.template as<TM>();
  cg_vop_ = opt_.cg_vcycle_operator && !std::is_same<TC, TM>::value &&
            opt_.precond == kPrecMultigrid && !L_[0].direct && L_[0].A.G == 0;
  // Partials: one per workgroup of the widest reduction (three segments of
  // the n-thread grid, or one per workgroup of a persistent launch).
  if (opt_.persist_grid < 8 || opt_.persist_grid % 8 != 0)
    throw std::runtime_error("persist_grid must be a positive multiple of 8.");
  partials_.alloc((std::max<size_t>(3 * static_cast<size_t>(grid_for(n_)),
                                    static_cast<size_t>(opt_.persist_grid)) + 64) *
                  sizeof(double));
  ctl_.alloc(sizeof(gsk::Ctl<TC>));
  tr_.alloc(2 * sizeof(double));
  if (!host_ctl_) GS_HIP(hipHostMalloc(reinterpret_cast<void **>(&host_ctl_), sizeof(gsk::Ctl<TC>)));
  std::memset(host_ctl_, 0, sizeof(gsk::Ctl<TC>));
  if (stage_n_ != n_) {
    if (stage_) GS_HIP(hipHostFree(stage_));
    stage_ = nullptr;
    stage_n_ = 0;
    GS_HIP(hipHostMalloc(reinterpret_cast<void **>(&stage_), 2 * n_ * sizeof(TC)));
    stage_n_ = n_;
  }
  if (!ev0_) GS_HIP(hipEventCreate(&ev0_));
  if (!ev1_) GS_HIP(hipEventCreate(&ev1_));
  drop_graphs();
  mark("vectors");
  setup_done_ = true;
}

// ---------------------------------------------------------------------------
// V-cycle encoding.
// ---------------------------------------------------------------------------
template <class TC, class TM>
void Engine<TC, TM>::encode_residual(Level<TM> &lv, const int32_t *stop) {
  const int64_t n = lv.n;
  TM *res = lv.res.template as<TM>();
  if (lv.A.G) {
    with_group(lv.A.G, [&](auto g) {
      constexpr int G = decltype(g)::value;
      if (lv.A.b3)
        gsk::k_residual_bw<TM, G><<<gsk::grid_rows<G>(n / 3), kBlock, 0, stream_>>>(
            lv.A.bw.view(), lv.rhs, lv.x, res, stop);
      else
        gsk::k_residual_w<TM, G><<<gsk::grid_rows<G>(n), kBlock, 0, stream_>>>(
            lv.A.w.view(), lv.rhs, lv.x, res, stop);
    });
  } else if (lv.A.b3) {
    gsk::k_residual_b3<TM><<<grid_for(n / 3), kBlock, 0, stream_>>>(lv.A.b.view(), lv.rhs, lv.x,
                                                                    res, stop);
  } else {
    gsk::k_residual<TM><<<grid_for(n), kBlock, 0, stream_>>>(lv.A.s.view(), lv.rhs, lv.x, res,
                                                             stop);
  }
}

// The colour sequence of `iters` symmetric sweeps (forward 0..C-1, backward
// C-1..0 each), with back-to-back repeats of one colour merged into a
// single launch that applies the update twice.
static std::vector<std::pair<int, int>> sweep_sequence(int ncolors, int iters) {
  std::vector<int> seq;
  for (int it = 0; it < iters; ++it) {
    for (int c = 0; c < ncolors; ++c) seq.push_back(c);
    for (int c = ncolors - 1; c >= 0; --c) seq.push_back(c);
  }
  std::vector<std::pair<int, int>> out;
  for (int c : seq) {
    if (!out.empty() && out.back().first == c && out.back().second == 1) out.back().second = 2;
    else out.emplace_back(c, 1);
  }
  return out;
}

template <class TC, class TM>
void Engine<TC, TM>::encode_smooth(Level<TM> &lv, int iters, bool zero_x, const int32_t *stop,
                                   bool x_zeroed, bool may_leave, bool first_done) {
  const int64_t n = lv.n;
  bool x_ready = !zero_x;
  lv.xout = lv.x;
  if ((lv.kind == kSGS || lv.kind == kBlockGS) && iters > 0) {
    const bool block = lv.kind == kBlockGS;
    if (block && !lv.A.b3)
      throw std::runtime_error("block Gauss-Seidel level is not 3x3-block representable.");
    const int64_t units = block ? n / 3 : n;
    // From x = 0, the first sweep's forward colours 1 .. C-1 (steps 1 .. C-1
    // of the sequence) see exact zeros at every later-colour neighbour, so
    // they read the earlier-colour entries only; the products they leave out
    // are +-0 and a row sum never is -0, so the sums are unchanged.
    const bool low_start = !x_ready && lv.has_low;
    int step_no = -1;
    for (const auto &step : sweep_sequence(lv.ncolors, iters)) {
      ++step_no;
      const int c = step.first;
      int reps = step.second;
      const int64_t b = lv.coff[c], e = lv.coff[c + 1];
      // The first sweep's forward colours 1 .. C-1 from x = 0 with the
      // earlier-colour operator: their own x is still zero (xz) and later
      // colours' x is never read before its own visit writes it, so only
      // colour 0 is written by the first kernel.
      const int xz = (low_start && step_no >= 1 && step_no <= lv.ncolors - 1) ? 1 : 0;
      const DOp<TM> &M = xz ? lv.Alow : lv.A;
      if (!x_ready && c == 0) {
        // First update from x = 0: closed form for colour 0, zero elsewhere
        // (nowhere else with the earlier-colour operator).
        const int64_t zn = low_start ? e : units;
        if (block)
          gsk::k_block_gs_color_zero<TM><<<grid_for(zn), kBlock, 0, stream_>>>(
              lv.dinv.template as<TM>(), e, zn, lv.rhs, lv.x, lv.omega, stop);
        else
          gsk::k_gs_color_zero<TM><<<grid_for(zn), kBlock, 0, stream_>>>(
              lv.diag.template as<TM>(), e, zn, lv.rhs, lv.x, lv.omega, stop);
        x_ready = true;
        if (reps == 1) continue;
        reps = 1;  // the second update of the pair reads the fresh values
      }
      if (e <= b) continue;
      if (M.G) {
        with_group(M.G, [&](auto g) {
          constexpr int G = decltype(g)::value;
          if (block)
            gsk::k_block_gs_color_bw<TM, G><<<gsk::grid_rows<G>(e - b), kBlock, 0, stream_>>>(
                M.bw.view(), lv.dinv.template as<TM>(), b, e, lv.rhs, lv.x, lv.omega, reps,
                stop, xz);
          else
            gsk::k_gs_color_w<TM, G><<<gsk::grid_rows<G>(e - b), kBlock, 0, stream_>>>(
                M.w.view(), lv.diag.template as<TM>(), b, e, lv.rhs, lv.x, lv.omega, reps,
                stop, xz);
        });
      } else if (block) {
        constexpr int CB = gsk::kColorBlock;
        const int64_t b0 = b & ~int64_t(gsk::kSlice - 1);  // slice-aligned start
        gsk::k_block_gs_color<TM><<<static_cast<uint32_t>((e - b0 + CB - 1) / CB), CB, 0, stream_>>>(
            M.b.view(), lv.dinv.template as<TM>(), b, e, lv.rhs, lv.x, lv.omega, reps, stop, xz);
      } else {
        gsk::k_gs_color<TM><<<grid_for(e - b), kBlock, 0, stream_>>>(
            M.s.view(), lv.diag.template as<TM>(), b, e, lv.rhs, lv.x, lv.omega, reps, stop, xz);
      }
    }
  } else if (lv.kind == kDampedJacobi || lv.kind == kBlockJacobi) {
    TM *dinv = lv.dinv.template as<TM>(), *tmp = lv.tmp.template as<TM>();
    int it0 = 0;
    if (first_done && iters > 0) {
      x_ready = true;
      it0 = 1;
    } else if (!x_ready && iters > 0) {
      // First sweep from x = 0 in closed form (k_jacobi_zero).
      if (lv.kind == kDampedJacobi)
        gsk::k_jacobi_zero<TM, 1><<<grid_for(n), kBlock, 0, stream_>>>(dinv, lv.rhs, lv.x, n,
                                                                       lv.omega, stop);
      else
        gsk::k_jacobi_zero<TM, 3><<<grid_for(n / 3), kBlock, 0, stream_>>>(dinv, lv.rhs, lv.x,
                                                                           n / 3, lv.omega, stop);
      x_ready = true;
      it0 = 1;
    }
    // Sweeps alternate between x and tmp (each reads one and writes the
    // other); a result left in tmp is copied back unless may_leave.
    TM *cur = lv.x, *oth = tmp;
    for (int it = it0; it < iters; ++it) {
      if (lv.A.G) {
        with_group(lv.A.G, [&](auto g) {
          constexpr int G = decltype(g)::value;
          if (lv.kind == kDampedJacobi && !lv.A.b3)
            gsk::k_jacobi_w<TM, G><<<gsk::grid_rows<G>(n), kBlock, 0, stream_>>>(
                lv.A.w.view(), dinv, lv.rhs, cur, oth, lv.omega, stop);
          else if (lv.kind == kBlockJacobi && lv.A.b3)
            gsk::k_block_jacobi_bw<TM, G><<<gsk::grid_rows<G>(n / 3), kBlock, 0, stream_>>>(
                lv.A.bw.view(), dinv, lv.rhs, cur, oth, lv.omega, stop);
          else
            throw std::runtime_error("Jacobi smoother: operator storage does not match the smoother.");
        });
      } else if (lv.kind == kDampedJacobi) {
        gsk::k_jacobi<TM><<<grid_for(n), kBlock, 0, stream_>>>(lv.A.s.view(), dinv, lv.rhs, cur,
                                                               oth, lv.omega, stop);
      } else if (lv.A.b3) {
        gsk::k_block_jacobi_b3<TM><<<grid_for(n / 3), kBlock, 0, stream_>>>(
            lv.A.b.view(), dinv, lv.rhs, cur, oth, lv.omega, stop);
      } else {
        gsk::k_block_jacobi<TM><<<grid_for(n / 3), kBlock, 0, stream_>>>(
            lv.A.s.view(), dinv, lv.rhs, cur, oth, lv.omega, stop);
      }
      std::swap(cur, oth);
    }
    if (cur != lv.x && !may_leave) {
      gsk::k_copy<TM><<<grid_for(n), kBlock, 0, stream_>>>(cur, lv.x, n, stop);
      cur = lv.x;
    }
    lv.xout = cur;
  }
  if (!x_ready && !x_zeroed) launch_fill<TM>(lv.x, TM(0), n, stream_, stop);
}

template <class TC, class TM>
void Engine<TC, TM>::encode_vcycle(int k, const int32_t *stop, bool zero_x, bool x_zeroed,
                                   bool first_done) {
  if (k == tail_k0_) {
    if (!zero_x) throw std::logic_error("V-cycle tail entered with a nonzero start.");
    encode_tail(stop);
    return;
  }
  Level<TM> &lv = L_[k];
  if (k == nlev_ - 1) {
    if (lv.direct) {
      double *prow = lv.sym_prow.template as<double>(), *pcol = lv.sym_pcol.template as<double>();
      gsk::k_dense_sym_strip<TM><<<static_cast<uint32_t>(lv.sym_nstrips), kBlock, 0, stream_>>>(
          lv.inv.template as<double>(), lv.sym_strips.template as<int32_t>(), lv.n, lv.sym_nt,
          lv.sym_kmax, lv.rhs, prow, pcol, stop);
      gsk::k_dense_sym_reduce<TM><<<static_cast<uint32_t>(lv.sym_nt), 1024, 0, stream_>>>(
          prow, pcol, lv.n, lv.sym_nt, lv.sym_kmax, lv.x, stop);
    } else {
      encode_smooth(lv, opt_.coarse_steps, zero_x, stop, x_zeroed, k > 0, first_done);
    }
    return;
  }
  encode_smooth(lv, lv.pre, zero_x, stop, x_zeroed, false, first_done);
  // With no pre-smoothing from x = 0 the residual is rhs bit for bit (the
  // row sums of A 0 are +0, and rhs - (+0) == rhs), so restrict rhs itself.
  const bool res_is_rhs = zero_x && lv.pre == 0;
  if (!res_is_rhs) encode_residual(lv, stop);
  const TM *res = res_is_rhs ? lv.rhs : lv.res.template as<TM>();
  Level<TM> &cl = L_[k + 1];
  // The coarse correction starts at zero; the coarse level's first kernel
  // zeroes it itself when it is a zero-start Gauss-Seidel colour.
  const bool coarse_self_zero =
      !cl.direct && ((k + 1 == nlev_ - 1) ? opt_.coarse_steps > 0 : cl.pre > 0) &&
      (cl.kind == kSGS || cl.kind == kBlockGS || cl.kind == kDampedJacobi ||
       cl.kind == kBlockJacobi);
  // A damped / block Jacobi coarse level that smooths from zero gets its
  // first sweep from the restriction (same bits, one launch fewer), when
  // the transfer's rows match the smoother's (scalar rows for damped
  // Jacobi, 3x3 block rows for block Jacobi).
  const bool cl_smooths = !cl.direct && ((k + 1 == nlev_ - 1) ? opt_.coarse_steps > 0 : cl.pre > 0);
  const bool fuse_first = cl_smooths && ((cl.kind == kDampedJacobi && !lv.R.b3) ||
                                         (cl.kind == kBlockJacobi && lv.R.b3 && cl.A.b3));
  TM *corr_c = fuse_first ? cl.x : coarse_self_zero ? nullptr : cl.x;
  encode_restrict(k, res, corr_c, stop, fuse_first ? cl.dinv.template as<TM>() : nullptr,
                  cl.omega);
  encode_vcycle(k + 1, stop, true, corr_c != nullptr && !fuse_first, fuse_first);
  encode_prolong(k, opt_.reference_vcycle && k > 0, stop);
  encode_smooth(lv, lv.post, false, stop, false, k > 0);
}

// Levels tail_k0_ .. nlev_ - 1 in one single-workgroup launch, from x = 0
// (the state encode_vcycle(tail_k0_, ...) starts from below the finest).
template <class TC, class TM>
void Engine<TC, TM>::encode_tail(const int32_t *stop) {
  gsk::TailArgs<TM> a{};
  a.nl = nlev_ - tail_k0_;
  a.coarse_steps = opt_.coarse_steps;
  // The prolongation into levels below the finest (every tail level).
  a.alias = opt_.reference_vcycle ? 1 : 0;
  int64_t off = 0;
  for (int j = 0; j < a.nl; ++j) {
    Level<TM> &lv = L_[tail_k0_ + j];
    gsk::TailLevel<TM> &t = a.L[j];
    t.A = lv.tA.view();
    if (j + 1 < a.nl) {
      t.R = lv.tR.view();
      t.P = lv.tP.view();
    }
    t.diag = lv.diag.template as<TM>();
    t.coff = lv.tcoff.template as<int64_t>();
    t.rhs = lv.rhs;
    t.x = lv.x;
    t.n = lv.n;
    t.lds_off = off;
    off += gsk::tail_level_bytes<TM>(lv.n);
    t.ncolors = lv.ncolors;
    t.pre = lv.pre;
    t.post = lv.post;
    t.omega = lv.omega;
  }
  a.lp_off = off;
  if (off + int64_t(gsk::kTailBlock) * gsk::kTailKS * static_cast<int64_t>(sizeof(TM)) != tail_lds_)
    throw std::logic_error("V-cycle tail: LDS layout does not match setup.");
  gsk::k_vcycle_tail<TM><<<1, gsk::kTailBlock, static_cast<size_t>(tail_lds_), stream_>>>(a, stop);
}

// rhs_{k+1} = R_k res (and corr_c = 0 when given).
template <class TC, class TM>
void Engine<TC, TM>::encode_restrict(int k, const TM *res, TM *corr_c, const int32_t *stop,
                                     const TM *dinv_c, TM omega_c) {
  Level<TM> &lv = L_[k];
  Level<TM> &cl = L_[k + 1];
  if (lv.R.b3) {
    if (lv.R.G)
      with_group(lv.R.G, [&](auto g) {
        with_bk(lv.R.K, [&](auto kk) {
          constexpr int G = decltype(g)::value, K = decltype(kk)::value;
          gsk::k_restrict_bw<TM, G, K><<<gsk::grid_rows<G>(cl.n / 3), kBlock, 0, stream_>>>(
              lv.R.bw.view(), res, cl.rhs, corr_c, stop, dinv_c, omega_c);
        });
      });
    else
      gsk::k_restrict_b3<TM><<<grid_for(cl.n / 3), kBlock, 0, stream_>>>(
          lv.R.sb.view(), res, cl.rhs, corr_c, stop, dinv_c, omega_c);
    return;
  }
  if (lv.R.G)
    with_group(lv.R.G, [&](auto g) {
      with_k(lv.R.K, [&](auto kk) {
        constexpr int G = decltype(g)::value, K = decltype(kk)::value;
        gsk::k_restrict_w<TM, G, K><<<gsk::grid_rows<G>(cl.n), kBlock, 0, stream_>>>(
            lv.R.w.view(), res, cl.rhs, corr_c, stop, dinv_c, omega_c);
      });
    });
  else
    gsk::k_restrict<TM><<<grid_for(cl.n), kBlock, 0, stream_>>>(lv.R.s.view(), res, cl.rhs,
                                                                corr_c, stop, dinv_c, omega_c);
}

// x_k += P_k x_{k+1} (alias: the reference's coarse-level x = 2 P e).
template <class TC, class TM>
void Engine<TC, TM>::encode_prolong(int k, bool alias, const int32_t *stop) {
  Level<TM> &lv = L_[k];
  Level<TM> &cl = L_[k + 1];
  // The coarse level's correction where its last smoothing left it.
  const TM *e = cl.xout ? cl.xout : cl.x;
  if (lv.P.b3) {
    if (lv.P.G)
      with_group(lv.P.G, [&](auto g) {
        with_bk(lv.P.K, [&](auto kk) {
          constexpr int G = decltype(g)::value, K = decltype(kk)::value;
          gsk::k_prolong_bw<TM, G, K><<<gsk::grid_rows<G>(lv.n / 3), kBlock, 0, stream_>>>(
              lv.P.bw.view(), e, lv.x, alias ? 1 : 0, stop);
        });
      });
    else
      gsk::k_prolong_b3<TM><<<grid_for(lv.n / 3), kBlock, 0, stream_>>>(
          lv.P.sb.view(), e, lv.x, alias ? 1 : 0, stop);
    return;
  }
  if (lv.P.G)
    with_group(lv.P.G, [&](auto g) {
      with_k(lv.P.K, [&](auto kk) {
        constexpr int G = decltype(g)::value, K = decltype(kk)::value;
        gsk::k_prolong_w<TM, G, K><<<gsk::grid_rows<G>(lv.n), kBlock, 0, stream_>>>(
            lv.P.w.view(), e, lv.x, alias ? 1 : 0, stop);
      });
    });
  else if (alias)
    gsk::k_prolong_alias<TM><<<grid_for(lv.n), kBlock, 0, stream_>>>(lv.P.s.view(), e, lv.x, stop);
  else
    gsk::k_prolong_add<TM><<<grid_for(lv.n), kBlock, 0, stream_>>>(lv.P.s.view(), e, lv.x, stop);
}

template <class TC, class TM>
void Engine<TC, TM>::apply_preconditioner(const int32_t *stop) {
  if (opt_.precond == kPrecIdentity) {
    gsk::k_copy<TM><<<grid_for(n_), kBlock, 0, stream_>>>(rvc_, z_.template as<TM>(), n_, stop);
    return;
  }
  for (int v = 0; v < opt_.vcycles; ++v) encode_vcycle(0, stop, v == 0);
  if (opt_.vcycles == 0) launch_fill<TM>(z_.template as<TM>(), TM(0), n_, stream_, stop);
}

template <class TC, class TM>
void Engine<TC, TM>::encode_iteration(const int32_t *stop, bool with_replace) {
  auto *ctl = ctl_.template as<gsk::Ctl<TC>>();
  TC *parts = partials_.template as<TC>();
  const DOp<TM> &Av = L_[0].A;
  const bool vb3 = cg_vop_ && Av.b3;
  const uint32_t g = (!cg_vop_ && cg_pairs_) ? static_cast<uint32_t>(opt_.persist_grid)
                                              : grid_for((cg_vop_ ? vb3 : Acg_.b3) ? n_ / 3 : n_);
  if (!cg_vop_ && cg_pairs_)
    gsk::k_cg_spmv_pap_b3_persist<TC, true><<<g, kBlock, 0, stream_>>>(
        Acg_.b.view(), p_.template as<TC>(), q_.template as<TC>(), ctl, parts);
  else if (cg_vop_ && vb3)
    gsk::k_cg_spmv_pap_b3<TC, TM><<<g, kBlock, 0, stream_>>>(Av.b.view(), p_.template as<TC>(),
                                                              q_.template as<TC>(), ctl, parts);
  else if (cg_vop_)
    gsk::k_cg_spmv_pap<TC, TM><<<g, kBlock, 0, stream_>>>(Av.s.view(), p_.template as<TC>(),
                                                           q_.template as<TC>(), ctl, parts);
  else if (Acg_.b3)
    gsk::k_cg_spmv_pap_b3<TC><<<g, kBlock, 0, stream_>>>(Acg_.b.view(), p_.template as<TC>(), q_.template as<TC>(),
                                                          ctl, parts);
  else
    gsk::k_cg_spmv_pap<TC><<<g, kBlock, 0, stream_>>>(Acg_.s.view(), p_.template as<TC>(), q_.template as<TC>(),
                                                       ctl, parts);
  gsk::k_cg_alpha<TC><<<1, kBlock, 0, stream_>>>(ctl, parts, g);
  TM *rm = std::is_same<TC, TM>::value ? nullptr : rm_.template as<TM>();
  const uint32_t gn = grid_for(n_);
  gsk::k_cg_update_xr<TC, TM><<<gn, kBlock, 0, stream_>>>(
      x_.template as<TC>(), r_.template as<TC>(), rm, p_.template as<TC>(), q_.template as<TC>(), n_, ctl, parts);
  gsk::k_cg_rr_final<TC><<<1, kBlock, 0, stream_>>>(ctl, parts, gn);
  if (with_replace && replace_every() > 0) {
    // Replacement (skipped on the device unless due or a convergence check
    // is pending): r = b - A x in FP64.
    const uint32_t gr = grid_for(Acg_.b3 ? n_ / 3 : n_);
    if (cg_pairs_)
      gsk::k_cg_replace_b3<TC, TM, true><<<gr, kBlock, 0, stream_>>>(
          Acg_.b.view(), b_.template as<TC>(), x_.template as<TC>(), r_.template as<TC>(), rm, ctl, parts);
    else if (Acg_.b3)
      gsk::k_cg_replace_b3<TC, TM><<<gr, kBlock, 0, stream_>>>(
          Acg_.b.view(), b_.template as<TC>(), x_.template as<TC>(), r_.template as<TC>(), rm, ctl, parts);
    else
      gsk::k_cg_replace<TC, TM><<<gr, kBlock, 0, stream_>>>(
          Acg_.s.view(), b_.template as<TC>(), x_.template as<TC>(), r_.template as<TC>(), rm, ctl, parts);
    gsk::k_cg_replace_final<TC><<<1, kBlock, 0, stream_>>>(ctl, parts, gr);
  }
  apply_preconditioner(stop);
  gsk::k_cg_rz<TC, TM><<<gn, kBlock, 0, stream_>>>(r_.template as<TC>(), z_.template as<TM>(), n_, ctl,
                                                   parts);
  gsk::k_cg_beta<TC><<<1, kBlock, 0, stream_>>>(ctl, parts, gn);
  gsk::k_cg_update_p<TC, TM><<<grid_for(n_), kBlock, 0, stream_>>>(p_.template as<TC>(), z_.template as<TM>(), n_,
                                                                   ctl);
}

template <class TC, class TM>
void Engine<TC, TM>::capture_graph() {
  if (graph_exec_) return;
  auto *ctl = ctl_.template as<gsk::Ctl<TC>>();
  const int32_t *stop = &ctl->stop;
  GS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  // The replacement pair sits in the last iteration of the graph (in every
  // one when the period is shorter): a due replacement waits there, and a
  // convergence check pending from an earlier iteration (whose remaining
  // kernels skipped) is resolved there, where a failed check lets that
  // iteration's preconditioner and direction update run.
  const int64_t every = replace_every();
  for (int i = 0; i < opt_.graph_iters; ++i)
    encode_iteration(stop, every > 0 && (every < opt_.graph_iters || i == opt_.graph_iters - 1));
  GS_HIP(hipStreamEndCapture(stream_, &graph_));
  size_t nodes = 0;
  GS_HIP(hipGraphGetNodes(graph_, nullptr, &nodes));
  graph_nodes_per_iter_ = static_cast<double>(nodes) / opt_.graph_iters;
  GS_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  const int gi = opt_.graph_iters;
  if (opt_.graph_tail && gi > 1 && (every == 0 || every < gi || every % gi == 0)) {
    // The single iteration is the graph's last, so it carries the pair.
    // With a period below graph_iters the chunk carries it in every
    // iteration as well. With a period that is a multiple of graph_iters,
    // no replacement falls due in a tail (a tail never reaches the next
    // multiple of graph_iters), and a pending convergence check is resolved
    // in the iteration that set it; the chunk resolves it in its last
    // iteration, after idle ones, from the same state. Either way the
    // results are the chunk graph's bit for bit. (Other periods keep whole
    // chunks: the chunk would defer a due replacement to its end.)
    GS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    encode_iteration(stop, every > 0);
    GS_HIP(hipStreamEndCapture(stream_, &graph1_));
    GS_HIP(hipGraphInstantiate(&graph1_exec_, graph1_, nullptr, nullptr, 0));
  }
}

template <class TC, class TM>
SolveResult Engine<TC, TM>::solve(const double *b, double *x, double tol, int64_t maxit,
                                  const LagGuard *guard) {
  if (!setup_done_) throw std::runtime_error("CGSolver: solver is not set up.");
  SolveResult res;
  // src/CGSolver.cc:113-118: a right-hand side below 1e-12 in norm returns
  // x = 0 without iterating.
  const double nb = gsx::sum_squares(b, n_);
  if (std::sqrt(nb) < 1e-12) {
    for (int64_t i = 0; i < n_; ++i) x[i] = 0.0;
    res.iterations = 0;
    res.zero_rhs = true;
    return res;
  }
  // Gather into pinned staging (internal order), one DMA each way.
  TC *bi = stage_, *xi = stage_ + n_;
  gsx::gather_cast(b, int2orig_.data(), n_, bi);
  gsx::gather_cast(x, int2orig_.data(), n_, xi);
  GS_HIP(hipMemcpyAsync(b_.get(), bi, n_ * sizeof(TC), hipMemcpyHostToDevice, stream_));
  GS_HIP(hipMemcpyAsync(x_.get(), xi, n_ * sizeof(TC), hipMemcpyHostToDevice, stream_));
  if (hist_cap_ < maxit + 1) {
    hist_.alloc((maxit + 1) * sizeof(TC));
    hist_cap_ = maxit + 1;
  }
  gsk::Ctl<TC> c0{};
  c0.tol2 = static_cast<TC>(tol) * static_cast<TC>(tol);
  c0.maxit = maxit;
  c0.hist = hist_.template as<TC>();
  c0.replace_every = replace_every();
  GS_HIP(hipMemcpyAsync(ctl_.get(), &c0, sizeof(c0), hipMemcpyHostToDevice, stream_));
  GS_HIP(hipStreamSynchronize(stream_));

  GS_HIP(hipEventRecord(ev0_, stream_));
  auto t0 = std::chrono::steady_clock::now();
  auto *ctl = ctl_.template as<gsk::Ctl<TC>>();
  TM *rm = std::is_same<TC, TM>::value ? nullptr : rm_.template as<TM>();
  if (cg_pairs_)
    gsk::k_cg_init_r_b3<TC, TM, true><<<grid_for(n_ / 3), kBlock, 0, stream_>>>(
        Acg_.b.view(), b_.template as<TC>(), x_.template as<TC>(), r_.template as<TC>(), rm);
  else if (Acg_.b3)
    gsk::k_cg_init_r_b3<TC, TM><<<grid_for(n_ / 3), kBlock, 0, stream_>>>(
        Acg_.b.view(), b_.template as<TC>(), x_.template as<TC>(), r_.template as<TC>(), rm);
  else
    gsk::k_cg_init_r<TC, TM><<<grid_for(n_), kBlock, 0, stream_>>>(
        Acg_.s.view(), b_.template as<TC>(), x_.template as<TC>(), r_.template as<TC>(), rm);
  apply_preconditioner(nullptr);
  gsk::k_cg_init_dots<TC, TM><<<grid_for(n_), kBlock, 0, stream_>>>(
      b_.template as<TC>(), r_.template as<TC>(), z_.template as<TM>(), p_.template as<TC>(), n_, ctl, partials_.template as<TC>());
  gsk::k_cg_init_final<TC><<<1, kBlock, 0, stream_>>>(ctl, partials_.template as<TC>(), grid_for(n_));
  const int32_t *stop = &ctl->stop;
  int64_t launches = 0;  // each runs at least one iteration while stop == 0
  double log_rel0 = 0.0;
  double tail_rel = 0.0;  // relres^2 and iteration at the previous host check
  int64_t tail_it = -1;
  if (maxit > 0) {
    if (opt_.graph_iters > 0) capture_graph();
    for (;;) {
      GS_HIP(hipMemcpyAsync(host_ctl_, ctl_.get(), sizeof(gsk::Ctl<TC>), hipMemcpyDeviceToHost,
                            stream_));
      GS_HIP(hipStreamSynchronize(stream_));
      if (host_ctl_->stop == 5)
        throw std::logic_error("CGSolver: convergence check left pending.");
      // The device's iteration count, not the launches: a graph launch in
      // which a convergence check failed runs fewer than graph_iters.
      if (host_ctl_->stop != 0 || host_ctl_->iter >= maxit || launches > maxit) break;
      if (guard) {
        // The first check sees the initial residual; later ones compare the
        // reduction so far with the guard's rate stretched by its lag.
        const double rel = static_cast<double>(host_ctl_->rel);
        const int64_t it = host_ctl_->iter;
        if (launches == 0) {
          log_rel0 = std::log(rel);
        } else if (it >= guard->min_iter && rel > 0.0 &&
                   std::log(rel) - log_rel0 > guard->log_rate * static_cast<double>(it) / guard->lag) {
          res.lagging = true;
          break;
        }
      }
      // Iterations in this launch: a whole chunk, or, when the rate of the
      // last chunk says fewer remain to the tolerance, that many single
      // iterations (a short estimate costs one more host check; a long one
      // runs no-op iterations, as a whole chunk would).
      int64_t k = opt_.graph_iters;
      {
        const double rel = static_cast<double>(host_ctl_->rel);
        const int64_t it = host_ctl_->iter;
        if (graph1_exec_ && tail_it < 0 && last_loop_iters_ > 0 && last_loop_iters_ < k) {
          k = last_loop_iters_;
        } else if (graph1_exec_ && tail_it >= 0 && it > tail_it && rel > 0.0 && tail_rel > 0.0 &&
                   host_ctl_->tol2 > 0) {
          const double rate = (std::log(rel) - std::log(tail_rel)) / static_cast<double>(it - tail_it);
          const double need = std::log(static_cast<double>(host_ctl_->tol2)) - std::log(rel);
          if (rate < 0.0 && need < 0.0) {
            const double est = std::ceil(need / rate);
            if (est < static_cast<double>(k)) k = std::max<int64_t>(1, static_cast<int64_t>(est));
          }
        }
        tail_rel = rel;
        tail_it = it;
      }
      if (graph_exec_ && k >= opt_.graph_iters) {
        GS_HIP(hipGraphLaunch(graph_exec_, stream_));
      } else if (graph_exec_) {
        for (int64_t i = 0; i < k; ++i) GS_HIP(hipGraphLaunch(graph1_exec_, stream_));
      } else {
        encode_iteration(stop);
        k = 1;
      }
      res.launched_iterations += k;
      ++launches;
    }
  } else {
    GS_HIP(hipMemcpyAsync(host_ctl_, ctl_.get(), sizeof(gsk::Ctl<TC>), hipMemcpyDeviceToHost,
                          stream_));
  }
  GS_HIP(hipEventRecord(ev1_, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
  auto t1 = std::chrono::steady_clock::now();
  float ms = 0.f;
  GS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  res.solve_ms = ms;
  res.wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  const gsk::Ctl<TC> &hc = *host_ctl_;
  if (hc.stop == 3) {
    if (hc.iter == 0)
      throw NonFiniteResidual("CGSolver: initial residual is NaN/Inf - check matrix and RHS.");
    throw NonFiniteResidual("CGSolver: solver diverged (NaN/Inf residual).");
  }
  res.iterations = (maxit == 0) ? 0 : hc.iters_out;
  res.loop_iterations = hc.iter;
  res.relres2 = static_cast<double>(hc.rel);
  res.converged = hc.stop == 1;
  if (graph1_exec_) last_loop_iters_ = res.converged ? hc.iter : 0;
  res.breakdown = hc.breakdown;
  // History.
  const int64_t hl = std::min<int64_t>(hc.iter + 1, maxit + 1);
  std::vector<TC> h(hl);
  GS_HIP(hipMemcpy(h.data(), hist_.get(), hl * sizeof(TC), hipMemcpyDeviceToHost));
  hist_host_.assign(h.begin(), h.end());
  // Solution back to the caller's ordering.
  GS_HIP(hipMemcpyAsync(xi, x_.get(), n_ * sizeof(TC), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
  gsx::scatter_cast(xi, int2orig_.data(), n_, x);
  // True residual of the CG operator in FP64 arithmetic.
  const uint32_t gt = grid_for(Acg_.b3 ? n_ / 3 : n_);
  if (cg_pairs_)
    gsk::k_true_residual_b3<TC, true><<<gt, kBlock, 0, stream_>>>(
        Acg_.b.view(), b_.template as<TC>(), x_.template as<TC>(), partials_.template as<double>());
  else if (Acg_.b3)
    gsk::k_true_residual_b3<TC><<<gt, kBlock, 0, stream_>>>(
        Acg_.b.view(), b_.template as<TC>(), x_.template as<TC>(), partials_.template as<double>());
  else
    gsk::k_true_residual<TC><<<gt, kBlock, 0, stream_>>>(
        Acg_.s.view(), b_.template as<TC>(), x_.template as<TC>(), partials_.template as<double>());
  gsk::k_true_final<double><<<1, kBlock, 0, stream_>>>(partials_.template as<double>(), gt,
                                                        tr_.template as<double>());
  double tr[2];
  GS_HIP(hipMemcpyAsync(tr, tr_.get(), sizeof(tr), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
  res.true_relres = std::sqrt(tr[0] / tr[1]);
  return res;
}

// z = M r for r given in level-0 (permuted) ordering; one preconditioner
// application exactly as inside the loop.
template <class TC, class TM>
void Engine<TC, TM>::vcycle_level_order(const double *r, double *z) {
  if (!setup_done_) throw std::runtime_error("solver is not set up.");
  std::vector<TM> ri(n_), zi(n_);
  for (int64_t i = 0; i < n_; ++i) ri[i] = static_cast<TM>(r[int2lvl0_[i]]);
  GS_HIP(hipMemcpy(rvc_, ri.data(), n_ * sizeof(TM), hipMemcpyHostToDevice));
  apply_preconditioner(nullptr);
  GS_HIP(hipStreamSynchronize(stream_));
  GS_HIP(hipMemcpy(zi.data(), z_.get(), n_ * sizeof(TM), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n_; ++i) z[int2lvl0_[i]] = static_cast<double>(zi[i]);
}

// y = A x with the CG operator, vectors in level-0 (permuted) ordering.
template <class TC, class TM>
void Engine<TC, TM>::spmv_level_order(const double *x, double *y) {
  if (!setup_done_) throw std::runtime_error("solver is not set up.");
  std::vector<TC> xi(n_), yi(n_);
  for (int64_t i = 0; i < n_; ++i) xi[i] = static_cast<TC>(x[int2lvl0_[i]]);
  GS_HIP(hipMemcpy(p_.get(), xi.data(), n_ * sizeof(TC), hipMemcpyHostToDevice));
  if (cg_pairs_)
    gsk::k_spmv_b3_persist<TC, true><<<static_cast<uint32_t>(opt_.persist_grid), kBlock, 0, stream_>>>(
        Acg_.b.view(), p_.template as<TC>(), q_.template as<TC>());
  else if (Acg_.b3)
    gsk::k_spmv_b3<TC><<<grid_for(n_ / 3), kBlock, 0, stream_>>>(Acg_.b.view(), p_.template as<TC>(),
                                                                 q_.template as<TC>(), nullptr);
  else
    gsk::k_spmv<TC><<<grid_for(n_), kBlock, 0, stream_>>>(Acg_.s.view(), p_.template as<TC>(), q_.template as<TC>(),
                                                          nullptr);
  GS_HIP(hipStreamSynchronize(stream_));
  GS_HIP(hipMemcpy(yi.data(), q_.get(), n_ * sizeof(TC), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n_; ++i) y[int2lvl0_[i]] = static_cast<double>(yi[i]);
}

// Average duration (ms) of one launch of a hot kernel, timed with HIP events
// on the solver's own stream over `reps` back-to-back launches. which: 0 =
// CG SpMV (y = A p), 1 = one V-cycle, 2 = one full PCG iteration, 3 / 4 =
// one smoothing sweep of level 0 / 1 from a nonzero x (Gauss-Seidel: the
// symmetric forward + backward sweep; Jacobi: the sweep and its copy back).
// *bytes receives the algorithmic bytes of one launch (0 when not defined);
// for a sweep, the sum over its colour launches of the colour's matrix
// entries and its rows' vector traffic (rhs, x read and write, diagonal or
// inverse block).
template <class TC, class TM>
double Engine<TC, TM>::time_kernel(int which, int reps, double *bytes) {
  if (!setup_done_) throw std::runtime_error("solver is not set up.");
  *bytes = 0.0;
  auto body = [&]() {
    if (which == 0) {
      if (cg_pairs_)
        gsk::k_spmv_b3_persist<TC, true><<<static_cast<uint32_t>(opt_.persist_grid), kBlock, 0, stream_>>>(
            Acg_.b.view(), p_.template as<TC>(), q_.template as<TC>());
      else if (Acg_.b3)
        gsk::k_spmv_b3<TC><<<grid_for(n_ / 3), kBlock, 0, stream_>>>(Acg_.b.view(), p_.template as<TC>(),
                                                                     q_.template as<TC>(), nullptr);
      else
        gsk::k_spmv<TC><<<grid_for(n_), kBlock, 0, stream_>>>(Acg_.s.view(), p_.template as<TC>(),
                                                              q_.template as<TC>(), nullptr);
    } else if (which == 1) {
      apply_preconditioner(nullptr);
    } else if (which == 2) {
      encode_iteration(nullptr);
    } else if (which == 5) {
      encode_restrict(0, L_[0].res.template as<TM>(), L_[1].x, nullptr);
    } else if (which == 6) {
      encode_prolong(0, false, nullptr);
    } else if (which == 7) {
      encode_residual(L_[0], nullptr);
    } else if (which == 8) {
      encode_tail(nullptr);
    } else {
      encode_smooth(L_[which - 3], 1, false, nullptr);
    }
  };
  if (which < 0 || which > 8) throw std::runtime_error("time_kernel: unknown kernel id.");
  if (which == 8 && tail_k0_ < 0) throw std::runtime_error("time_kernel: no V-cycle tail.");
  if (which == 7 && (opt_.precond != kPrecMultigrid || nlev_ < 2 || L_[0].direct))
    throw std::runtime_error("time_kernel: no level-0 residual.");
  if ((which == 5 || which == 6) && (opt_.precond != kPrecMultigrid || nlev_ < 2))
    throw std::runtime_error("time_kernel: no level-0 transfer.");
  if ((which == 3 || which == 4) && (which - 3 >= nlev_ || L_[which - 3].direct))
    throw std::runtime_error("time_kernel: that level has no smoother sweep.");
  if (which == 0) *bytes = Acg_.spmv_bytes();
  // Level-0 residual: the operator's SpMV bytes plus the right-hand side read.
  if (which == 7) *bytes = L_[0].A.spmv_bytes() + L_[0].n * static_cast<double>(sizeof(TM));
  if (which == 5 || which == 6) {
    // Entries (value + column) + the input vector once + the output vector
    // (restriction: rhs_c and the zeroed correction written; prolongation:
    // x read and written).
    const double s = sizeof(TM);
    const auto &T = which == 5 ? L_[0].R : L_[0].P;
    const double entries = T.b3 ? T.nblocks * (9 * s + 4) : T.nnz * (s + 4);
    *bytes = entries + (which == 5 ? L_[0].n * s + 2.0 * L_[1].n * s
                                   : L_[1].n * s + 2.0 * L_[0].n * s);
  }
  if ((which == 3 || which == 4) &&
      (L_[which - 3].kind == kDampedJacobi || L_[which - 3].kind == kBlockJacobi)) {
    // One Jacobi sweep into the second buffer (entries, rhs, x, the new x,
    // the inverse diagonal or its 3x3 blocks) and the copy back.
    const double s = sizeof(TM);
    const auto &lv = L_[which - 3];
    const double n = static_cast<double>(lv.n);
    const double entries = lv.A.b3 ? lv.A.nblocks * (9 * s + 4) : lv.A.nnz * (s + 4);
    *bytes = entries + 3 * n * s + (lv.kind == kBlockJacobi ? 3 * n * s : n * s) + 2 * n * s;
  } else if (which == 3 || which == 4) {
    const double s = sizeof(TM);
    const auto &lv = L_[which - 3];
    const bool blk = lv.kind == kBlockGS;
    double by = 0.0;
    for (const auto &step : sweep_sequence(lv.ncolors, 1)) {
      const int c = step.first;
      const double rows = static_cast<double>(lv.coff[c + 1] - lv.coff[c]);
      by += blk ? lv.cunits[c] * (9 * s + 4) + rows * 18 * s  // rhs 3, x 6, D^-1 9
                : lv.cunits[c] * (s + 4) + rows * 4 * s;      // rhs, x 2, diagonal
    }
    *bytes = by;
  }
  // Keep the iteration kernels live: a zeroed control block with stop = 0.
  if (which == 2) {
    gsk::Ctl<TC> c0{};
    c0.tol2 = TC(0);
    c0.maxit = int64_t(1) << 40;
    c0.rho = TC(1);
    c0.bb = TC(1);
    c0.hist = nullptr;
    c0.replace_every = replace_every();
    GS_HIP(hipMemcpyAsync(ctl_.get(), &c0, sizeof(c0), hipMemcpyHostToDevice, stream_));
  }
  body();  // warm-up
  GS_HIP(hipStreamSynchronize(stream_));
  GS_HIP(hipEventRecord(ev0_, stream_));
  for (int i = 0; i < reps; ++i) body();
  GS_HIP(hipEventRecord(ev1_, stream_));
  GS_HIP(hipEventSynchronize(ev1_));
  float ms = 0.f;
  GS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  return ms / reps;
}

template <class TC, class TM>
void Engine<TC, TM>::level_info(std::vector<LevelInfo> &out) const {
  out.clear();
  for (int k = 0; k < nlev_ && k < static_cast<int>(L_.size()); ++k) {
    const auto &lv = L_[k];
    LevelInfo li;
    li.n = lv.n;
    li.nnz = lv.A.nnz;
    li.blocks3 = lv.A.nblocks;
    li.group = lv.A.G;
    li.r_group = k + 1 < nlev_ ? lv.R.G : 0;
    li.p_group = k + 1 < nlev_ ? lv.P.G : 0;
    li.transfer_b3 = k + 1 < nlev_ && lv.P.b3;
    li.ncolors = lv.ncolors;
    for (int c = 0; c < lv.ncolors; ++c)
      li.max_color_rows = std::max<int64_t>(li.max_color_rows, lv.coff[c + 1] - lv.coff[c]);
    li.kind = lv.kind;
    li.direct = lv.direct;
    li.omega = static_cast<double>(lv.omega);
    li.pre = lv.pre;
    li.post = lv.post;
    out.push_back(li);
  }
}

template <class TC, class TM>
std::string Engine<TC, TM>::describe() const {
  std::ostringstream os;
  os << "{\"n\": " << n_ << ", \"levels\": [";
  std::vector<LevelInfo> li;
  level_info(li);
  for (size_t k = 0; k < li.size(); ++k) {
    if (k) os << ", ";
    os << "{\"n\": " << li[k].n << ", \"nnz\": " << li[k].nnz << ", \"blocks3\": " << li[k].blocks3
       << ", \"colors\": " << li[k].ncolors << ", \"max_color_rows\": " << li[k].max_color_rows
       << ", \"kind\": " << li[k].kind
       << ", \"direct\": " << (li[k].direct ? "true" : "false") << ", \"omega\": " << li[k].omega
       << ", \"pre\": " << li[k].pre << ", \"post\": " << li[k].post
       << ", \"group\": " << li[k].group << ", \"r_group\": " << li[k].r_group
       << ", \"p_group\": " << li[k].p_group
       << ", \"transfer_b3\": " << (li[k].transfer_b3 ? "true" : "false") << "}";
  }
  os << "], \"setup_ms\": {";
  for (size_t i = 0; i < setup_ms_.size(); ++i)
    os << (i ? ", " : "") << "\"" << setup_ms_[i].first << "\": " << setup_ms_[i].second;
  os << "}, \"tail_level\": " << tail_k0_ << ", \"tail_lds_bytes\": " << tail_lds_
     << ", \"cg_vcycle_operator\": " << (cg_vop_ ? "true" : "false")
     << ", \"replace_every\": " << replace_every()
     << ", \"cg_pair_layout\": " << (cg_pairs_ ? "true" : "false")
     << ", \"cg_block3\": " << (Acg_.b3 ? "true" : "false")
     << ", \"cg_spmv_bytes\": " << Acg_.spmv_bytes() << ", \"graph_iters\": " << opt_.graph_iters
     << ", \"launches_per_iteration\": " << graph_nodes_per_iter_ << "}";
  return os.str();
}

std::unique_ptr<EngineBase> make_engine(int precision, const Options &o) {
  switch (precision) {
    case kPrecDouble: return std::unique_ptr<EngineBase>(new Engine<double, double>(o));
    case kPrecMixed: return std::unique_ptr<EngineBase>(new Engine<double, float>(o));
    case kPrecSingle: return std::unique_ptr<EngineBase>(new Engine<float, float>(o));
    default: throw std::runtime_error("unknown precision mode.");
  }
}

}  // namespace gse
