// Group-per-row kernels for operators with long rows (the Galerkin coarse
// levels and the restriction operators of an AMG hierarchy).
//
// A thread-per-row kernel over a row of w entries waits on w dependent
// (index -> x) load chains one after another, and a coarse-level colour
// holds too few rows to hide that latency. Here a group of G lanes (G
// divides 64, so a group never straddles a wavefront) owns one row: lane j
// loads entries j, j + G, ... of the row in parallel and forms the products
// a_ij * x_j, stages them in LDS, and one lane per scalar row then adds the
// staged products in CSR order. Every product and every partial sum is the
// same floating-point operation, in the same order, as in the thread-per-row
// kernels and the reference's CPU loop (src/CPU/CPUBackend.cc:62-86), so
// results are bit-identical to them.
//
// Storage is row-contiguous: scalar rows as CSR (int64 row offsets), 3x3
// block rows as block CSR with the nine values of a block stored row-major
// and contiguous, so a group's loads of consecutive entries are coalesced.
#pragma once

#include "kernels.hpp"

namespace gsk {

template <class T>
struct Rows {
  int64_t rows, cols;
  const int64_t *ptr;
  const int32_t *col;
  const T *val;
  // Stored position -> row (transfer operators executed in another row
  // order); null: position == row.
  const int32_t *row_map = nullptr;
};

template <class T>
struct BRows {
  int64_t nb;
  const int64_t *ptr;   // block offsets, nb + 1
  const int32_t *bcol;  // block column per block
  const T *val;         // 9 per block, row-major
  // Stored position -> block row (transfer operators); null: identity.
  const int32_t *row_map = nullptr;
};

// Orders the LDS traffic of one wavefront: stores issued by lanes of the
// wave before the call are visible to loads issued after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Entries per lane and round for scalar rows: a round covers G * G / 16
// entries (16, 64, 256), so a long row needs few dependent load rounds.
template <int G>
constexpr int grp_unroll() { return G >= 16 ? G / 16 : 1; }
// Blocks per lane and round for block rows: two, so each lane keeps two
// independent block loads in flight (a round covers 2 G blocks).
template <int G>
constexpr int grp_bunroll() { return 2; }

// s + p[0] + p[1] + ... in this order: a pure chain of adds, unrolled by
// eight so the LDS reads of a batch issue before its adds.
template <class T>
__device__ __forceinline__ T chain_add(const T *p, int cnt, T s) {
  int j = 0;
  for (; j + 8 <= cnt; j += 8) {
    T q[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q[u] = p[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += q[u];
  }
  for (; j < cnt; ++j) s += p[j];
  return s;
}

// An entry left out of a row sum (the diagonal of a Gauss-Seidel row, the
// diagonal block of a block row) is staged as -0.0: s + (-0.0) == s for
// every s, signed zeros included, so the sum is bit-identical to one that
// skips the entry, without a data-dependent select in the add chain.
template <class T>
__device__ __forceinline__ T neg_zero() { return -T(0); }

// Lane 0 of a G-lane group receives the sum over the entries [rb, re) of a
// row in CSR order of val * x[col], leaving out entries whose column equals
// `excl` (-1: none). lp: the group's G * grp_unroll<G>() LDS slots.
template <class T, int G>
__device__ __forceinline__ T grp_row_dot(const Rows<T> &A, int64_t rb, int64_t re, const T *x,
                                         int64_t excl, T *lp, int lane) {
  constexpr int U = grp_unroll<G>();
  constexpr int C = G * U;
  T s = T(0);
  for (int64_t base = rb; base < re; base += C) {
    int32_t c[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = base + lane + u * G;
      c[u] = 0;
      v[u] = T(0);
      if (k < re) {
        c[u] = A.col[k];
        v[u] = A.val[k];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = base + lane + u * G;
      if (k < re) {
        const T xv = x[c[u]];
        lp[lane + u * G] = c[u] == excl ? neg_zero<T>() : v[u] * xv;
      }
    }
    wave_sync();
    if (lane == 0) s = chain_add(lp, static_cast<int>(re - base < C ? re - base : C), s);
    wave_sync();
  }
  return s;
}

// Transfer-operator rows: as grp_row_dot with K entries per lane and round
// (a round covers G * K entries, chosen so most rows take one round) and
// no excluded entry. Every lane loads from a clamped, in-bounds index, so
// the loads carry no branch; entries past the row end are staged but not
// summed.
template <class T, int G, int K>
__device__ __forceinline__ T grp_row_dot_k(const Rows<T> &A, int64_t rb, int64_t re, const T *x,
                                           T *lp, int lane) {
  constexpr int C = G * K;
  T s = T(0);
  for (int64_t base = rb; base < re; base += C) {
    int32_t c[K];
    T v[K], xv[K];
#pragma unroll
    for (int u = 0; u < K; ++u) {
      int64_t k = base + lane + u * G;
      k = k < re ? k : re - 1;
      c[u] = A.col[k];
      v[u] = A.val[k];
    }
#pragma unroll
    for (int u = 0; u < K; ++u) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < K; ++u) lp[lane + u * G] = v[u] * xv[u];
    wave_sync();
    if (lane == 0) s = chain_add(lp, static_cast<int>(re - base < C ? re - base : C), s);
    wave_sync();
  }
  return s;
}

// Lanes 0, 1, 2 of a G-lane group receive the three line sums of a block
// row (blocks [rb, re)) in stored order (columns 0, 1, 2 within a block),
// leaving out the block in block column `excl` (-1: none). A round covers
// C = G * grp_bunroll<G>() blocks; products are staged line-major (line l
// of round slot j at lp[3 C l + 3 j + c]) so each summing lane adds one
// contiguous run. lp: 9 C values.
// U: blocks per lane and round (transfer rows may take more than the
// smoothers' grp_bunroll, so a long row needs one round). A round issues
// every block-column index and value load first, then the x gathers, then
// the products: slots past the row end reload the row's last block (an
// in-bounds index, so no load is behind a branch) and are staged but not
// summed; the excluded block becomes -0.0 by a select.
template <class T, int G, int U = grp_bunroll<G>()>
__device__ __forceinline__ T grp_brow_dot(const BRows<T> &A, int64_t rb, int64_t re, const T *x,
                                          int64_t excl, T *lp, int lane) {
  constexpr int C = G * U;
  T s = T(0);
  for (int64_t base = rb; base < re; base += C) {
    int32_t bc[U];
    T v[U][9];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int64_t k = base + lane + u * G;
      k = k < re ? k : re - 1;
      bc[u] = A.bcol[k];
      const T *vp = A.val + 9 * k;
#pragma unroll
      for (int e = 0; e < 9; ++e) v[u][e] = vp[e];
    }
    T xv[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T *xp = x + 3 * static_cast<int64_t>(bc[u]);
#pragma unroll
      for (int c = 0; c < 3; ++c) xv[u][c] = xp[c];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = lane + u * G;
      const bool d = bc[u] == excl;
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          lp[3 * C * l + 3 * j + c] = d ? neg_zero<T>() : v[u][3 * l + c] * xv[u][c];
    }
    wave_sync();
    if (lane < 3) {
      const int cnt = static_cast<int>(re - base < C ? re - base : C);
      s = chain_add(lp + 3 * C * lane, 3 * cnt, s);
    }
    wave_sync();
  }
  return s;
}

// Row index handled by this thread's group.
template <int G>
__device__ __forceinline__ int64_t grp_row(int64_t begin) {
  return begin + xcd_block() * (kBlock / G) + threadIdx.x / G;
}

template <int G>
inline uint32_t grid_rows(int64_t rows) {
  constexpr int64_t R = kBlock / G;
  return static_cast<uint32_t>((rows + R - 1) / R);
}

// Each kernel below first issues every load that does not depend on the
// row's entries (row bounds, right-hand side, diagonal data, the stop
// word), so they share one memory round trip, and only then tests the stop
// word.
#define GSK_GROUP_SCALAR_LDS(G)              \
  __shared__ T lp[kBlock * grp_unroll<G>()]; \
  const int lane = threadIdx.x % G;          \
  const int g0 = (threadIdx.x - lane) * grp_unroll<G>()

#define GSK_GROUP_BLOCK_LDS(G)                       \
  __shared__ T lp[9 * kBlock * grp_bunroll<G>()];    \
  __shared__ T lr[3 * (kBlock / G)];                 \
  const int lane = threadIdx.x % G;                  \
  const int g0 = (threadIdx.x - lane) * grp_bunroll<G>(); \
  T *const glr = lr + 3 * (threadIdx.x / G)

// ---------------------------------------------------------------------------
// Scalar rows.
// ---------------------------------------------------------------------------
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_residual_w(Rows<T> A, const T *__restrict__ rhs,
                                                       const T *__restrict__ x,
                                                       T *__restrict__ res, const int32_t *stop) {
  GSK_GROUP_SCALAR_LDS(G);
  const int64_t row = grp_row<G>(0);
  const bool active = row < A.rows;
  int64_t rb = 0, re = 0;
  T bi = T(0);
  if (active) {
    rb = A.ptr[row];
    re = A.ptr[row + 1];
    bi = rhs[row];
  }
  if (stop_word(stop)) return;
  const T s = grp_row_dot<T, G>(A, rb, re, x, -1, lp + g0, lane);
  if (active && lane == 0) res[row] = bi - s;
}

// rhs_c = R res, corr_c = 0 or, with dinv_c, the coarse level's first
// damped Jacobi sweep in closed form (as k_restrict).
template <class T, int G, int K>
__global__ __launch_bounds__(kBlock) void k_restrict_w(Rows<T> R, const T *__restrict__ res,
                                                       T *__restrict__ rhs_c, T *__restrict__ corr_c,
                                                       const int32_t *stop,
                                                       const T *__restrict__ dinv_c = nullptr,
                                                       T omega_c = T(0)) {
  __shared__ T lp[kBlock * K];
  const int lane = threadIdx.x % G;
  const int g0 = (threadIdx.x - lane) * K;
  const int64_t pos = grp_row<G>(0);
  const bool active = pos < R.rows;
  int64_t rb = 0, re = 0, row = pos;
  T di = T(0);
  if (active) {
    rb = R.ptr[pos];
    re = R.ptr[pos + 1];
    if (R.row_map) row = R.row_map[pos];
  }
  if (active && lane == 0 && dinv_c) di = dinv_c[row];
  if (stop_word(stop)) return;
  const T s = grp_row_dot_k<T, G, K>(R, rb, re, res, lp + g0, lane);
  if (active && lane == 0) {
    rhs_c[row] = s;
    if (corr_c) corr_c[row] = dinv_c ? jacobi_zero1(s, di, omega_c) : T(0);
  }
}

// alias = 0: x += P e (as k_prolong_add); alias = 1: x = c + 1 * c with
// c = P e (as k_prolong_alias).
template <class T, int G, int K>
__global__ __launch_bounds__(kBlock) void k_prolong_w(Rows<T> P, const T *__restrict__ e,
                                                      T *__restrict__ x, int alias,
                                                      const int32_t *stop) {
  __shared__ T lp[kBlock * K];
  const int lane = threadIdx.x % G;
  const int g0 = (threadIdx.x - lane) * K;
  const int64_t pos = grp_row<G>(0);
  const bool active = pos < P.rows;
  int64_t rb = 0, re = 0, row = pos;
  T xi = T(0);
  if (active) {
    rb = P.ptr[pos];
    re = P.ptr[pos + 1];
    if (P.row_map) row = P.row_map[pos];
    if (!alias) xi = x[row];
  }
  if (stop_word(stop)) return;
  const T c = grp_row_dot_k<T, G, K>(P, rb, re, e, lp + g0, lane);
  if (active && lane == 0) x[row] = alias ? c + T(1) * c : xi + T(1) * c;
}

// One colour of multicolour Gauss-Seidel (as k_gs_color). The diagonal
// comes from `diag` (the level's diagonal, the last entry of each row with
// column == row, as the reference's row scan finds it), and x[row] is not
// read by the off-diagonal sum, so both load up front.
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_gs_color_w(Rows<T> A, const T *__restrict__ diag,
                                                       int64_t begin, int64_t end,
                                                       const T *__restrict__ rhs, T *x, T omega,
                                                       int reps, const int32_t *stop,
                                                       int xz = 0) {
  GSK_GROUP_SCALAR_LDS(G);
  const int64_t row = grp_row<G>(begin);
  const bool active = row < end;
  int64_t rb = 0, re = 0;
  T bi = T(0), xi = T(0), di = T(1);
  if (active) {
    rb = A.ptr[row];
    re = A.ptr[row + 1];
    bi = rhs[row];
    xi = xz ? T(0) : x[row];
    di = diag[row];
  }
  if (stop_word(stop)) return;
  const T off = grp_row_dot<T, G>(A, rb, re, x, row, lp + g0, lane);
  if (active && lane == 0) {
    const T nv = (bi - off) / di;
    xi = (T(1) - omega) * xi + omega * nv;
    if (reps > 1) xi = (T(1) - omega) * xi + omega * nv;
    x[row] = xi;
  }
}

// Damped Jacobi sweep into a second buffer (as k_jacobi).
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_jacobi_w(Rows<T> A, const T *__restrict__ dinv,
                                                     const T *__restrict__ rhs,
                                                     const T *__restrict__ x, T *__restrict__ xo,
                                                     T omega, const int32_t *stop) {
  GSK_GROUP_SCALAR_LDS(G);
  const int64_t row = grp_row<G>(0);
  const bool active = row < A.rows;
  int64_t rb = 0, re = 0;
  T bi = T(0), di = T(0), xi = T(0);
  if (active) {
    rb = A.ptr[row];
    re = A.ptr[row + 1];
    bi = rhs[row];
    di = dinv[row];
    xi = x[row];
  }
  if (stop_word(stop)) return;
  const T s = grp_row_dot<T, G>(A, rb, re, x, -1, lp + g0, lane);
  if (active && lane == 0) {
    T c = bi - s;
    c = c * di;
    xo[row] = xi + omega * c;
  }
}

// ---------------------------------------------------------------------------
// 3x3 block rows.
// ---------------------------------------------------------------------------
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_residual_bw(BRows<T> A, const T *__restrict__ rhs,
                                                        const T *__restrict__ x,
                                                        T *__restrict__ res, const int32_t *stop) {
  GSK_GROUP_BLOCK_LDS(G);
  (void)glr;
  const int64_t b = grp_row<G>(0);
  const bool active = b < A.nb;
  int64_t rb = 0, re = 0;
  T bl = T(0);
  if (active) {
    rb = A.ptr[b];
    re = A.ptr[b + 1];
    if (lane < 3) bl = rhs[3 * b + lane];
  }
  if (stop_word(stop)) return;
  const T s = grp_brow_dot<T, G>(A, rb, re, x, -1, lp + 9 * g0, lane);
  if (active && lane < 3) res[3 * b + lane] = bl - s;
}

// Loads row `lane` (< 3) of block b's inverse diagonal block.
template <class T>
__device__ __forceinline__ void load_dinv_row(const T *__restrict__ dinv, int64_t b, int lane,
                                              T d[3]) {
  const T *bi = dinv + 9 * b + 3 * lane;
  d[0] = bi[0];
  d[1] = bi[1];
  d[2] = bi[2];
}

// 3x3-block transfer rows, one G-lane group per block row (as
// k_restrict_b3 / k_prolong_b3), K blocks per lane and round.
#define GSK_TRANSFER_BLOCK_LDS(G, K)        \
  __shared__ T lp[9 * kBlock * K];          \
  const int lane = threadIdx.x % G;         \
  const int g0 = (threadIdx.x - lane) * K

// dinv_c as in k_restrict_b3: lane l < 3 forms line l of the closed-form
// first block Jacobi sweep from the group's three line sums.
template <class T, int G, int K>
__global__ __launch_bounds__(kBlock) void k_restrict_bw(BRows<T> R, const T *__restrict__ res,
                                                        T *__restrict__ rhs_c,
                                                        T *__restrict__ corr_c,
                                                        const int32_t *stop,
                                                        const T *__restrict__ dinv_c = nullptr,
                                                        T omega_c = T(0)) {
  GSK_TRANSFER_BLOCK_LDS(G, K);
  const int64_t pos = grp_row<G>(0);
  const bool active = pos < R.nb;
  int64_t rb = 0, re = 0, b = pos;
  if (active) {
    rb = R.ptr[pos];
    re = R.ptr[pos + 1];
    if (R.row_map) b = R.row_map[pos];
  }
  T d[3] = {T(0), T(0), T(0)};
  if (active && lane < 3 && dinv_c) load_dinv_row(dinv_c, b, lane, d);
  if (stop_word(stop)) return;
  const T s = grp_brow_dot<T, G, K>(R, rb, re, res, -1, lp + 9 * g0, lane);
  T xl = T(0);
  if (dinv_c) {
    // The group's three line sums, from its lanes 0..2, in every lane.
    const int base = static_cast<int>(threadIdx.x & 63) & ~(G - 1);
    const T s0 = __shfl(s, base, 64), s1 = __shfl(s, base + 1, 64), s2 = __shfl(s, base + 2, 64);
    xl = jacobi_zero3(d, s0, s1, s2, omega_c);
  }
  if (active && lane < 3) {
    rhs_c[3 * b + lane] = s;
    if (corr_c) corr_c[3 * b + lane] = xl;
  }
}

template <class T, int G, int K>
__global__ __launch_bounds__(kBlock) void k_prolong_bw(BRows<T> P, const T *__restrict__ e,
                                                       T *__restrict__ x, int alias,
                                                       const int32_t *stop) {
  GSK_TRANSFER_BLOCK_LDS(G, K);
  const int64_t pos = grp_row<G>(0);
  const bool active = pos < P.nb;
  int64_t rb = 0, re = 0, b = pos;
  T xl = T(0);
  if (active) {
    rb = P.ptr[pos];
    re = P.ptr[pos + 1];
    if (P.row_map) b = P.row_map[pos];
    if (!alias && lane < 3) xl = x[3 * b + lane];
  }
  if (stop_word(stop)) return;
  const T c = grp_brow_dot<T, G, K>(P, rb, re, e, -1, lp + 9 * g0, lane);
  if (active && lane < 3) x[3 * b + lane] = alias ? c + T(1) * c : xl + T(1) * c;
}

// nv_l = sum_c Dinv_b[l][c] r_c for lane l < 3 of the group, from the three
// line residuals held by lanes 0..2 (d: this lane's row of Dinv_b).
template <class T>
__device__ __forceinline__ T block_apply(const T d[3], bool active, T rl, T *lr, int lane) {
  if (lane < 3) lr[lane] = rl;
  wave_sync();
  T nv = T(0);
  if (active && lane < 3) {
    nv += d[0] * lr[0];
    nv += d[1] * lr[1];
    nv += d[2] * lr[2];
  }
  return nv;
}

// One colour of 3x3 block Gauss-Seidel over block rows [begin, end)
// (as k_block_gs_color).
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_block_gs_color_bw(BRows<T> A, const T *__restrict__ dinv,
                                                              int64_t begin, int64_t end,
                                                              const T *__restrict__ rhs, T *x,
                                                              T omega, int reps,
                                                              const int32_t *stop, int xz = 0) {
  GSK_GROUP_BLOCK_LDS(G);
  const int64_t b = grp_row<G>(begin);
  const bool active = b < end;
  int64_t rb = 0, re = 0;
  T bl = T(0), xl = T(0), d[3] = {T(0), T(0), T(0)};
  if (active) {
    rb = A.ptr[b];
    re = A.ptr[b + 1];
    if (lane < 3) {
      bl = rhs[3 * b + lane];
      xl = xz ? T(0) : x[3 * b + lane];
      load_dinv_row(dinv, b, lane, d);
    }
  }
  if (stop_word(stop)) return;
  const T off = grp_brow_dot<T, G>(A, rb, re, x, b, lp + 9 * g0, lane);
  const T nv = block_apply(d, active, active && lane < 3 ? bl - off : T(0), glr, lane);
  if (active && lane < 3) {
    xl = (T(1) - omega) * xl + omega * nv;
    if (reps > 1) xl = (T(1) - omega) * xl + omega * nv;
    x[3 * b + lane] = xl;
  }
}

// 3x3 block Jacobi sweep into a second buffer (as k_block_jacobi_b3).
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_block_jacobi_bw(BRows<T> A, const T *__restrict__ dinv,
                                                            const T *__restrict__ rhs,
                                                            const T *__restrict__ x,
                                                            T *__restrict__ xo, T omega,
                                                            const int32_t *stop) {
  GSK_GROUP_BLOCK_LDS(G);
  const int64_t b = grp_row<G>(0);
  const bool active = b < A.nb;
  int64_t rb = 0, re = 0;
  T bl = T(0), xl = T(0), d[3] = {T(0), T(0), T(0)};
  if (active) {
    rb = A.ptr[b];
    re = A.ptr[b + 1];
    if (lane < 3) {
      bl = rhs[3 * b + lane];
      xl = x[3 * b + lane];
      load_dinv_row(dinv, b, lane, d);
    }
  }
  if (stop_word(stop)) return;
  const T s = grp_brow_dot<T, G>(A, rb, re, x, -1, lp + 9 * g0, lane);
  const T v = block_apply(d, active, active && lane < 3 ? bl - s : T(0), glr, lane);
  if (active && lane < 3) xo[3 * b + lane] = xl + omega * v;
}

#undef GSK_GROUP_SCALAR_LDS
#undef GSK_GROUP_BLOCK_LDS

}  // namespace gsk
