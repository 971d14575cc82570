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
};

template <class T>
struct BRows {
  int64_t nb;
  const int64_t *ptr;   // block offsets, nb + 1
  const int32_t *bcol;  // block column per block
  const T *val;         // 9 per block, row-major
};

// Orders the LDS traffic of one wavefront: stores issued by lanes of the
// wave before the call are visible to loads issued after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane 0 of a G-lane group receives the sum over the entries of `row` in
// CSR order of val * x[col], leaving out entries whose column equals
// `excl` (-1: none); `diag` receives the value of the last such entry.
// lp / lf: the group's G product / flag slots in LDS.
template <class T, int G>
__device__ __forceinline__ T grp_row_dot(const Rows<T> &A, int64_t row, bool active, const T *x,
                                         int64_t excl, T &diag, T *lp, uint8_t *lf, int lane) {
  T s = T(0);
  diag = T(0);
  int64_t rb = 0, re = 0;
  if (active) {
    rb = A.ptr[row];
    re = A.ptr[row + 1];
  }
  for (int64_t base = rb; base < re; base += G) {
    const int64_t k = base + lane;
    if (k < re) {
      const int32_t c = A.col[k];
      const T v = A.val[k];
      const bool d = c == excl;
      lp[lane] = d ? v : v * x[c];
      lf[lane] = d;
    }
    wave_sync();
    if (lane == 0) {
      const int cnt = static_cast<int>(re - base < G ? re - base : G);
      for (int j = 0; j < cnt; ++j) {
        const T p = lp[j];
        if (lf[j]) diag = p;
        else s += p;
      }
    }
    wave_sync();
  }
  return s;
}

// Lanes 0, 1, 2 of a G-lane group receive the three line sums of block row
// b over its blocks in stored order (columns 0, 1, 2 within a block),
// leaving out the block in block column `excl` (-1: none). lp: 9 G values,
// lf: G flags.
template <class T, int G>
__device__ __forceinline__ T grp_brow_dot(const BRows<T> &A, int64_t b, bool active, const T *x,
                                          int64_t excl, T *lp, uint8_t *lf, int lane) {
  T s = T(0);
  int64_t rb = 0, re = 0;
  if (active) {
    rb = A.ptr[b];
    re = A.ptr[b + 1];
  }
  for (int64_t base = rb; base < re; base += G) {
    const int64_t k = base + lane;
    if (k < re) {
      const int32_t bc = A.bcol[k];
      const bool d = bc == excl;
      lf[lane] = d;
      if (!d) {
        const T *v = A.val + 9 * k;
        const T x0 = x[3 * static_cast<int64_t>(bc)];
        const T x1 = x[3 * static_cast<int64_t>(bc) + 1];
        const T x2 = x[3 * static_cast<int64_t>(bc) + 2];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          lp[9 * lane + 3 * l + 0] = v[3 * l + 0] * x0;
          lp[9 * lane + 3 * l + 1] = v[3 * l + 1] * x1;
          lp[9 * lane + 3 * l + 2] = v[3 * l + 2] * x2;
        }
      }
    }
    wave_sync();
    if (lane < 3) {
      const int cnt = static_cast<int>(re - base < G ? re - base : G);
      for (int j = 0; j < cnt; ++j) {
        if (lf[j]) continue;
        s += lp[9 * j + 3 * lane + 0];
        s += lp[9 * j + 3 * lane + 1];
        s += lp[9 * j + 3 * lane + 2];
      }
    }
    wave_sync();
  }
  return s;
}

// Row index handled by this thread's group, and the group's lane.
template <int G>
__device__ __forceinline__ int64_t grp_row(int64_t begin) {
  return begin + xcd_block() * (kBlock / G) + threadIdx.x / G;
}

template <int G>
inline uint32_t grid_rows(int64_t rows) {
  constexpr int64_t R = kBlock / G;
  return static_cast<uint32_t>((rows + R - 1) / R);
}

// ---------------------------------------------------------------------------
// Scalar rows.
// ---------------------------------------------------------------------------
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_residual_w(Rows<T> A, const T *__restrict__ rhs,
                                                       const T *__restrict__ x,
                                                       T *__restrict__ res, const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ T lp[kBlock];
  __shared__ uint8_t lf[kBlock];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t row = grp_row<G>(0);
  const bool active = row < A.rows;
  T diag;
  const T s = grp_row_dot<T, G>(A, row, active, x, -1, diag, lp + g0, lf + g0, lane);
  if (active && lane == 0) res[row] = rhs[row] - s;
}

// rhs_c = R res, corr_c = 0 (as k_restrict).
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_restrict_w(Rows<T> R, const T *__restrict__ res,
                                                       T *__restrict__ rhs_c, T *__restrict__ corr_c,
                                                       const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ T lp[kBlock];
  __shared__ uint8_t lf[kBlock];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t row = grp_row<G>(0);
  const bool active = row < R.rows;
  T diag;
  const T s = grp_row_dot<T, G>(R, row, active, res, -1, diag, lp + g0, lf + g0, lane);
  if (active && lane == 0) {
    rhs_c[row] = s;
    if (corr_c) corr_c[row] = T(0);
  }
}

// alias = 0: x += P e (as k_prolong_add); alias = 1: x = c + 1 * c with
// c = P e (as k_prolong_alias).
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_prolong_w(Rows<T> P, const T *__restrict__ e,
                                                      T *__restrict__ x, int alias,
                                                      const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ T lp[kBlock];
  __shared__ uint8_t lf[kBlock];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t row = grp_row<G>(0);
  const bool active = row < P.rows;
  T diag;
  const T c = grp_row_dot<T, G>(P, row, active, e, -1, diag, lp + g0, lf + g0, lane);
  if (active && lane == 0) x[row] = alias ? c + T(1) * c : x[row] + T(1) * c;
}

// One colour of multicolour Gauss-Seidel (as k_gs_color).
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_gs_color_w(Rows<T> A, int64_t begin, int64_t end,
                                                       const T *__restrict__ rhs, T *x, T omega,
                                                       int reps, const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ T lp[kBlock];
  __shared__ uint8_t lf[kBlock];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t row = grp_row<G>(begin);
  const bool active = row < end;
  T diag;
  const T off = grp_row_dot<T, G>(A, row, active, x, row, diag, lp + g0, lf + g0, lane);
  if (active && lane == 0) {
    const T nv = (rhs[row] - off) / diag;
    T xi = x[row];
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
  if (stopped(stop)) return;
  __shared__ T lp[kBlock];
  __shared__ uint8_t lf[kBlock];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t row = grp_row<G>(0);
  const bool active = row < A.rows;
  T diag;
  const T s = grp_row_dot<T, G>(A, row, active, x, -1, diag, lp + g0, lf + g0, lane);
  if (active && lane == 0) {
    T c = rhs[row] - s;
    c = c * dinv[row];
    xo[row] = x[row] + omega * c;
  }
}

// ---------------------------------------------------------------------------
// 3x3 block rows.
// ---------------------------------------------------------------------------
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_residual_bw(BRows<T> A, const T *__restrict__ rhs,
                                                        const T *__restrict__ x,
                                                        T *__restrict__ res, const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ T lp[9 * kBlock];
  __shared__ uint8_t lf[kBlock];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t b = grp_row<G>(0);
  const bool active = b < A.nb;
  const T s = grp_brow_dot<T, G>(A, b, active, x, -1, lp + 9 * g0, lf + g0, lane);
  if (active && lane < 3) res[3 * b + lane] = rhs[3 * b + lane] - s;
}

// Applies x_b update of block GS / block Jacobi from the three line
// residuals held by lanes 0..2 of the group: nv_l = sum_c Dinv_b[l][c] r_c.
template <class T>
__device__ __forceinline__ T block_apply(const T *__restrict__ bi, bool active, T rl, T *lr,
                                         int lane) {
  if (lane < 3) lr[lane] = rl;
  wave_sync();
  T nv = T(0);
  if (active && lane < 3) {
    nv += bi[3 * lane + 0] * lr[0];
    nv += bi[3 * lane + 1] * lr[1];
    nv += bi[3 * lane + 2] * lr[2];
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
                                                              const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ T lp[9 * kBlock];
  __shared__ uint8_t lf[kBlock];
  __shared__ T lr[3 * (kBlock / G)];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t b = grp_row<G>(begin);
  const bool active = b < end;
  const T off = grp_brow_dot<T, G>(A, b, active, x, b, lp + 9 * g0, lf + g0, lane);
  const T rl = (active && lane < 3) ? rhs[3 * b + lane] - off : T(0);
  const T nv = block_apply(dinv + 9 * b, active, rl, lr + 3 * (threadIdx.x / G), lane);
  if (active && lane < 3) {
    T xi = x[3 * b + lane];
    xi = (T(1) - omega) * xi + omega * nv;
    if (reps > 1) xi = (T(1) - omega) * xi + omega * nv;
    x[3 * b + lane] = xi;
  }
}

// 3x3 block Jacobi sweep into a second buffer (as k_block_jacobi_b3).
template <class T, int G>
__global__ __launch_bounds__(kBlock) void k_block_jacobi_bw(BRows<T> A, const T *__restrict__ dinv,
                                                            const T *__restrict__ rhs,
                                                            const T *__restrict__ x,
                                                            T *__restrict__ xo, T omega,
                                                            const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ T lp[9 * kBlock];
  __shared__ uint8_t lf[kBlock];
  __shared__ T lr[3 * (kBlock / G)];
  const int lane = threadIdx.x % G, g0 = threadIdx.x - lane;
  const int64_t b = grp_row<G>(0);
  const bool active = b < A.nb;
  const T s = grp_brow_dot<T, G>(A, b, active, x, -1, lp + 9 * g0, lf + g0, lane);
  const T rl = (active && lane < 3) ? rhs[3 * b + lane] - s : T(0);
  const T v = block_apply(dinv + 9 * b, active, rl, lr + 3 * (threadIdx.x / G), lane);
  if (active && lane < 3) xo[3 * b + lane] = x[3 * b + lane] + omega * v;
}

}  // namespace gsk
