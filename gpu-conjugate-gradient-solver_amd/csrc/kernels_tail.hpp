// The V-cycle below a level whose operators are small, in one workgroup.
//
// On the coarse levels of a scalar hierarchy a Gauss-Seidel colour holds a
// few dozen rows: one launch per colour costs its launch plus three
// dependent memory round trips (row offsets, then indices and values, then
// x) for a few microseconds of work, and a symmetric sweep of a 40-colour
// level is 80 launches. k_vcycle_tail runs every level from the tail's
// first down to the coarsest and back (pre-smoothing, residual,
// restriction, coarse sweeps, prolongation, post-smoothing) in one
// workgroup of 1024 threads, with a workgroup barrier where the launches
// had their boundaries. Each level's x, right-hand side, residual and row
// offsets live in LDS for the whole kernel, so a colour step's gathers are
// LDS reads; the matrix rows stream from global memory, and a group loads
// its next row's indices and values (which no colour changes) before it
// forms the current row's sum, so those loads overlap the sum and the
// barrier.
//
// Every row sum is the one the per-colour kernels form: the products of a
// row in CSR order, added one after another from +0 by one lane, the
// diagonal product (Gauss-Seidel's excluded entry) staged as -0.0, which
// leaves every sum unchanged. The updates are those of k_gs_color_zero,
// k_gs_color, k_residual, k_restrict, k_prolong_add and k_prolong_alias, in
// the colour sequence of the engine's symmetric sweeps, so the results are
// bit-identical to the launch-per-colour path
// (src/Smoothers/SymmetricGaussSeidelSmoother.cc:80-128,
// src/Smoothers/MultigridSmoother.cc:337-412).
#pragma once

#include "kernels_wide.hpp"

namespace gsk {

constexpr int kTailBlock = 1024;
constexpr int kTailMaxLevels = 8;
constexpr int kTailG = 16;                      // lanes per row
constexpr int kTailNG = kTailBlock / kTailG;    // rows per round
constexpr int kTailKP = 8;   // entries per lane held in registers: a row's first 128
constexpr int kTailKS = 8;   // entries per lane per LDS staging round (all of the registers')
constexpr int64_t kTailLdsMax = 160 * 1024;     // LDS of one workgroup on gfx950

template <class T>
struct TailLevel {
  Rows<T> A;  // level operator in the internal (colour) order, row-contiguous CSR
  Rows<T> R;  // restriction to the next level (rows: the next level's unknowns)
  Rows<T> P;  // prolongation from the next level (rows: this level's unknowns)
  const T *diag;
  const int64_t *coff;  // colour offsets, ncolors + 1
  const T *rhs;         // the first level's right-hand side (global)
  T *x;                 // the first level's solution (global)
  int64_t n;
  int64_t lds_off;      // byte offset of x, rhs, res (n each) and row offsets (n + 1, int32)
  int ncolors, pre, post;
  T omega;
};

template <class T>
struct TailArgs {
  TailLevel<T> L[kTailMaxLevels];
  int nl;            // levels in the tail; the last is the coarsest
  int coarse_steps;  // symmetric sweeps on the coarsest level
  int alias;         // prolongation in the reference's aliased form (x = c + 1 * c)
  int64_t lp_off;    // byte offset of the staging slots (kTailBlock * kTailKS values)
};

// LDS bytes of one level's block (x, rhs, res, row offsets), 16-aligned.
template <class T>
constexpr int64_t tail_level_bytes(int64_t n) {
  return ((3 * n * static_cast<int64_t>(sizeof(T)) + 4 * (n + 1)) + 15) / 16 * 16;
}

template <class T>
struct TailLds {
  T *x, *rhs, *res;
  int32_t *ptr;
};

template <class T>
__device__ __forceinline__ TailLds<T> tail_lds(const TailLevel<T> &L, unsigned char *base) {
  T *x = reinterpret_cast<T *>(base + L.lds_off);
  return TailLds<T>{x, x + L.n, x + 2 * L.n, reinterpret_cast<int32_t *>(x + 3 * L.n)};
}

// The number of steps of `iters` symmetric sweeps over C colours (forward
// 0..C-1, backward C-1..0, back-to-back repeats of a colour merged into one
// two-update step) and step i of that sequence.
__device__ __forceinline__ int tail_steps(int C, int iters) {
  return C == 1 ? iters : iters * (2 * C - 2) + 1;
}
__device__ __forceinline__ void tail_step(int C, int iters, int i, int &colour, int &reps) {
  if (C == 1) {
    colour = 0;
    reps = 2;
    return;
  }
  const int m = 2 * C - 2;
  if (i == iters * m) {
    colour = 0;
    reps = 1;
    return;
  }
  const int t = i / m, u = i % m;
  if (u == 0) {
    colour = 0;
    reps = t == 0 ? 1 : 2;
  } else if (u < C - 1) {
    colour = u;
    reps = 1;
  } else if (u == C - 1) {
    colour = C - 1;
    reps = 2;
  } else {
    colour = 2 * C - 2 - u;
    reps = 1;
  }
}

// A group's row: offsets, its first kTailG * kTailKP indices and values
// (lane j holds entries j, j + G, ...; loads from clamped in-bounds
// indices), and the diagonal.
template <class T>
struct TailRow {
  int32_t c[kTailKP];
  T v[kTailKP];
  T diag;
  int64_t rb, re;
};

// ptr32: the row offsets in LDS, or null for M's own (global) offsets.
template <class T>
__device__ __forceinline__ void tail_load(TailRow<T> &g, const Rows<T> &M, const int32_t *ptr32,
                                          const T *diag, int64_t row, bool active, int lane) {
  g.rb = 0;
  g.re = 0;
  if (active) {
    g.rb = ptr32 ? ptr32[row] : M.ptr[row];
    g.re = ptr32 ? ptr32[row + 1] : M.ptr[row + 1];
  }
  const bool any = g.re > g.rb;
#pragma unroll
  for (int u = 0; u < kTailKP; ++u) {
    int64_t k = g.rb + lane + u * kTailG;
    k = k < g.re ? k : g.re - 1;
    g.c[u] = any ? M.col[k] : 0;
    g.v[u] = any ? M.val[k] : T(0);
  }
  g.diag = (active && diag) ? diag[row] : T(1);
}

// Lane 0 receives the row's sum of val * xs[col] in CSR order, the entry in
// column excl staged as -0.0 (-1: none); xs in LDS. Staging rounds of
// kTailG * kTailKS entries (entry j of a round at lp[j]); entries past the
// registers' first kTailG * kTailKP are loaded here.
template <class T>
__device__ __forceinline__ T tail_dot(const TailRow<T> &g, const Rows<T> &M, const T *xs,
                                      int64_t excl, T *lp, int lane) {
  constexpr int C = kTailG * kTailKS;
  T s = T(0);
#pragma unroll
  for (int h = 0; h < kTailKP / kTailKS; ++h) {
    const int64_t base = g.rb + h * C;
    if (base < g.re) {
#pragma unroll
      for (int u = 0; u < kTailKS; ++u) {
        const int32_t c = g.c[h * kTailKS + u];
        lp[lane + u * kTailG] = c == excl ? neg_zero<T>() : g.v[h * kTailKS + u] * xs[c];
      }
      wave_sync();
      if (lane == 0) s = chain_add(lp, static_cast<int>(g.re - base < C ? g.re - base : C), s);
      wave_sync();
    }
  }
  for (int64_t base = g.rb + kTailG * kTailKP; base < g.re; base += C) {
#pragma unroll
    for (int u = 0; u < kTailKS; ++u) {
      int64_t k = base + lane + u * kTailG;
      k = k < g.re ? k : g.re - 1;
      const int32_t c = M.col[k];
      lp[lane + u * kTailG] = c == excl ? neg_zero<T>() : M.val[k] * xs[c];
    }
    wave_sync();
    if (lane == 0) s = chain_add(lp, static_cast<int>(g.re - base < C ? g.re - base : C), s);
    wave_sync();
  }
  return s;
}

struct TailCtx {
  int lane, grp;
};

// `iters` symmetric sweeps of one level. From x = 0 the first step is
// colour 0's closed form with every other row zeroed (k_gs_color_zero); a
// colour step applies x_i = (1 - w) x_i + w (b_i - sum_{j != i} a_ij x_j) /
// a_ii once or twice (k_gs_color). Items are (step, round of kTailNG rows);
// each group loads its next item's row before summing the current one.
template <class T>
__device__ void tail_sweep(const TailLevel<T> &L, const TailLds<T> &S, int iters, bool zero_x,
                           T *lp, const TailCtx &t) {
  if (iters <= 0) {
    if (zero_x) {
      for (int64_t i = threadIdx.x; i < L.n; i += kTailBlock) S.x[i] = T(0);
      __syncthreads();
    }
    return;
  }
  const int C = L.ncolors;
  const int nsteps = tail_steps(C, iters);
  int i = 0, reps0 = 0;  // reps0 > 0: step 0's update count after the closed form
  if (zero_x) {
    const int64_t e0 = L.coff[1];
    for (int64_t r = threadIdx.x; r < L.n; r += kTailBlock) {
      if (r < e0) {
        const T nv = S.rhs[r] / L.diag[r];
        S.x[r] = L.omega * nv;
      } else {
        S.x[r] = T(0);
      }
    }
    __syncthreads();
    int c0, r0;
    tail_step(C, iters, 0, c0, r0);
    if (r0 == 1) i = 1;
    else reps0 = 1;
  }
  if (i >= nsteps) return;
  int c, reps;
  tail_step(C, iters, i, c, reps);
  if (i == 0 && reps0 > 0) reps = reps0;
  int64_t b = L.coff[c], e = L.coff[c + 1], r = 0;
  TailRow<T> cur, nxt;
  tail_load(cur, L.A, S.ptr, L.diag, b + t.grp, b + t.grp < e, t.lane);
  while (i < nsteps) {
    // The next item (uniform across the workgroup).
    int ni = i, nc = c, nreps = reps;
    int64_t nb = b, ne = e, nr = r + 1;
    if (b + nr * kTailNG >= e) {
      ni = i + 1;
      nr = 0;
      if (ni < nsteps) {
        tail_step(C, iters, ni, nc, nreps);
        nb = L.coff[nc];
        ne = L.coff[nc + 1];
      }
    }
    if (ni < nsteps) {
      const int64_t nrow = nb + nr * kTailNG + t.grp;
      tail_load(nxt, L.A, S.ptr, L.diag, nrow, nrow < ne, t.lane);
    }
    const int64_t row = b + r * kTailNG + t.grp;
    const T off = tail_dot(cur, L.A, S.x, row, lp, t.lane);
    if (row < e && t.lane == 0) {
      const T nv = (S.rhs[row] - off) / cur.diag;
      T xi = S.x[row];
      xi = (T(1) - L.omega) * xi + L.omega * nv;
      if (reps > 1) xi = (T(1) - L.omega) * xi + L.omega * nv;
      S.x[row] = xi;
    }
    if (ni != i) __syncthreads();  // a colour's rows are all updated before the next colour
    cur = nxt;
    i = ni;
    c = nc;
    reps = nreps;
    b = nb;
    e = ne;
    r = nr;
  }
}

// res = rhs - A x (k_residual).
template <class T>
__device__ void tail_residual(const TailLevel<T> &L, const TailLds<T> &S, T *lp, const TailCtx &t) {
  for (int64_t base = 0; base < L.n; base += kTailNG) {
    const int64_t row = base + t.grp;
    TailRow<T> g;
    tail_load(g, L.A, S.ptr, static_cast<const T *>(nullptr), row, row < L.n, t.lane);
    const T s = tail_dot(g, L.A, S.x, -1, lp, t.lane);
    if (row < L.n && t.lane == 0) S.res[row] = S.rhs[row] - s;
  }
  __syncthreads();
}

// y = M v (k_restrict); with `add`, y = y + 1 * (M v) (k_prolong_add); with
// `alias`, y = c + 1 * c (k_prolong_alias). v and y in LDS.
template <class T>
__device__ void tail_apply(const Rows<T> &M, const T *v, T *y, bool add, bool alias, T *lp,
                           const TailCtx &t) {
  for (int64_t base = 0; base < M.rows; base += kTailNG) {
    const int64_t row = base + t.grp;
    TailRow<T> g;
    tail_load(g, M, static_cast<const int32_t *>(nullptr), static_cast<const T *>(nullptr), row,
              row < M.rows, t.lane);
    const T c = tail_dot(g, M, v, -1, lp, t.lane);
    if (row < M.rows && t.lane == 0) {
      if (alias) y[row] = c + T(1) * c;
      else if (add) y[row] = y[row] + T(1) * c;
      else y[row] = c;
    }
  }
  __syncthreads();
}

template <class T>
__global__ __launch_bounds__(kTailBlock) void k_vcycle_tail(const TailArgs<T> a,
                                                           const int32_t *stop) {
  if (stop_word(stop)) return;
  extern __shared__ __align__(16) unsigned char tail_mem[];
  const TailCtx t{static_cast<int>(threadIdx.x % kTailG), static_cast<int>(threadIdx.x / kTailG)};
  T *lp = reinterpret_cast<T *>(tail_mem + a.lp_off) + t.grp * kTailG * kTailKS;
  const int m = a.nl;
  // Row offsets of every level's operator and the first level's
  // right-hand side into LDS.
  for (int j = 0; j < m; ++j) {
    const TailLds<T> S = tail_lds(a.L[j], tail_mem);
    for (int64_t i = threadIdx.x; i <= a.L[j].n; i += kTailBlock)
      S.ptr[i] = static_cast<int32_t>(a.L[j].A.ptr[i]);
  }
  {
    const TailLds<T> S = tail_lds(a.L[0], tail_mem);
    for (int64_t i = threadIdx.x; i < a.L[0].n; i += kTailBlock) S.rhs[i] = a.L[0].rhs[i];
  }
  __syncthreads();
  for (int j = 0; j + 1 < m; ++j) {
    const TailLevel<T> &L = a.L[j];
    const TailLds<T> S = tail_lds(L, tail_mem);
    tail_sweep(L, S, L.pre, true, lp, t);
    // With no pre-smoothing from x = 0 the residual is rhs bit for bit.
    const T *res = S.rhs;
    if (L.pre > 0) {
      tail_residual(L, S, lp, t);
      res = S.res;
    }
    tail_apply(L.R, res, tail_lds(a.L[j + 1], tail_mem).rhs, false, false, lp, t);
  }
  tail_sweep(a.L[m - 1], tail_lds(a.L[m - 1], tail_mem), a.coarse_steps, true, lp, t);
  for (int j = m - 2; j >= 0; --j) {
    const TailLevel<T> &L = a.L[j];
    const TailLds<T> S = tail_lds(L, tail_mem);
    tail_apply(L.P, tail_lds(a.L[j + 1], tail_mem).x, S.x, true, a.alias != 0, lp, t);
    tail_sweep(L, S, L.post, false, lp, t);
  }
  {
    const TailLds<T> S = tail_lds(a.L[0], tail_mem);
    for (int64_t i = threadIdx.x; i < a.L[0].n; i += kTailBlock) a.L[0].x[i] = S.x[i];
  }
}

}  // namespace gsk
