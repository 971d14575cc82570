// HIP kernels of the AMG-PCG solve phase for gfx950 (CDNA4, wave64).
//
// Matrix storage on the device is SELL-64: rows are cut into slices of 64
// (one wavefront), each slice stores its rows slot-major with the slice's
// own width, so lane l of a wave reads entry j of its row at
// slice_ptr[s] + 64 * j + l -- unit-stride 64-wide loads for column
// indices and values, and no padding beyond the widest row of a slice. A
// row's entries keep their CSR order, so every row sum here is computed in
// exactly the reference's order (src/CPU/CPUBackend.cc:62-86). Padding
// slots hold value 0 and a valid column, and add nothing.
//
// 3x3-block operators (elasticity) use BSR3-SELL-64: one thread per block
// row, 64 block rows per slice, per block slot one block-column index and
// nine value planes of 64 lanes each.
//
// Reductions are deterministic: each workgroup reduces its chunk with wave
// shuffles and LDS in a fixed tree and stores one partial; a one-workgroup
// finalize kernel launched next on the same stream sums the partials in a
// fixed order and applies the scalar epilogue. The kernel boundary makes
// the partials visible across XCDs; a last-block-arrives scheme would need
// a device-scope release (an L2 writeback on this part) in every block.
//
// Every kernel that runs inside the PCG loop takes a pointer to the
// solver's stop word and returns at once when it is set, so iterations
// captured past convergence are no-ops.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gsk {

constexpr int kBlock = 256;
constexpr int kSlice = 64;

template <class T>
struct Sell {
  int64_t rows;
  int64_t cols;
  const int64_t *slice_ptr;  // entry offset of each slice (multiple of 64)
  const int32_t *slice_w;    // width of each slice
  const int32_t *col;
  const T *val;
  // Stored position -> row (rows sorted by length within windows, transfer
  // operators only); null: position == row.
  const int32_t *row_map = nullptr;
};

// Row held at stored position `pos` of a SELL operator.
template <class T>
__device__ __forceinline__ int64_t sell_row_of(const Sell<T> &A, int64_t pos) {
  return A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
}

template <class T>
struct Bsr3 {
  int64_t nb;                // block rows
  const int64_t *slice_ptr;  // block-slot offset of each slice (multiple of 64)
  const int32_t *slice_w;    // width (blocks) of each slice
  const int32_t *bcol;       // block column per block slot
  const T *val;              // 9 planes of 64 per block slot
  // Stored position -> block row (transfer operators); null: identity.
  const int32_t *row_map = nullptr;
};

// Solver control block, one per solve, in device memory.
template <class T>
struct Ctl {
  T bb;        // b . b
  T rho;       // r . z of the current iterate
  T rho_new;
  T alpha;
  T beta;
  T pap;
  T rr;
  T rel;       // rr / bb, the recursive relres^2
  T tol2;
  int64_t iter;       // loop index of the next iteration
  int64_t maxit;
  int64_t iters_out;  // reported iteration count (reference convention)
  // 0 run, 1 converged, 2 maxit, 3 non-finite, 4 breakdown, 5 the
  // recursive residual passed the tolerance and waits for the replaced one
  // (every kernel but the replacement pair skips; the pair sets 1, 2 or
  // back to 0, and with 0 the kernels after it finish the iteration)
  int32_t stop;
  int32_t breakdown;  // 1: p.Ap <= 0, 2: r.z <= 0 (recorded, not fatal)
  T *hist;            // relres^2 per iteration (length maxit + 1) or null
  // Residual replacement (mixed precision): every replace_every
  // iterations, and whenever the recursive residual passes the tolerance,
  // r is recomputed as b - A x with the full-precision operator and the
  // stop test uses it. 0: off.
  int64_t replace_every;
  int32_t replace;    // a periodic replacement is due (kept until done)
};

// The replacement pair runs for a due periodic replacement or a pending
// convergence check.
template <class T>
__device__ __forceinline__ bool replace_runs(const Ctl<T> *ctl) {
  const int32_t s = ctl->stop;
  return s == 5 || (s == 0 && ctl->replace);
}

// Blocks are dispatched round-robin over the eight XCDs; remap so each XCD
// walks one contiguous range of rows, which keeps stencil neighbours of x
// inside the same L2.
__device__ __forceinline__ int64_t xcd_block() {
  const uint32_t g = gridDim.x, b = blockIdx.x;
  const uint32_t full = g & ~7u;
  if (b >= full) return b;
  const uint32_t per = full >> 3;
  return static_cast<int64_t>(b & 7u) * per + (b >> 3);
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a 256-thread block in a fixed tree; result valid in thread 0.
template <class T>
__device__ __forceinline__ T block_sum(T v) {
  __shared__ T lds[kBlock / 64];
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  T r = T(0);
  if (threadIdx.x == 0) r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return r;
}

// Stores this block's sum as partials[blockIdx.x].
template <class T>
__device__ __forceinline__ void block_partial(T v, T *partials) {
  v = block_sum(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// Sum of nparts partials in a fixed order, for a one-block launch of kBlock
// threads; result valid in thread 0.
template <class T>
__device__ __forceinline__ T fixed_sum(const T *partials, uint32_t nparts) {
  // Thread t adds partials t, t + kBlock, ... in increasing order; the loads
  // of eight such terms are issued before their adds.
  T s = T(0);
  uint32_t i = threadIdx.x;
  for (; i + 7 * kBlock < nparts; i += 8 * kBlock) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partials[i + u * kBlock];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; i < nparts; i += kBlock) s += partials[i];
  return block_sum(s);
}

// Deterministic dot product (Backend::encodeDot): each workgroup of a grid
// of at most kDotMaxGrid strides over the vectors and stores one partial
// (fixed tree), then one workgroup (k_dot_final) adds the partials in
// index order into the output scalar.
constexpr uint32_t kDotMaxGrid = 2048;

inline uint32_t dot_grid(int64_t n) {
  const int64_t g = (n + kBlock - 1) / kBlock;
  return static_cast<uint32_t>(g < 1 ? 1 : (g > kDotMaxGrid ? kDotMaxGrid : g));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_dot(const T *x, const T *y, int64_t n, T *partials) {
  T v = T(0);
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock)
    v += x[i] * y[i];
  block_partial(v, partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_dot_final(const T *partials, uint32_t G, T *out) {
  const T s = fixed_sum(partials, G);
  if (threadIdx.x == 0) *out = s;
}

template <class T>
__device__ __forceinline__ bool finite_val(T v) {
  return v == v && v - v == T(0);
}

// Matrix streams (values, indices, dense tiles) read once per pass:
// non-temporal loads keep them from displacing the vectors the gathers
// reuse. Used where measured faster: the FP64 CG SpMV (202 -> 188 us), the
// coarse tile product (137 -> 119 us) and the block residual; the
// smoothers' colours and the transfers measured slower with them. The
// scalar CG operator's row dot uses them too.
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

// ---------------------------------------------------------------------------
// Row kernels (device helpers).
// ---------------------------------------------------------------------------
// X: the type of x and of the products and sum (the stored values are
// widened to it; X = T for every operator but the CG one read in the V-cycle
// type).
// The slots of one SELL row (first slot at `base`, slice width w) in
// batches of kSellBatch: the batch's column indices and values are loaded
// together, then its x gathers, then the products are added in slot order
// -- two dependent memory round trips per batch instead of two per slot,
// with the arithmetic (order included) of a plain slot loop. Slots past
// the width reload the last slot (an in-bounds address) and add nothing.
// SKIP: the entry in column `row` contributes -0.0 instead of its product
// (s + -0.0 == s for every s, signed zeros included), so the sum is the
// reference's loop that skips it (src/CPU/CPUBackend.cc:62-86) without a
// branch around the loads. NT: non-temporal matrix loads.
constexpr int kSellBatch = 8;
template <class T, class X, bool NT, bool SKIP>
__device__ __forceinline__ X sell_slots(const Sell<T> &A, int64_t base, int w, const X *x,
                                        int64_t row) {
  X sum = X(0);
  for (int j0 = 0; j0 < w; j0 += kSellBatch) {
    int32_t c[kSellBatch];
    T v[kSellBatch];
#pragma unroll
    for (int u = 0; u < kSellBatch; ++u) {
      const int j = j0 + u < w ? j0 + u : w - 1;
      const int64_t e = base + static_cast<int64_t>(j) * kSlice;
      c[u] = NT ? ld_stream(A.col + e) : A.col[e];
      v[u] = NT ? ld_stream(A.val + e) : A.val[e];
    }
    X xv[kSellBatch];
#pragma unroll
    for (int u = 0; u < kSellBatch; ++u) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < kSellBatch; ++u) {
      if (j0 + u < w) {
        const X prod = static_cast<X>(v[u]) * xv[u];
        sum += (SKIP && c[u] == row) ? X(-0.0) : prod;
      }
    }
  }
  return sum;
}

template <class T, class X = T>
__device__ __forceinline__ X sell_row_dot(const Sell<T> &A, int64_t row, const X *x) {
  const int64_t s = row >> 6;
  return sell_slots<T, X, true, false>(A, A.slice_ptr[s] + (row & 63), A.slice_w[s], x, row);
}

// y(3) = A(3 rows of block row b) x, scalar-row order kept: blocks in slot
// order, within a block columns 0, 1, 2.
// The nine values of the block at `slot` for lane `lane`: nine planes of 64
// (PAIRS false), or four planes of value pairs then one plane (PAIRS true,
// gsx::bsr3_pair_planes: one 16-byte load per pair in FP64).
template <bool PAIRS, class T>
__device__ __forceinline__ void bsr3_block_values(const T *val, int64_t slot, int64_t lane, T vv[9]) {
  if constexpr (PAIRS) {
    typedef T T2 __attribute__((ext_vector_type(2)));
    const T *v = val + slot * 9;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const T2 t = ld_stream(reinterpret_cast<const T2 *>(v + q * 2 * kSlice) + lane);
      vv[2 * q] = t.x;
      vv[2 * q + 1] = t.y;
    }
    vv[8] = ld_stream(v + 8 * kSlice + lane);
  } else {
    const T *v = val + slot * 9 + lane;
#pragma unroll
    for (int e = 0; e < 9; ++e) vv[e] = ld_stream(v + e * kSlice);
  }
}

// X as in sell_row_dot; PAIRS as in bsr3_block_values.
template <class T, class X = T, bool PAIRS = false>
__device__ __forceinline__ void bsr3_row_dot(const Bsr3<T> &A, int64_t b, const X *x, X y[3],
                                             bool skip_diag) {
  const int64_t s = b >> 6;
  const int64_t lane = b & 63;
  const int64_t base = A.slice_ptr[s];
  const int w = A.slice_w[s];
  X y0 = X(0), y1 = X(0), y2 = X(0);
#pragma unroll 2
  for (int j = 0; j < w; ++j) {
    const int64_t slot = base + static_cast<int64_t>(j) * kSlice;
    const int32_t bc = ld_stream(A.bcol + slot + lane);
    if (skip_diag && bc == b) continue;
    const X x0 = x[3 * static_cast<int64_t>(bc)];
    const X x1 = x[3 * static_cast<int64_t>(bc) + 1];
    const X x2 = x[3 * static_cast<int64_t>(bc) + 2];
    T tv[9];
    bsr3_block_values<PAIRS>(A.val, slot, lane, tv);
    X vv[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) vv[e] = static_cast<X>(tv[e]);
    y0 += vv[0] * x0;
    y0 += vv[1] * x1;
    y0 += vv[2] * x2;
    y1 += vv[3] * x0;
    y1 += vv[4] * x1;
    y1 += vv[5] * x2;
    y2 += vv[6] * x0;
    y2 += vv[7] * x1;
    y2 += vv[8] * x2;
  }
  y[0] = y0;
  y[1] = y1;
  y[2] = y2;
}

__device__ __forceinline__ bool stopped(const int32_t *stop) { return stop && *stop; }
__device__ __forceinline__ int32_t stop_word(const int32_t *stop) { return stop ? *stop : 0; }

// Slice head of a row: the offset of its first slot and the slice width.
// V-cycle kernels load it together with their other row-independent
// operands and the stop word, then test the stop word, so the stop test
// costs no extra memory round trip.
struct SliceHead {
  int64_t base;
  int w;
};
template <class T>
__device__ __forceinline__ SliceHead sell_head(const Sell<T> &A, int64_t row) {
  return SliceHead{A.slice_ptr[row >> 6] + (row & 63), A.slice_w[row >> 6]};
}
template <class T>
__device__ __forceinline__ SliceHead bsr3_head(const Bsr3<T> &A, int64_t b) {
  return SliceHead{A.slice_ptr[b >> 6], A.slice_w[b >> 6]};
}
template <class T>
__device__ __forceinline__ T sell_dot_h(const Sell<T> &A, SliceHead h, const T *x) {
  return sell_slots<T, T, false, false>(A, h.base, h.w, x, -1);
}
// Off-diagonal row sum (rowDotExcluding with exclusion [row, row + 1),
// src/CPU/CPUBackend.cc:62-86).
template <class T>
__device__ __forceinline__ T sell_offdiag_h(const Sell<T> &A, SliceHead h, int64_t row,
                                            const T *x) {
  return sell_slots<T, T, false, true>(A, h.base, h.w, x, row);
}
// Row sums of block row b, slots in order. The slots are processed in
// batches (8 in FP32, 4 in FP64): the block indices of a batch are loaded
// together, then its x gathers and block values, then the products are
// added in slot order --
// two dependent memory round trips per batch instead of one per slot, with
// the arithmetic (order included) of a plain slot loop. A skipped slot
// (past the row's slice width, or the diagonal block when skip_diag) adds
// nothing.
// Keeps a loaded value live at this point: an empty asm statement that
// takes it as a register input, so the compiler cannot sink the load into a
// later conditional block (the load itself is ordinary compiled code).
template <class T>
__device__ __forceinline__ void pin(const T &v) {
  asm volatile("" ::"v"(v));
}

// NT: matrix loads non-temporal (the level-0 residual, measured 1.5 ms
// faster over the headline solve; the smoother colours and the block
// transfers measured slower with them).
template <class T, bool NT = false>
__device__ __forceinline__ void bsr3_dot_h(const Bsr3<T> &A, SliceHead h, int64_t b, const T *x,
                                           T y[3], bool skip_diag) {
  constexpr int CH = sizeof(T) == 4 ? 8 : 4;
  const int64_t lane = b & 63;
  T y0 = T(0), y1 = T(0), y2 = T(0);
  for (int j0 = 0; j0 < h.w; j0 += CH) {
    int32_t bc[CH];
    bool use[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int j = j0 + u < h.w ? j0 + u : h.w - 1;  // in-bounds address
      const int32_t *pb = A.bcol + h.base + static_cast<int64_t>(j) * kSlice + lane;
      bc[u] = NT ? ld_stream(pb) : *pb;
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) use[u] = j0 + u < h.w && !(skip_diag && bc[u] == b);
    T xv[CH][3];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int64_t c = 3 * static_cast<int64_t>(bc[u]);
      xv[u][0] = x[c];
      xv[u][1] = x[c + 1];
      xv[u][2] = x[c + 2];
    }
    T vv[CH][9];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int j = j0 + u < h.w ? j0 + u : h.w - 1;
      const T *v = A.val + (h.base + static_cast<int64_t>(j) * kSlice) * 9 + lane;
#pragma unroll
      for (int e = 0; e < 9; ++e) vv[u][e] = NT ? ld_stream(v + e * kSlice) : v[e * kSlice];
    }
#pragma unroll
    for (int u = 0; u < CH; ++u)
#pragma unroll
      for (int e = 0; e < 9; ++e) pin(vv[u][e]);
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (use[u]) {
        y0 += vv[u][0] * xv[u][0];
        y0 += vv[u][1] * xv[u][1];
        y0 += vv[u][2] * xv[u][2];
        y1 += vv[u][3] * xv[u][0];
        y1 += vv[u][4] * xv[u][1];
        y1 += vv[u][5] * xv[u][2];
        y2 += vv[u][6] * xv[u][0];
        y2 += vv[u][7] * xv[u][1];
        y2 += vv[u][8] * xv[u][2];
      }
    }
  }
  y[0] = y0;
  y[1] = y1;
  y[2] = y2;
}

// ---------------------------------------------------------------------------
// SpMV family.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_spmv(Sell<T> A, const T *__restrict__ x,
                                                 T *__restrict__ y, const int32_t *stop) {
  if (stopped(stop)) return;
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  if (row >= A.rows) return;
  y[row] = sell_row_dot(A, row, x);
}

template <class T, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_spmv_b3(Bsr3<T> A, const T *__restrict__ x,
                                                    T *__restrict__ y, const int32_t *stop) {
  if (stopped(stop)) return;
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  if (pos >= A.nb) return;
  T v[3];
  bsr3_row_dot<T, T, PAIRS>(A, pos, x, v, false);
  const int64_t b = A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
  y[3 * b] = v[0];
  y[3 * b + 1] = v[1];
  y[3 * b + 2] = v[2];
}

// Persistent launches over the SELL slices of the CG operator: the grid is
// a multiple of 8 workgroups; XCD x (workgroups x, x + 8, ...) walks slices
// [x ns / 8, (x + 1) ns / 8), its wavefronts striding by the number it
// holds, one block row per lane.
constexpr uint32_t kPersistGrid = 4096;
struct SliceWalk {
  int64_t s, end, stride;
};
__device__ __forceinline__ SliceWalk slice_walk(int64_t ns) {
  const int64_t x = blockIdx.x & 7;
  const int64_t per = gridDim.x >> 3;
  constexpr int64_t kWaves = kBlock / 64;
  return SliceWalk{x * ns / 8 + (blockIdx.x >> 3) * kWaves + (threadIdx.x >> 6), (x + 1) * ns / 8,
                   per * kWaves};
}

template <class T, bool PAIRS>
__global__ __launch_bounds__(kBlock) void k_spmv_b3_persist(Bsr3<T> A, const T *__restrict__ x,
                                                            T *__restrict__ y) {
  const int64_t lane = threadIdx.x & 63;
  const SliceWalk w = slice_walk((A.nb + 63) >> 6);
  for (int64_t s = w.s; s < w.end; s += w.stride) {
    const int64_t pos = (s << 6) + lane;
    if (pos < A.nb) {
      T v[3];
      bsr3_row_dot<T, T, PAIRS>(A, pos, x, v, false);
      const int64_t b = A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
      y[3 * b] = v[0];
      y[3 * b + 1] = v[1];
      y[3 * b + 2] = v[2];
    }
  }
}

// res = rhs - A x, as rhs + (-1) * (A x) (src/Smoothers/MultigridSmoother.cc:367-371).
template <class T>
__global__ __launch_bounds__(kBlock) void k_residual(Sell<T> A, const T *__restrict__ rhs,
                                                     const T *__restrict__ x,
                                                     T *__restrict__ res, const int32_t *stop) {
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  if (row >= A.rows) return;
  const SliceHead h = sell_head(A, row);
  const T bi = rhs[row];
  if (stop_word(stop)) return;
  res[row] = bi - sell_dot_h(A, h, x);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_residual_b3(Bsr3<T> A, const T *__restrict__ rhs,
                                                        const T *__restrict__ x,
                                                        T *__restrict__ res, const int32_t *stop) {
  const int64_t b = xcd_block() * kBlock + threadIdx.x;
  if (b >= A.nb) return;
  const SliceHead h = bsr3_head(A, b);
  const T b0 = rhs[3 * b], b1 = rhs[3 * b + 1], b2 = rhs[3 * b + 2];
  if (stop_word(stop)) return;
  T v[3];
  bsr3_dot_h<T, true>(A, h, b, x, v, false);
  res[3 * b] = b0 - v[0];
  res[3 * b + 1] = b1 - v[1];
  res[3 * b + 2] = b2 - v[2];
}

// First damped Jacobi sweep from x = 0 of one row (k_jacobi_zero, UNIT 1):
// b - A 0 == b bit for bit, so x = 0 + w ((b - 0) d^-1).
template <class T>
__device__ __forceinline__ T jacobi_zero1(T b, T di, T omega) {
  T c = b - T(0);
  c = c * di;
  return T(0) + omega * c;
}
// Line l of the first block Jacobi sweep from x = 0 (k_jacobi_zero, UNIT
// 3): x_l = 0 + w (sum_c Dinv[l][c] (b_c - 0)); d: row l of D_b^-1.
template <class T>
__device__ __forceinline__ T jacobi_zero3(const T d[3], T b0, T b1, T b2, T omega) {
  const T r0 = b0 - T(0), r1 = b1 - T(0), r2 = b2 - T(0);
  T s = T(0);
  s += d[0] * r0;
  s += d[1] * r1;
  s += d[2] * r2;
  return T(0) + omega * s;
}

// rhs_c = R res and corr_c = 0 (the coarse correction is zeroed in the same
// pass, src/Smoothers/MultigridSmoother.cc:381-388). With dinv_c (a coarse
// level smoothed by damped Jacobi from x = 0) corr_c instead receives that
// level's first sweep in closed form (jacobi_zero1 of the row just summed),
// bit for bit what the sweep would give.
template <class T>
__global__ __launch_bounds__(kBlock) void k_restrict(Sell<T> R, const T *__restrict__ res,
                                                     T *__restrict__ rhs_c, T *__restrict__ corr_c,
                                                     const int32_t *stop,
                                                     const T *__restrict__ dinv_c = nullptr,
                                                     T omega_c = T(0)) {
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  if (pos >= R.rows) return;
  const SliceHead h = sell_head(R, pos);
  const int64_t row = sell_row_of(R, pos);
  const T di = dinv_c ? dinv_c[row] : T(0);
  if (stop_word(stop)) return;
  const T v = sell_dot_h(R, h, res);
  rhs_c[row] = v;
  if (corr_c) corr_c[row] = dinv_c ? jacobi_zero1(v, di, omega_c) : T(0);
}

// x += P e (src/Smoothers/MultigridSmoother.cc:401-404).
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong_add(Sell<T> P, const T *__restrict__ e,
                                                        T *__restrict__ x, const int32_t *stop) {
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  if (pos >= P.rows) return;
  const SliceHead h = sell_head(P, pos);
  const int64_t row = sell_row_of(P, pos);
  const T xi = x[row];
  if (stop_word(stop)) return;
  const T c = sell_dot_h(P, h, e);
  x[row] = xi + T(1) * c;
}

// The reference's coarse-level correction storage: on levels below the
// finest, MultigridSmoother::vCycle prolongates into the level's own
// `correction` vector, which is also the `solution` it was called with
// (src/Smoothers/MultigridSmoother.cc:401-404), and then adds that vector to
// itself: x = P e, x = x + 1 * x.
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong_alias(Sell<T> P, const T *__restrict__ e,
                                                          T *__restrict__ x, const int32_t *stop) {
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  if (pos >= P.rows) return;
  const SliceHead h = sell_head(P, pos);
  const int64_t row = sell_row_of(P, pos);
  if (stop_word(stop)) return;
  const T c = sell_dot_h(P, h, e);
  x[row] = c + T(1) * c;
}

// 3x3-block transfer operators (both levels keep their unknowns in node
// triples): the same row sums as the scalar CSR rows, block by block in
// the rows' column order, three entries per block.
// dinv_c as in k_restrict (a coarse level smoothed by block Jacobi: 9
// values per block row).
template <class T>
__global__ __launch_bounds__(kBlock) void k_restrict_b3(Bsr3<T> R, const T *__restrict__ res,
                                                        T *__restrict__ rhs_c,
                                                        T *__restrict__ corr_c,
                                                        const int32_t *stop,
                                                        const T *__restrict__ dinv_c = nullptr,
                                                        T omega_c = T(0)) {
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  if (pos >= R.nb) return;
  const SliceHead h = bsr3_head(R, pos);
  const int64_t b = R.row_map ? static_cast<int64_t>(R.row_map[pos]) : pos;
  if (stop_word(stop)) return;
  T v[3];
  bsr3_dot_h(R, h, pos, res, v, false);
#pragma unroll
  for (int l = 0; l < 3; ++l) rhs_c[3 * b + l] = v[l];
  if (!corr_c) return;
  if (dinv_c) {
    T bi[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) bi[e] = dinv_c[9 * b + e];
#pragma unroll
    for (int l = 0; l < 3; ++l) corr_c[3 * b + l] = jacobi_zero3(bi + 3 * l, v[0], v[1], v[2], omega_c);
  } else {
#pragma unroll
    for (int l = 0; l < 3; ++l) corr_c[3 * b + l] = T(0);
  }
}

// alias = 0: x += P e; alias = 1: x = c + 1 * c with c = P e.
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong_b3(Bsr3<T> P, const T *__restrict__ e,
                                                       T *__restrict__ x, int alias,
                                                       const int32_t *stop) {
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  if (pos >= P.nb) return;
  const SliceHead h = bsr3_head(P, pos);
  const int64_t b = P.row_map ? static_cast<int64_t>(P.row_map[pos]) : pos;
  T xi[3] = {T(0), T(0), T(0)};
  if (!alias) {
    xi[0] = x[3 * b];
    xi[1] = x[3 * b + 1];
    xi[2] = x[3 * b + 2];
  }
  if (stop_word(stop)) return;
  T c[3];
  bsr3_dot_h(P, h, pos, e, c, false);
#pragma unroll
  for (int l = 0; l < 3; ++l) x[3 * b + l] = alias ? c[l] + T(1) * c[l] : xi[l] + T(1) * c[l];
}

// ---------------------------------------------------------------------------
// Smoothers.
// ---------------------------------------------------------------------------
// One colour of multicolour Gauss-Seidel over contiguous (colour-permuted)
// rows [begin, end): x_i = (1 - w) x_i + w (b_i - sum_{j != i} a_ij x_j) / a_ii
// (src/CPU/CPUBackend.cc:238-268).
//
// `reps` applies the update 1 or 2 times in a row: a symmetric sweep visits
// its last colour twice back to back (end of the forward pass, start of the
// backward pass) with no other colour in between, so the off-diagonal sum
// and b_i / a_ii are the same both times and one read of the row serves
// both updates, with the reference's operations in the reference's order.
// The diagonal comes from `diag` (the last entry of each row with column ==
// row, the value the reference's row scan finds).
// xz: the rows' own x is still the zero the sweep started from (the
// forward colours of a sweep from x = 0 on a level whose later colours were
// not zeroed); it is not read.
template <class T>
__global__ __launch_bounds__(kBlock) void k_gs_color(Sell<T> A, const T *__restrict__ diag,
                                                     int64_t begin, int64_t end,
                                                     const T *__restrict__ rhs, T *x, T omega,
                                                     int reps, const int32_t *stop, int xz = 0) {
  const int64_t row = begin + xcd_block() * kBlock + threadIdx.x;
  if (row >= end) return;
  const SliceHead h = sell_head(A, row);
  const T bi = rhs[row], di = diag[row];
  T xi = xz ? T(0) : x[row];  // not read by the off-diagonal sum
  if (stop_word(stop)) return;
  const T off = sell_offdiag_h(A, h, row, x);
  const T nv = (bi - off) / di;
  xi = (T(1) - omega) * xi + omega * nv;
  if (reps > 1) xi = (T(1) - omega) * xi + omega * nv;
  x[row] = xi;
}

// First colour of a forward sweep started from x = 0: the off-diagonal sum
// of a colour-0 row is exactly zero, so x_i = w * (b_i / a_ii) bit for bit
// without reading the row; rows [end0, n) of later colours are zeroed in the
// same pass (the reference's setZero before the V-cycle). A level whose
// forward colours read only earlier colours and not their own x (xz)
// passes n = end0: nothing reads a later colour's x before its own visit
// writes it.
template <class T>
__global__ __launch_bounds__(kBlock) void k_gs_color_zero(const T *__restrict__ diag,
                                                          int64_t end0, int64_t n,
                                                          const T *__restrict__ rhs, T *x,
                                                          T omega, const int32_t *stop) {
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  if (row >= n) return;
  T bi = T(0), di = T(1);
  if (row < end0) {
    bi = rhs[row];
    di = diag[row];
  }
  if (stop_word(stop)) return;
  if (row < end0) {
    const T nv = bi / di;
    x[row] = omega * nv;
  } else {
    x[row] = T(0);
  }
}

// One colour of 3x3 block Gauss-Seidel over contiguous block rows
// [begin, end) (src/CPU/CPUBackend.cc:350-387): r_l = b_l - sum over
// off-diagonal blocks, x_b = (1 - w) x_b + w D_b^-1 r.
// `reps` as for k_gs_color.
// Workgroup size of the thread-per-row block Gauss-Seidel colour launches
// (one wavefront, one SELL slice): a colour holds ~50k-165k block rows, and
// 64-thread workgroups spread them over the 256 CUs more evenly than
// 256-thread ones (measured: level-0 sweep 0.249 -> 0.239 ms; 128 threads
// no gain).
constexpr int kColorBlock = 64;

template <class T>
__global__ __launch_bounds__(kColorBlock) void k_block_gs_color(
    Bsr3<T> A, const T *__restrict__ dinv, int64_t begin, int64_t end, const T *__restrict__ rhs,
    T *x, T omega, int reps, const int32_t *stop, int xz = 0) {
  // Threads start at the SELL slice boundary below `begin`, so every
  // wavefront covers exactly one slice (one head, one width); rows before
  // `begin` belong to another colour and exit.
  const int64_t b = (begin & ~int64_t(kSlice - 1)) + xcd_block() * kColorBlock + threadIdx.x;
  if (b < begin || b >= end) return;
  const SliceHead h = bsr3_head(A, b);
  T bl[3], xl[3], bi[9];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    bl[l] = rhs[3 * b + l];
    xl[l] = xz ? T(0) : x[3 * b + l];  // not read by the off-block sum
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) bi[e] = dinv[9 * b + e];
  if (stop_word(stop)) return;
  T off[3];
  bsr3_dot_h(A, h, b, x, off, true);
  T lr[3];
#pragma unroll
  for (int l = 0; l < 3; ++l) lr[l] = bl[l] - off[l];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T nv = T(0);
    nv += bi[3 * l + 0] * lr[0];
    nv += bi[3 * l + 1] * lr[1];
    nv += bi[3 * l + 2] * lr[2];
    T xi = xl[l];
    xi = (T(1) - omega) * xi + omega * nv;
    if (reps > 1) xi = (T(1) - omega) * xi + omega * nv;
    x[3 * b + l] = xi;
  }
}

// Block GS colour started from x = 0: r = b, x_b = w D_b^-1 b for the
// blocks of colour 0, zero for the rest.
template <class T>
__global__ __launch_bounds__(kBlock) void k_block_gs_color_zero(const T *__restrict__ dinv,
                                                                int64_t end0, int64_t nb,
                                                                const T *__restrict__ rhs, T *x,
                                                                T omega, const int32_t *stop) {
  const int64_t b = xcd_block() * kBlock + threadIdx.x;
  if (b >= nb) return;
  if (b >= end0) {
    if (stop_word(stop)) return;
    x[3 * b] = T(0);
    x[3 * b + 1] = T(0);
    x[3 * b + 2] = T(0);
    return;
  }
  T bi[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) bi[e] = dinv[9 * b + e];
  const T r0 = rhs[3 * b], r1 = rhs[3 * b + 1], r2 = rhs[3 * b + 2];
  if (stop_word(stop)) return;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T nv = T(0);
    nv += bi[3 * l + 0] * r0;
    nv += bi[3 * l + 1] * r1;
    nv += bi[3 * l + 2] * r2;
    x[3 * b + l] = omega * nv;
  }
}

// Damped Jacobi sweep into a second buffer:
// xo = x + w * (D^-1 (b + (-1) A x)) (src/Smoothers/DampedJacobiSmoother.cc:79-87).
template <class T>
__global__ __launch_bounds__(kBlock) void k_jacobi(Sell<T> A, const T *__restrict__ dinv,
                                                   const T *__restrict__ rhs,
                                                   const T *__restrict__ x, T *__restrict__ xo,
                                                   T omega, const int32_t *stop) {
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  if (row >= A.rows) return;
  const SliceHead h = sell_head(A, row);
  const T bi = rhs[row], di = dinv[row], xi = x[row];
  if (stop_word(stop)) return;
  T c = bi - sell_dot_h(A, h, x);
  c = c * di;
  xo[row] = xi + omega * c;
}

// First Jacobi sweep from x = 0 (damped, scalar or 3x3 block): every
// product a_ij * 0 is +-0 and a row sum that starts at +0 never is -0, so
// b - A x == b bit for bit and the sweep is x = 0 + w * (D^-1 b) with the
// reference's operations in the reference's order -- no matrix read, no
// second buffer. unit = 1 (scalar, dinv one per row) or 3 (dinv 9 per
// block row).
template <class T, int UNIT>
__global__ __launch_bounds__(kBlock) void k_jacobi_zero(const T *__restrict__ dinv,
                                                        const T *__restrict__ rhs, T *x,
                                                        int64_t units, T omega,
                                                        const int32_t *stop) {
  const int64_t u = xcd_block() * kBlock + threadIdx.x;
  if (u >= units) return;
  if (UNIT == 1) {
    const T bi = rhs[u], di = dinv[u];
    if (stop_word(stop)) return;
    T c = bi - T(0);
    c = c * di;
    x[u] = T(0) + omega * c;
  } else {
    T r[3], bi[9];
#pragma unroll
    for (int l = 0; l < 3; ++l) r[l] = rhs[3 * u + l];
#pragma unroll
    for (int e = 0; e < 9; ++e) bi[e] = dinv[9 * u + e];
    if (stop_word(stop)) return;
#pragma unroll
    for (int l = 0; l < 3; ++l) r[l] = r[l] - T(0);
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      T s = T(0);
      s += bi[3 * l + 0] * r[0];
      s += bi[3 * l + 1] * r[1];
      s += bi[3 * l + 2] * r[2];
      x[3 * u + l] = T(0) + omega * s;
    }
  }
}

// 3x3 block Jacobi sweep into a second buffer
// (src/Smoothers/BlockJacobiSmoother.cc:17-87).
template <class T>
__global__ __launch_bounds__(kBlock) void k_block_jacobi(Sell<T> A, const T *__restrict__ dinv,
                                                         const T *__restrict__ rhs,
                                                         const T *__restrict__ x,
                                                         T *__restrict__ xo, T omega,
                                                         const int32_t *stop) {
  if (stopped(stop)) return;
  const int64_t b = xcd_block() * kBlock + threadIdx.x;
  if (3 * b >= A.rows) return;
  T r[3];
#pragma unroll
  for (int l = 0; l < 3; ++l) r[l] = rhs[3 * b + l] - sell_row_dot(A, 3 * b + l, x);
  const T *bi = dinv + 9 * b;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T s = T(0);
    s += bi[3 * l + 0] * r[0];
    s += bi[3 * l + 1] * r[1];
    s += bi[3 * l + 2] * r[2];
    xo[3 * b + l] = x[3 * b + l] + omega * s;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_block_jacobi_b3(Bsr3<T> A, const T *__restrict__ dinv,
                                                            const T *__restrict__ rhs,
                                                            const T *__restrict__ x,
                                                            T *__restrict__ xo, T omega,
                                                            const int32_t *stop) {
  const int64_t b = xcd_block() * kBlock + threadIdx.x;
  if (b >= A.nb) return;
  const SliceHead h = bsr3_head(A, b);
  T bl[3], bi[9];
#pragma unroll
  for (int l = 0; l < 3; ++l) bl[l] = rhs[3 * b + l];
#pragma unroll
  for (int e = 0; e < 9; ++e) bi[e] = dinv[9 * b + e];
  if (stop_word(stop)) return;
  T v[3];
  bsr3_dot_h(A, h, b, x, v, false);
  T r[3];
#pragma unroll
  for (int l = 0; l < 3; ++l) r[l] = bl[l] - v[l];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T s = T(0);
    s += bi[3 * l + 0] * r[0];
    s += bi[3 * l + 1] * r[1];
    s += bi[3 * l + 2] * r[2];
    xo[3 * b + l] = x[3 * b + l] + omega * s;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_copy(const T *__restrict__ x, T *__restrict__ y,
                                                 int64_t n, const int32_t *stop) {
  if (stopped(stop)) return;
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i < n) y[i] = x[i];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill(T *__restrict__ y, T v, int64_t n,
                                                 const int32_t *stop) {
  if (stopped(stop)) return;
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i < n) y[i] = v;
}

// Coarsest level: x = A_c^-1 rhs with the symmetric FP64 dense inverse,
// stored as its lower 64 x 64 tiles only: tile (I, J), I >= J,
// at tiles + (I (I + 1) / 2 + J) * 4096, row-major. Each tile is read once
// and applied twice: its row sums T x_J go to y_I and, off the diagonal, its
// column sums T^T x_I to y_J. Workgroups take strips of kSymStrip tiles of
// one tile column J (tile rows I0..I1); the partial sums land in
//   prow[(I (I + 1) / 2 + J) 64 + r]   row sums of tile (I, J),
//   pcol[(J kmax + s) 64 + c]          column sums of strip s of column J,
// and k_dense_sym_reduce adds them per output block in a fixed order.
constexpr int kSymStrip = 4;

// Sums 16 row-vectors p[0..15] across the 64 lanes of a wavefront: each
// halving step keeps half of the vectors in a lane and adds the partner
// lane's copy of them. Lane l ends with the sum of vector l >> 2 when
// (l & 3) == 0.
__device__ __forceinline__ double wave_rowsums16(const double p[16]) {
  const int lane = threadIdx.x & 63;
  const bool u5 = lane & 32, u4 = lane & 16, u3 = lane & 8, u2 = lane & 4;
  double q8[8], q4[4], q2[2];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double send = u5 ? p[i] : p[8 + i];
    const double keep = u5 ? p[8 + i] : p[i];
    q8[i] = keep + __shfl_xor(send, 32, 64);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double send = u4 ? q8[i] : q8[4 + i];
    const double keep = u4 ? q8[4 + i] : q8[i];
    q4[i] = keep + __shfl_xor(send, 16, 64);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double send = u3 ? q4[i] : q4[2 + i];
    const double keep = u3 ? q4[2 + i] : q4[i];
    q2[i] = keep + __shfl_xor(send, 8, 64);
  }
  const double send = u2 ? q2[0] : q2[1];
  const double keep = u2 ? q2[1] : q2[0];
  double r = keep + __shfl_xor(send, 4, 64);
  r += __shfl_xor(r, 2, 64);
  r += __shfl_xor(r, 1, 64);
  return r;
}

// One workgroup per strip; strips[3 s .. 3 s + 2] = (J, I0, I1). Wave w
// holds rows w, w + 4, ..., w + 60 of each tile, one column per lane, so
// every load is one contiguous 512-byte tile row.
template <class T>
__global__ __launch_bounds__(kBlock) void k_dense_sym_strip(const double *__restrict__ tiles,
                                                            const int32_t *__restrict__ strips,
                                                            int64_t n, int64_t nt, int64_t kmax,
                                                            const T *__restrict__ rhs,
                                                            double *__restrict__ prow,
                                                            double *__restrict__ pcol,
                                                            const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ double csum[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wu = __builtin_amdgcn_readfirstlane(w);
  const int64_t J = strips[3 * blockIdx.x], I0 = strips[3 * blockIdx.x + 1],
                I1 = strips[3 * blockIdx.x + 2];
  // x entries past n are zero (clamped in-bounds load, then select).
  const int64_t gj = 64 * J + lane;
  const double xjv = static_cast<double>(rhs[gj < n ? gj : n - 1]);
  const double xj = gj < n ? xjv : 0.0;
  double v = 0.0;
  for (int64_t I = I0; I <= I1; ++I) {
    const double *t = tiles + (I * (I + 1) / 2 + J) * 4096;
    double a[16], xi[16], p[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = ld_stream(t + (w + 4 * k) * 64 + lane);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int64_t gi = 64 * I + wu + 4 * k;  // wave-uniform: scalar loads
      const double xv = static_cast<double>(rhs[gi < n ? gi : n - 1]);
      xi[k] = gi < n ? xv : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) p[k] = a[k] * xj;
    if (I != J) {
#pragma unroll
      for (int k = 0; k < 16; ++k) v += a[k] * xi[k];
    }
    const double rs = wave_rowsums16(p);
    if ((lane & 3) == 0) prow[(I * (I + 1) / 2 + J) * 64 + w + 4 * (lane >> 2)] = rs;
  }
  csum[w][lane] = v;
  __syncthreads();
  if (w == 0)
    pcol[(J * kmax + (I0 - J) / kSymStrip) * 64 + lane] =
        (csum[0][lane] + csum[1][lane]) + (csum[2][lane] + csum[3][lane]);
}

// y[64 I + r] = sum over J <= I of prow[I][J][r] plus the column-sum
// strips of tile column I, 16 waves per block I, combined in a fixed order.
template <class T>
__global__ __launch_bounds__(1024) void k_dense_sym_reduce(const double *__restrict__ prow,
                                                           const double *__restrict__ pcol,
                                                           int64_t n, int64_t nt, int64_t kmax,
                                                           T *__restrict__ y,
                                                           const int32_t *stop) {
  if (stopped(stop)) return;
  __shared__ double accw[16][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t I = blockIdx.x;
  const int64_t ncs = (nt - I + kSymStrip - 1) / kSymStrip;
  double acc = 0.0;
#pragma unroll 4
  for (int64_t J = w; J <= I; J += 16) acc += prow[(I * (I + 1) / 2 + J) * 64 + lane];
  for (int64_t k = w; k < ncs; k += 16) acc += pcol[(I * kmax + k) * 64 + lane];
  accw[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) t += accw[u][lane];
    const int64_t row = I * 64 + lane;
    if (row < n) y[row] = static_cast<T>(t);
  }
}

// ---------------------------------------------------------------------------
// PCG kernels (src/CGSolver.cc:208-265), CG type TC, V-cycle type TM.
// ---------------------------------------------------------------------------
// Initial residual r = b + (-1) A x (src/CGSolver.cc:134-140) and, for a
// V-cycle in another precision, its rounded copy.
template <class TC, class TM>
__global__ __launch_bounds__(kBlock) void k_cg_init_r(Sell<TC> A, const TC *__restrict__ b,
                                                      const TC *__restrict__ x, TC *__restrict__ r,
                                                      TM *__restrict__ rm) {
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  if (row >= A.rows) return;
  const TC v = b[row] - sell_row_dot(A, row, x);
  r[row] = v;
  if (rm) rm[row] = static_cast<TM>(v);
}

template <class TC, class TM, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_cg_init_r_b3(Bsr3<TC> A, const TC *__restrict__ b,
                                                         const TC *__restrict__ x,
                                                         TC *__restrict__ r, TM *__restrict__ rm) {
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  if (pos >= A.nb) return;
  TC v[3];
  bsr3_row_dot<TC, TC, PAIRS>(A, pos, x, v, false);
  const int64_t bb = A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const TC t = b[3 * bb + l] - v[l];
    r[3 * bb + l] = t;
    if (rm) rm[3 * bb + l] = static_cast<TM>(t);
  }
}

// Initial scalars: bb = b.b, rho = r.z, rr = r.r, rel = rr / bb, p = z
// (src/CGSolver.cc:148-176). z may be in the V-cycle type.
template <class TC, class TM>
__global__ __launch_bounds__(kBlock) void k_cg_init_dots(const TC *__restrict__ b,
                                                         const TC *__restrict__ r,
                                                         const TM *__restrict__ z,
                                                         TC *__restrict__ p, int64_t n,
                                                         Ctl<TC> *ctl, TC *partials) {
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  TC sbb = TC(0), srz = TC(0), srr = TC(0);
  if (i < n) {
    const TC zi = static_cast<TC>(z[i]);
    const TC ri = r[i];
    sbb = b[i] * b[i];
    srz = ri * zi;
    srr = ri * ri;
    p[i] = zi;
  }
  const uint32_t G = gridDim.x;
  block_partial(sbb, partials);
  block_partial(srz, partials + G);
  block_partial(srr, partials + 2 * G);
}

// One block: bb, rho, rr from the partials of k_cg_init_dots (G each).
template <class TC>
__global__ __launch_bounds__(kBlock) void k_cg_init_final(Ctl<TC> *ctl, const TC *partials,
                                                          uint32_t G) {
  const TC a = fixed_sum(partials, G);
  const TC c = fixed_sum(partials + G, G);
  const TC d = fixed_sum(partials + 2 * G, G);
  if (threadIdx.x == 0) {
    ctl->bb = a;
    ctl->rho = c;
    ctl->rr = d;
    const TC rel = d / a;
    ctl->rel = rel;
    if (ctl->hist) ctl->hist[0] = rel;
    ctl->iter = 0;
    ctl->iters_out = ctl->maxit;
    ctl->breakdown = (c <= TC(0)) ? 2 : 0;
    if (!finite_val(rel)) {
      ctl->stop = 3;
    } else if (rel < ctl->tol2) {
      ctl->stop = 1;
      ctl->iters_out = 0;
    } else {
      ctl->stop = 0;
    }
  }
}

// q = A p with per-block partials of p.q; k_cg_alpha then forms
// alpha = rho / pAp (src/CGSolver.cc:210-215).
template <class TC>
__global__ __launch_bounds__(kBlock) void k_cg_alpha(Ctl<TC> *ctl, const TC *partials,
                                                     uint32_t G) {
  if (ctl->stop) return;
  const TC pap = fixed_sum(partials, G);
  if (threadIdx.x == 0) {
    ctl->pap = pap;
    ctl->alpha = ctl->rho / pap;
    if (!(pap > TC(0)) && ctl->breakdown == 0) ctl->breakdown = 1;
  }
}

// TA: the stored operator's type (TC, or the V-cycle type with products and
// sums in TC).
template <class TC, class TA = TC>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_pap(Sell<TA> A, const TC *__restrict__ p,
                                                        TC *__restrict__ q, Ctl<TC> *ctl,
                                                        TC *partials) {
  if (ctl->stop) return;
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  TC v = TC(0);
  if (row < A.rows) {
    const TC qi = sell_row_dot(A, row, p);
    q[row] = qi;
    v = p[row] * qi;
  }
  block_partial(v, partials);
}

template <class TC, class TA = TC>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_pap_b3(Bsr3<TA> A, const TC *__restrict__ p,
                                                           TC *__restrict__ q, Ctl<TC> *ctl,
                                                           TC *partials) {
  if (ctl->stop) return;
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  TC v = TC(0);
  if (pos < A.nb) {
    TC y[3];
    bsr3_row_dot(A, pos, p, y, false);
    const int64_t b = A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
    q[3 * b] = y[0];
    q[3 * b + 1] = y[1];
    q[3 * b + 2] = y[2];
    v = p[3 * b] * y[0];
    v += p[3 * b + 1] * y[1];
    v += p[3 * b + 2] * y[2];
  }
  block_partial(v, partials);
}

// Persistent form of k_cg_spmv_pap_b3 for the CG operator (one partial per
// workgroup, gridDim.x partials).
template <class TC, bool PAIRS>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_pap_b3_persist(Bsr3<TC> A,
                                                                   const TC *__restrict__ p,
                                                                   TC *__restrict__ q,
                                                                   Ctl<TC> *ctl, TC *partials) {
  if (ctl->stop) return;
  const int64_t lane = threadIdx.x & 63;
  const SliceWalk w = slice_walk((A.nb + 63) >> 6);
  TC v = TC(0);
  for (int64_t s = w.s; s < w.end; s += w.stride) {
    const int64_t pos = (s << 6) + lane;
    if (pos < A.nb) {
      TC y[3];
      bsr3_row_dot<TC, TC, PAIRS>(A, pos, p, y, false);
      const int64_t b = A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
      q[3 * b] = y[0];
      q[3 * b + 1] = y[1];
      q[3 * b + 2] = y[2];
      v += p[3 * b] * y[0];
      v += p[3 * b + 1] * y[1];
      v += p[3 * b + 2] * y[2];
    }
  }
  block_partial(v, partials);
}

// x += alpha p; r = r + (-alpha) q; partials of r.r (k_cg_rr_final then
// forms rel = rr / bb and the stop test, src/CGSolver.cc:218-230, 252-268).
// Writes the V-cycle-type copy of r.
template <class TC, class TM>
__global__ __launch_bounds__(kBlock) void k_cg_update_xr(TC *__restrict__ x, TC *__restrict__ r,
                                                         TM *__restrict__ rm,
                                                         const TC *__restrict__ p,
                                                         const TC *__restrict__ q, int64_t n,
                                                         Ctl<TC> *ctl, TC *partials) {
  if (ctl->stop) return;
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  const TC alpha = ctl->alpha;
  const TC na = -alpha;
  TC v = TC(0);
  if (i < n) {
    x[i] = x[i] + alpha * p[i];
    const TC ri = r[i] + na * q[i];
    r[i] = ri;
    if (rm) rm[i] = static_cast<TM>(ri);
    v = ri * ri;
  }
  block_partial(v, partials);
}

template <class TC>
__global__ __launch_bounds__(kBlock) void k_cg_rr_final(Ctl<TC> *ctl, const TC *partials,
                                                        uint32_t G) {
  if (ctl->stop) return;
  const TC rr = fixed_sum(partials, G);
  if (threadIdx.x == 0) {
    const TC rel = rr / ctl->bb;
    ctl->rr = rr;
    ctl->rel = rel;
    const int64_t it = ctl->iter;
    if (ctl->hist) ctl->hist[it + 1] = rel;
    ctl->iter = it + 1;
    const int64_t every = ctl->replace_every;
    if (!finite_val(rel)) {
      ctl->stop = 3;
    } else if (rel < ctl->tol2 && every == 0) {
      ctl->stop = 1;
      ctl->iters_out = it;
    } else if (rel < ctl->tol2) {
      ctl->stop = 5;  // confirm with the replaced residual
    } else if (it + 1 >= ctl->maxit) {
      ctl->stop = 2;
      ctl->iters_out = ctl->maxit;
    }
    if (every > 0 && (it + 1) % every == 0) ctl->replace = 1;
  }
}

// Residual replacement, when ctl->replace is set: r = b - A x with the
// full-precision operator, its V-cycle-type copy, and partials of r.r.
template <class TC, class TM>
__global__ __launch_bounds__(kBlock) void k_cg_replace(Sell<TC> A, const TC *__restrict__ b,
                                                       const TC *__restrict__ x,
                                                       TC *__restrict__ r, TM *__restrict__ rm,
                                                       const Ctl<TC> *ctl, TC *partials) {
  if (!replace_runs(ctl)) return;
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  TC v = TC(0);
  if (row < A.rows) {
    const TC t = b[row] - sell_row_dot(A, row, x);
    r[row] = t;
    if (rm) rm[row] = static_cast<TM>(t);
    v = t * t;
  }
  block_partial(v, partials);
}

template <class TC, class TM, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_cg_replace_b3(Bsr3<TC> A, const TC *__restrict__ b,
                                                          const TC *__restrict__ x,
                                                          TC *__restrict__ r, TM *__restrict__ rm,
                                                          const Ctl<TC> *ctl, TC *partials) {
  if (!replace_runs(ctl)) return;
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  TC v = TC(0);
  if (pos < A.nb) {
    TC y[3];
    bsr3_row_dot<TC, TC, PAIRS>(A, pos, x, y, false);
    const int64_t bb = A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const TC t = b[3 * bb + l] - y[l];
      r[3 * bb + l] = t;
      if (rm) rm[3 * bb + l] = static_cast<TM>(t);
      v += t * t;
    }
  }
  block_partial(v, partials);
}

// rr, rel and the stop test from the replaced residual (the iteration was
// counted by k_cg_rr_final; its history entry is overwritten). A pending
// convergence check that fails resumes the iteration (stop back to 0).
template <class TC>
__global__ __launch_bounds__(kBlock) void k_cg_replace_final(Ctl<TC> *ctl, const TC *partials,
                                                             uint32_t G) {
  if (!replace_runs(ctl)) return;
  const TC rr = fixed_sum(partials, G);
  if (threadIdx.x == 0) {
    const TC rel = rr / ctl->bb;
    const int64_t it = ctl->iter;  // iterations done
    ctl->rr = rr;
    ctl->rel = rel;
    ctl->replace = 0;
    if (ctl->hist) ctl->hist[it] = rel;
    if (!finite_val(rel)) {
      ctl->stop = 3;
    } else if (rel < ctl->tol2) {
      ctl->stop = 1;
      ctl->iters_out = it - 1;
    } else if (it >= ctl->maxit) {
      ctl->stop = 2;
      ctl->iters_out = ctl->maxit;
    } else {
      ctl->stop = 0;
    }
  }
}

// Partials of r.z; k_cg_beta forms rho_new and beta = rho_new / rho
// (src/CGSolver.cc:236-239).
template <class TC, class TM>
__global__ __launch_bounds__(kBlock) void k_cg_rz(const TC *__restrict__ r,
                                                  const TM *__restrict__ z, int64_t n,
                                                  Ctl<TC> *ctl, TC *partials) {
  if (ctl->stop) return;
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  TC v = TC(0);
  if (i < n) v = r[i] * static_cast<TC>(z[i]);
  block_partial(v, partials);
}

template <class TC>
__global__ __launch_bounds__(kBlock) void k_cg_beta(Ctl<TC> *ctl, const TC *partials, uint32_t G) {
  if (ctl->stop) return;
  const TC rz = fixed_sum(partials, G);
  if (threadIdx.x == 0) {
    ctl->rho_new = rz;
    ctl->beta = rz / ctl->rho;
    ctl->rho = rz;
    if (!(rz > TC(0)) && ctl->breakdown == 0) ctl->breakdown = 2;
  }
}

// p = beta p + 1 * z (src/CGSolver.cc:241-245).
template <class TC, class TM>
__global__ __launch_bounds__(kBlock) void k_cg_update_p(TC *__restrict__ p,
                                                        const TM *__restrict__ z, int64_t n,
                                                        const Ctl<TC> *ctl) {
  if (ctl->stop) return;
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i >= n) return;
  const TC beta = ctl->beta;
  const TC pi = p[i] * beta;
  p[i] = pi + TC(1) * static_cast<TC>(z[i]);
}

// True residual ||b - A x||^2 and ||b||^2 in FP64 (verification only).
template <class TC>
__global__ __launch_bounds__(kBlock) void k_true_residual(Sell<TC> A, const TC *__restrict__ b,
                                                          const TC *__restrict__ x,
                                                          double *partials) {
  const int64_t row = xcd_block() * kBlock + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (row < A.rows) {
    const int64_t s = row >> 6;
    const int64_t base = A.slice_ptr[s] + (row & 63);
    const int w = A.slice_w[s];
    double sum = 0.0;
    for (int j = 0; j < w; ++j) {
      const int64_t e = base + static_cast<int64_t>(j) * kSlice;
      sum += static_cast<double>(A.val[e]) * static_cast<double>(x[A.col[e]]);
    }
    const double bi = static_cast<double>(b[row]);
    const double d = bi - sum;
    rr = d * d;
    bb = bi * bi;
  }
  block_partial(rr, partials);
  block_partial(bb, partials + gridDim.x);
}

template <class TC, bool PAIRS = false>
__global__ __launch_bounds__(kBlock) void k_true_residual_b3(Bsr3<TC> A, const TC *__restrict__ b,
                                                             const TC *__restrict__ x,
                                                             double *partials) {
  const int64_t pos = xcd_block() * kBlock + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (pos < A.nb) {
    const int64_t blk = A.row_map ? static_cast<int64_t>(A.row_map[pos]) : pos;
    const int64_t s = pos >> 6;
    const int64_t lane = pos & 63;
    const int w = A.slice_w[s];
    double y[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < w; ++j) {
      const int64_t slot = A.slice_ptr[s] + static_cast<int64_t>(j) * kSlice;
      const int64_t bc = A.bcol[slot + lane];
      TC v[9];
      bsr3_block_values<PAIRS>(A.val, slot, lane, v);
      for (int l = 0; l < 3; ++l)
        for (int c = 0; c < 3; ++c)
          y[l] += static_cast<double>(v[3 * l + c]) * static_cast<double>(x[3 * bc + c]);
    }
    for (int l = 0; l < 3; ++l) {
      const double bi = static_cast<double>(b[3 * blk + l]);
      const double d = bi - y[l];
      rr += d * d;
      bb += bi * bi;
    }
  }
  block_partial(rr, partials);
  block_partial(bb, partials + gridDim.x);
}

// One block: out2 = {||b - A x||^2, ||b||^2} from the partials (G each).
template <class T = double>
__global__ __launch_bounds__(kBlock) void k_true_final(const T *partials, uint32_t G, T *out2) {
  const T a = fixed_sum(partials, G);
  const T c = fixed_sum(partials + G, G);
  if (threadIdx.x == 0) {
    out2[0] = a;
    out2[1] = c;
  }
}

inline uint32_t grid_for(int64_t n) {
  return static_cast<uint32_t>((n + kBlock - 1) / kBlock);
}

}  // namespace gsk
