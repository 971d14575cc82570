// GPU side of the smoothed-aggregation setup: the sparse products of one
// coarsening step (P = M P_tent, R = P^T, A_c = R (A P)), which dominate the
// parallel part of building a hierarchy. The reference obtains its
// hierarchy from amgcl (examples/amgSolver/src/SmoothedAggregationCoarsener.cc:
// 93-97: transfer_operators, then coarse_operator); the host restatement is
// gsx::GalerkinOps. The results here equal the host's bit for bit:
//
//  * SpGEMM C = A B gives every output column the host's Gustavson sum:
//    A's row entries in order, each one's B row in order; a column's first
//    product starts its sum and later products are added in that order. A
//    symbolic pass counts each row's distinct columns first. Rows whose A
//    row and column count fit (at most 1024 each) run one wavefront per row
//    with the column table, the sorted columns and the sums in LDS; the
//    others run one thread per row with open-addressing tables in a
//    scratch buffer (rows processed in chunks that fit it) and a heap sort
//    of the finished row.
//  * Transpose: entries scattered into their column's segment by atomic
//    slot, then every segment sorted by source row, which restores the
//    host's stable (ascending row) order; the pairs are unique, so the order
//    is the same whatever order the atomics handed out the slots in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "device.hpp"
#include "host/setup.h"
#include "setup_device.hpp"

namespace gsdev {

using gsd::DevBuf;

namespace {

constexpr int kThreads = 256;
constexpr int32_t kEmpty = -1;
// Scratch for the per-row column tables (bytes); rows are processed in
// chunks whose tables fit.
constexpr size_t kScratchBytes = size_t(1) << 31;

uint32_t blocks_for(int64_t n) {
  return static_cast<uint32_t>(std::max<int64_t>(1, (n + kThreads - 1) / kThreads));
}

}  // namespace

DCsr upload(const gsx::Csr &A) {
  DCsr d;
  d.rows = A.rows;
  d.cols = A.cols;
  d.nnz = A.nnz();
  d.ptr.upload(A.ptr);
  d.col.upload(A.col.data(), A.col.size());
  d.val.upload(A.val.data(), A.val.size());
  return d;
}

namespace {

gsx::Csr download(const DCsr &d) {
  gsx::Csr A;
  A.rows = d.rows;
  A.cols = d.cols;
  A.ptr.resize(d.rows + 1);
  A.col.resize(d.nnz);
  A.val.resize(d.nnz);
  gsx::touch_pages(A.col.data(), A.col.size() * sizeof(int32_t));
  gsx::touch_pages(A.val.data(), A.val.size() * sizeof(double));
  GS_HIP(hipMemcpy(A.ptr.data(), d.ptr.get(), (d.rows + 1) * sizeof(int32_t),
                   hipMemcpyDeviceToHost));
  if (d.nnz) {
    GS_HIP(hipMemcpy(A.col.data(), d.col.get(), d.nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    GS_HIP(hipMemcpy(A.val.data(), d.val.get(), d.nnz * sizeof(double), hipMemcpyDeviceToHost));
  }
  return A;
}

__device__ __forceinline__ uint32_t hash_col(int32_t c, uint32_t mask) {
  return (static_cast<uint32_t>(c) * 2654435761u) & mask;
}

// In-place heapsort of keys[0, n) carrying vals (keys unique).
__device__ void heap_sort(int32_t *keys, double *vals, int64_t n) {
  auto sift = [&](int64_t root, int64_t end) {
    for (;;) {
      int64_t child = 2 * root + 1;
      if (child >= end) return;
      if (child + 1 < end && keys[child + 1] > keys[child]) ++child;
      if (keys[root] >= keys[child]) return;
      const int32_t tk = keys[root];
      keys[root] = keys[child];
      keys[child] = tk;
      const double tv = vals[root];
      vals[root] = vals[child];
      vals[child] = tv;
      root = child;
    }
  };
  for (int64_t s = n / 2 - 1; s >= 0; --s) sift(s, n);
  for (int64_t e = n - 1; e > 0; --e) {
    const int32_t tk = keys[0];
    keys[0] = keys[e];
    keys[e] = tk;
    const double tv = vals[0];
    vals[0] = vals[e];
    vals[e] = tv;
    sift(0, e);
  }
}

// Products per row of A B.
__global__ void k_products(const int32_t *ap, const int32_t *ac, const int32_t *bp, int64_t rows,
                           int64_t *np) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= rows) return;
  int64_t s = 0;
  for (int32_t k = ap[r]; k < ap[r + 1]; ++k) s += bp[ac[k] + 1] - bp[ac[k]];
  np[r] = s;
}

// Distinct columns of rows [r0, r1) of A B; the table of row r holds
// tsize[r] int32 keys at toff[r] (power of two, at least twice the
// product count or the column count, whichever is smaller).
__global__ void k_symbolic(const int32_t *ap, const int32_t *ac, const int32_t *bp,
                           const int32_t *bc, int64_t r0, int64_t r1, const int64_t *toff,
                           const int32_t *tsize, int32_t *table, int32_t *rowlen) {
  const int64_t r = r0 + static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= r1) return;
  const uint32_t H = static_cast<uint32_t>(tsize[r]);
  if (!H) return;  // a row without products, or one the LDS path counts
  int32_t cnt = 0;
  {
    int32_t *t = table + toff[r];
    for (uint32_t i = 0; i < H; ++i) t[i] = kEmpty;
    const uint32_t mask = H - 1;
    for (int32_t k = ap[r]; k < ap[r + 1]; ++k) {
      const int32_t j = ac[k];
      for (int32_t kb = bp[j]; kb < bp[j + 1]; ++kb) {
        const int32_t c = bc[kb];
        uint32_t h = hash_col(c, mask);
        while (t[h] != kEmpty && t[h] != c) h = (h + 1) & mask;
        if (t[h] == kEmpty) {
          t[h] = c;
          ++cnt;
        }
      }
    }
  }
  rowlen[r] = cnt;
}

// Values of rows [r0, r1) of C = A B in the host's order (see the file
// comment); the table of row r holds tsize[r] (key, slot) pairs at
// 2 * toff[r].
__global__ void k_numeric(const int32_t *ap, const int32_t *ac, const double *av,
                          const int32_t *bp, const int32_t *bc, const double *bv, int64_t r0,
                          int64_t r1, const int64_t *toff, const int32_t *tsize, int32_t *table,
                          const int32_t *cp, int32_t *cc, double *cv) {
  const int64_t r = r0 + static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= r1) return;
  const uint32_t H = static_cast<uint32_t>(tsize[r]);
  if (!H) return;
  int32_t *t = table + 2 * toff[r];
  for (uint32_t i = 0; i < H; ++i) t[2 * i] = kEmpty;
  const uint32_t mask = H - 1;
  const int64_t base = cp[r];
  int32_t cnt = 0;
  for (int32_t k = ap[r]; k < ap[r + 1]; ++k) {
    const int32_t j = ac[k];
    const double a = av[k];
    for (int32_t kb = bp[j]; kb < bp[j + 1]; ++kb) {
      const int32_t c = bc[kb];
      uint32_t h = hash_col(c, mask);
      while (t[2 * h] != kEmpty && t[2 * h] != c) h = (h + 1) & mask;
      if (t[2 * h] == kEmpty) {
        t[2 * h] = c;
        t[2 * h + 1] = cnt;
        cc[base + cnt] = c;
        cv[base + cnt] = a * bv[kb];
        ++cnt;
      } else {
        cv[base + t[2 * h + 1]] += a * bv[kb];
      }
    }
  }
  heap_sort(cc + base, cv + base, cnt);
}

// ---------------------------------------------------------------------------
// One wavefront per output row, tables in LDS: the path for rows whose
// A row (NCAP entries) and distinct columns (DCAP) fit. The row's products
// are numbered t = 0 .. np-1 in the host's loop order (A entry k, then its
// B row entry); pref[k] is the number of products before A entry k and
// bst[k] the start of its B row, so lane t finds (k, kb) by a binary
// search. Within one A entry the columns are distinct (a B row holds each
// column once), so the products of one k can be added in parallel; the
// A entries are applied one after the other, which is the host's order for
// every column. A sum starts at -0.0, the additive identity: -0 + v == v
// bit for bit for every v (signed zeros, infinities and NaN included), the
// same as the host's "first product starts the sum".
// ---------------------------------------------------------------------------
constexpr int kRowWave = 64;

// Fills pref[0 .. na] and bst[0 .. na) for the A row (columns ac_row,
// na <= NCAP); returns the product count.
__device__ int32_t row_products(const int32_t *ac_row, int32_t na, const int32_t *bp,
                                int32_t *pref, int32_t *bst) {
  const int lane = threadIdx.x;
  int32_t carry = 0;
  if (lane == 0) pref[0] = 0;
  for (int32_t i0 = 0; i0 < na; i0 += kRowWave) {
    const int32_t i = i0 + lane;
    int32_t len = 0, st = 0;
    if (i < na) {
      const int32_t j = ac_row[i];
      st = bp[j];
      len = bp[j + 1] - st;
    }
    int32_t s = len;
#pragma unroll
    for (int d = 1; d < kRowWave; d <<= 1) {
      const int32_t u = __shfl_up(s, d, kRowWave);
      if (lane >= d) s += u;
    }
    if (i < na) {
      pref[i + 1] = carry + s;
      bst[i] = st;
    }
    carry += __shfl(s, kRowWave - 1, kRowWave);
  }
  __syncthreads();
  return carry;
}

// The A entry k of product t: pref[k] <= t < pref[k + 1] (pref[0] = 0,
// pref[na] = np > t; empty B rows are passed over).
__device__ __forceinline__ int32_t product_entry(const int32_t *pref, int32_t na, int32_t t) {
  int32_t lo = 0, hi = na;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (pref[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
  for (int d = kRowWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kRowWave);
  return v;
}

// Distinct columns of the listed rows of A B. Stops a row as soon as more
// than DCAP columns were found and reports -1 (the row goes to k_symbolic);
// TS > DCAP + 64 keeps a free slot for every probe of a round.
template <int TS, int DCAP, int NCAP>
__global__ __launch_bounds__(kRowWave) void k_symbolic_lds(const int32_t *ap, const int32_t *ac,
                                                           const int32_t *bp, const int32_t *bc,
                                                           const int32_t *rows, int32_t *rowlen) {
  static_assert(TS > DCAP + kRowWave && (TS & (TS - 1)) == 0, "table size");
  __shared__ int32_t key[TS];
  __shared__ int32_t pref[NCAP + 1];
  __shared__ int32_t bst[NCAP];
  const int lane = threadIdx.x;
  const int32_t r = rows[blockIdx.x];
  const int32_t a0 = ap[r], na = ap[r + 1] - a0;
  for (int i = lane; i < TS; i += kRowWave) key[i] = kEmpty;
  const int32_t np = row_products(ac + a0, na, bp, pref, bst);
  int32_t cnt = 0, total = 0;
  bool over = false;
  for (int32_t t0 = 0; t0 < np; t0 += kRowWave) {
    const int32_t t = t0 + lane;
    if (t < np) {
      const int32_t k = product_entry(pref, na, t);
      const int32_t c = bc[bst[k] + (t - pref[k])];
      uint32_t h = hash_col(c, TS - 1);
      for (;;) {
        const int32_t old = atomicCAS(&key[h], kEmpty, c);
        if (old == kEmpty) {
          ++cnt;
          break;
        }
        if (old == c) break;
        h = (h + 1) & (TS - 1);
      }
    }
    total = wave_sum(cnt);
    if (total > DCAP) {
      over = true;
      break;
    }
  }
  if (lane == 0) rowlen[r] = over ? -1 : total;
}

// Values of the listed rows of C = A B (rowlen[r] in 1 .. DCAP distinct
// columns, na <= NCAP): the columns are collected in a table, sorted
// (bitonic, in LDS), each given its sorted position, and the products
// added A entry by A entry.
template <int TS, int DCAP, int NCAP>
__global__ __launch_bounds__(kRowWave) void k_numeric_lds(
    const int32_t *ap, const int32_t *ac, const double *av, const int32_t *bp, const int32_t *bc,
    const double *bv, const int32_t *rows, const int32_t *rowlen, const int32_t *cp, int32_t *cc,
    double *cv) {
  static_assert(TS >= 2 * DCAP && (TS & (TS - 1)) == 0 && (DCAP & (DCAP - 1)) == 0, "sizes");
  __shared__ int32_t key[TS];
  __shared__ uint16_t pos[TS];
  __shared__ int32_t col[DCAP];
  __shared__ double acc[DCAP];
  __shared__ int32_t pref[NCAP + 1];
  __shared__ int32_t bst[NCAP];
  __shared__ int32_t fill;
  const int lane = threadIdx.x;
  const int32_t r = rows[blockIdx.x];
  const int32_t m = rowlen[r];
  const int32_t a0 = ap[r], na = ap[r + 1] - a0;
  for (int i = lane; i < TS; i += kRowWave) key[i] = kEmpty;
  if (lane == 0) fill = 0;
  const int32_t np = row_products(ac + a0, na, bp, pref, bst);
  // The row's distinct columns.
  for (int32_t t0 = 0; t0 < np; t0 += kRowWave) {
    const int32_t t = t0 + lane;
    if (t < np) {
      const int32_t k = product_entry(pref, na, t);
      const int32_t c = bc[bst[k] + (t - pref[k])];
      uint32_t h = hash_col(c, TS - 1);
      for (;;) {
        const int32_t old = atomicCAS(&key[h], kEmpty, c);
        if (old == kEmpty || old == c) break;
        h = (h + 1) & (TS - 1);
      }
    }
  }
  __syncthreads();
  for (int i = lane; i < TS; i += kRowWave) {
    const int32_t c = key[i];
    if (c != kEmpty) col[atomicAdd(&fill, 1)] = c;
  }
  int32_t ns = 1;
  while (ns < m) ns <<= 1;
  __syncthreads();
  for (int32_t i = m + lane; i < ns; i += kRowWave) col[i] = INT_MAX;
  __syncthreads();
  for (int32_t size = 2; size <= ns; size <<= 1) {
    for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (int32_t i = lane; i < ns / 2; i += kRowWave) {
        const int32_t lo = 2 * i - (i & (stride - 1));
        const int32_t hi = lo + stride;
        const int32_t x = col[lo], y = col[hi];
        if ((x > y) == ((lo & size) == 0)) {
          col[lo] = y;
          col[hi] = x;
        }
      }
      __syncthreads();
    }
  }
  for (int32_t p = lane; p < m; p += kRowWave) {
    const int32_t c = col[p];
    uint32_t h = hash_col(c, TS - 1);
    while (key[h] != c) h = (h + 1) & (TS - 1);
    pos[h] = static_cast<uint16_t>(p);
    acc[p] = -0.0;
  }
  __syncthreads();
  for (int32_t t0 = 0; t0 < np; t0 += kRowWave) {
    const int32_t t = t0 + lane;
    const bool valid = t < np;
    int32_t k = -1, p = 0;
    double v = 0.0;
    if (valid) {
      k = product_entry(pref, na, t);
      const int32_t kb = bst[k] + (t - pref[k]);
      const int32_t c = bc[kb];
      v = av[a0 + k] * bv[kb];
      uint32_t h = hash_col(c, TS - 1);
      while (key[h] != c) h = (h + 1) & (TS - 1);
      p = pos[h];
    }
    const int32_t last = min(kRowWave - 1, np - 1 - t0);
    const int32_t klo = __shfl(k, 0, kRowWave), khi = __shfl(k, last, kRowWave);
    for (int32_t s = klo; s <= khi; ++s) {
      if (valid && k == s) acc[p] += v;
      __syncthreads();
    }
  }
  const int64_t base = cp[r];
  for (int32_t p = lane; p < m; p += kRowWave) {
    cc[base + p] = col[p];
    cv[base + p] = acc[p];
  }
}

// Column counts of A (for its transpose).
__global__ void k_col_count(const int32_t *ac, int64_t nnz, int32_t *cnt) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (k < nnz) atomicAdd(&cnt[ac[k]], 1);
}

__global__ void k_transpose_scatter(const int32_t *ap, const int32_t *ac, const double *av,
                                    int64_t rows, int32_t *next, int32_t *tc, double *tv) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= rows) return;
  for (int32_t k = ap[r]; k < ap[r + 1]; ++k) {
    const int32_t d = atomicAdd(&next[ac[k]], 1);
    tc[d] = static_cast<int32_t>(r);
    tv[d] = av[k];
  }
}

__global__ void k_sort_rows(const int32_t *p, int64_t rows, int32_t *c, double *v) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= rows) return;
  heap_sort(c + p[r], v + p[r], p[r + 1] - p[r]);
}

std::vector<int32_t> row_ptr(const std::vector<int32_t> &len, const char *what) {
  std::vector<int32_t> ptr(len.size() + 1, 0);
  int64_t acc = 0;
  for (size_t i = 0; i < len.size(); ++i) {
    acc += len[i];
    if (acc > INT32_MAX) throw std::runtime_error(what);
    ptr[i + 1] = static_cast<int32_t>(acc);
  }
  return ptr;
}

uint32_t pow2_at_least(int64_t v) {
  uint32_t h = 1;
  while (static_cast<int64_t>(h) < v) h <<= 1;
  return h;
}

// Row chunks [r0, r1) whose tables (entry_bytes per table slot) fit the
// scratch buffer; table offsets (in slots) restart at 0 in each chunk.
std::vector<std::pair<int64_t, int64_t>> chunks(const std::vector<int32_t> &tsize,
                                                std::vector<int64_t> &toff, size_t entry_bytes) {
  const int64_t cap = static_cast<int64_t>(kScratchBytes / entry_bytes);
  std::vector<std::pair<int64_t, int64_t>> out;
  toff.assign(tsize.size(), 0);
  int64_t r0 = 0, used = 0;
  for (int64_t r = 0; r < static_cast<int64_t>(tsize.size()); ++r) {
    if (tsize[r] > cap) throw std::runtime_error("device spgemm: row table exceeds scratch.");
    if (used + tsize[r] > cap) {
      out.emplace_back(r0, r);
      r0 = r;
      used = 0;
    }
    toff[r] = used;
    used += tsize[r];
  }
  if (r0 < static_cast<int64_t>(tsize.size())) out.emplace_back(r0, tsize.size());
  return out;
}

// scratch: the global-table kernels' tables, allocated (kScratchBytes) at
// their first use.
// b_unique: no row of B repeats a column (the LDS kernels add one A entry's
// products in parallel; otherwise every row takes the sequential kernels).
DCsr spgemm(const DCsr &A, const DCsr &B, DevBuf &scratch, bool b_unique) {
  if (A.cols != B.rows) throw std::runtime_error("spgemm: shape mismatch.");
  const int64_t n = A.rows;
  DCsr C;
  C.rows = n;
  C.cols = B.cols;
  if (n == 0) {
    C.ptr.upload(std::vector<int32_t>(1, 0));
    return C;
  }
  DevBuf np_d(n * sizeof(int64_t));
  k_products<<<blocks_for(n), kThreads>>>(A.p(), A.c(), B.p(), n, np_d.as<int64_t>());
  std::vector<int64_t> np(n);
  GS_HIP(hipMemcpy(np.data(), np_d.get(), n * sizeof(int64_t), hipMemcpyDeviceToHost));
  std::vector<int32_t> aptr(n + 1);
  GS_HIP(hipMemcpy(aptr.data(), A.p(), (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  auto na = [&](int64_t r) { return static_cast<int64_t>(aptr[r + 1]) - aptr[r]; };
  // Row lists of the LDS kernels (one wavefront per row); a row that fits
  // none of them uses the global-table kernels.
  // (Each list keeps its own buffer until the function returns: a launch
  // may still read it when the next list is uploaded.)
  std::vector<DevBuf> lists;
  auto launch_rows = [&lists](const std::vector<int32_t> &list, auto &&go) {
    if (list.empty()) return;
    lists.emplace_back();
    lists.back().upload(list);
    go(static_cast<uint32_t>(list.size()), lists.back().as<int32_t>());
    GS_HIP(hipGetLastError());
  };
  DevBuf rowlen_d(n * sizeof(int32_t));
  GS_HIP(hipMemset(rowlen_d.get(), 0, n * sizeof(int32_t)));
  // (The LDS kernels count a row's products in int32: below 2^30.)
  constexpr int64_t kNpMax = int64_t(1) << 30;
  std::vector<int32_t> sym_s, sym_l;
  for (int64_t r = 0; r < n && b_unique; ++r) {
    if (np[r] == 0 || np[r] >= kNpMax) continue;
    const int64_t bound = std::min<int64_t>(np[r], B.cols);
    if (bound <= 128 && na(r) <= 256) sym_s.push_back(static_cast<int32_t>(r));
    else if (na(r) <= 1024) sym_l.push_back(static_cast<int32_t>(r));
  }
  launch_rows(sym_s, [&](uint32_t g, const int32_t *rows) {
    k_symbolic_lds<256, 128, 256><<<g, kRowWave>>>(A.p(), A.c(), B.p(), B.c(), rows,
                                                   rowlen_d.as<int32_t>());
  });
  launch_rows(sym_l, [&](uint32_t g, const int32_t *rows) {
    k_symbolic_lds<2048, 1024, 1024><<<g, kRowWave>>>(A.p(), A.c(), B.p(), B.c(), rows,
                                                      rowlen_d.as<int32_t>());
  });
  std::vector<int32_t> rowlen(n);
  GS_HIP(hipMemcpy(rowlen.data(), rowlen_d.get(), n * sizeof(int32_t), hipMemcpyDeviceToHost));
  // Symbolic pass of the remaining rows (more than 1024 columns or A
  // entries): tables in global memory sized for the distinct-column bound.
  std::vector<int32_t> tsize(n, 0);
  bool any_global = false;
  {
    std::vector<char> lds(n, 0);
    for (int32_t r : sym_s) lds[r] = 1;
    for (int32_t r : sym_l) lds[r] = rowlen[r] >= 0;
    for (int64_t r = 0; r < n; ++r) {
      if (np[r] == 0 || lds[r]) continue;
      const int64_t bound = std::min<int64_t>(np[r], B.cols);
      tsize[r] = static_cast<int32_t>(pow2_at_least(2 * bound));
      any_global = true;
    }
  }
  std::vector<int64_t> toff;
  DevBuf toff_d, tsize_d;
  if (any_global) {
    if (!scratch.get()) scratch.alloc(kScratchBytes);
    tsize_d.upload(tsize);
    for (const auto &ch : chunks(tsize, toff, sizeof(int32_t))) {
      toff_d.upload(toff);
      const int64_t m = ch.second - ch.first;
      k_symbolic<<<blocks_for(m), kThreads>>>(A.p(), A.c(), B.p(), B.c(), ch.first, ch.second,
                                             toff_d.as<int64_t>(), tsize_d.as<int32_t>(),
                                             scratch.as<int32_t>(), rowlen_d.as<int32_t>());
      GS_HIP(hipGetLastError());
    }
    GS_HIP(hipMemcpy(rowlen.data(), rowlen_d.get(), n * sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  const std::vector<int32_t> cptr = row_ptr(rowlen, "spgemm: nnz exceeds int32.");
  C.nnz = cptr[n];
  C.ptr.upload(cptr);
  C.col.alloc(C.nnz * sizeof(int32_t));
  C.val.alloc(C.nnz * sizeof(double));
  // Numeric pass: LDS kernels by distinct-column count and A row length;
  // the rest with global tables sized for the exact distinct count.
  std::vector<int32_t> num_s, num_m, num_l;
  any_global = false;
  for (int64_t r = 0; r < n; ++r) {
    tsize[r] = 0;
    const int64_t m = rowlen[r];
    if (m == 0) continue;
    if (!b_unique || np[r] >= kNpMax) {
      tsize[r] = static_cast<int32_t>(pow2_at_least(2 * m));
      any_global = true;
    } else if (m <= 128 && na(r) <= 256) num_s.push_back(static_cast<int32_t>(r));
    else if (m <= 256 && na(r) <= 256) num_m.push_back(static_cast<int32_t>(r));
    else if (m <= 1024 && na(r) <= 1024) num_l.push_back(static_cast<int32_t>(r));
    else {
      tsize[r] = static_cast<int32_t>(pow2_at_least(2 * m));
      any_global = true;
    }
  }
  auto numeric_lds = [&](const std::vector<int32_t> &list, auto kernel) {
    launch_rows(list, [&](uint32_t g, const int32_t *rows) {
      kernel<<<g, kRowWave>>>(A.p(), A.c(), A.v(), B.p(), B.c(), B.v(), rows,
                              rowlen_d.as<int32_t>(), C.ptr.as<int32_t>(), C.col.as<int32_t>(),
                              C.val.as<double>());
    });
  };
  numeric_lds(num_s, k_numeric_lds<256, 128, 256>);
  numeric_lds(num_m, k_numeric_lds<512, 256, 256>);
  numeric_lds(num_l, k_numeric_lds<2048, 1024, 1024>);
  if (any_global) {
    if (!scratch.get()) scratch.alloc(kScratchBytes);
    tsize_d.upload(tsize);
    for (const auto &ch : chunks(tsize, toff, 2 * sizeof(int32_t))) {
      toff_d.upload(toff);
      const int64_t m = ch.second - ch.first;
      k_numeric<<<blocks_for(m), kThreads>>>(A.p(), A.c(), A.v(), B.p(), B.c(), B.v(), ch.first,
                                            ch.second, toff_d.as<int64_t>(),
                                            tsize_d.as<int32_t>(), scratch.as<int32_t>(),
                                            C.ptr.as<int32_t>(), C.col.as<int32_t>(),
                                            C.val.as<double>());
      GS_HIP(hipGetLastError());
    }
  }
  GS_HIP(hipDeviceSynchronize());
  return C;
}

DCsr transpose(const DCsr &A) {
  DCsr T;
  T.rows = A.cols;
  T.cols = A.rows;
  T.nnz = A.nnz;
  DevBuf cnt(std::max<int64_t>(A.cols, 1) * sizeof(int32_t));
  GS_HIP(hipMemset(cnt.get(), 0, std::max<int64_t>(A.cols, 1) * sizeof(int32_t)));
  if (A.nnz) k_col_count<<<blocks_for(A.nnz), kThreads>>>(A.c(), A.nnz, cnt.as<int32_t>());
  std::vector<int32_t> len(A.cols);
  if (A.cols)
    GS_HIP(hipMemcpy(len.data(), cnt.get(), A.cols * sizeof(int32_t), hipMemcpyDeviceToHost));
  const std::vector<int32_t> tptr = row_ptr(len, "transpose: nnz exceeds int32.");
  T.ptr.upload(tptr);
  T.col.alloc(T.nnz * sizeof(int32_t));
  T.val.alloc(T.nnz * sizeof(double));
  if (A.nnz) {
    DevBuf next;
    next.upload(tptr.data(), static_cast<size_t>(A.cols));
    k_transpose_scatter<<<blocks_for(A.rows), kThreads>>>(A.p(), A.c(), A.v(), A.rows,
                                                         next.as<int32_t>(), T.col.as<int32_t>(),
                                                         T.val.as<double>());
    k_sort_rows<<<blocks_for(T.rows), kThreads>>>(T.ptr.as<int32_t>(), T.rows,
                                                 T.col.as<int32_t>(), T.val.as<double>());
    GS_HIP(hipGetLastError());
  }
  GS_HIP(hipDeviceSynchronize());
  return T;
}

}  // namespace

void DeviceGalerkinOps::run(const gsx::Csr &A, const gsx::Csr &M, const gsx::Csr &Pt,
                            gsx::Csr &P, gsx::Csr &R, gsx::Csr &Ac) const {
  DevBuf scratch;
  const DCsr dM = upload(M), dPt = upload(Pt);
  // The products' own rows are sorted and unique.
  const DCsr dP = spgemm(dM, dPt, scratch, gsx::rows_sorted_unique(Pt));
  const DCsr dR = transpose(dP);
  const DCsr dA = upload(A);
  const DCsr dAP = spgemm(dA, dP, scratch, true);
  const DCsr dAc = spgemm(dR, dAP, scratch, true);
  P = download(dP);
  R = download(dR);
  Ac = download(dAc);
}

gsx::Csr device_spgemm(const gsx::Csr &A, const gsx::Csr &B) {
  DevBuf scratch;
  return download(spgemm(upload(A), upload(B), scratch, gsx::rows_sorted_unique(B)));
}

gsx::Csr device_transpose(const gsx::Csr &A) { return download(transpose(upload(A))); }


// ---------------------------------------------------------------------------
// Operator packing on the device: the relabelling and the 3x3-block layouts
// of a level operator, equal to gsx::relabel, gsx::bsr3_pack and
// gsx::bsr3_sell64_pack (+ gsx::bsr3_pair_planes) array for array.
// ---------------------------------------------------------------------------
namespace {

__global__ void k_relabel_len(const int32_t *ap, const int32_t *rq, int64_t n, int32_t *len) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) len[i] = ap[rq[i] + 1] - ap[rq[i]];
}

__global__ void k_relabel_copy(const int32_t *ap, const int32_t *ac, const double *av,
                               const int32_t *rq, const int32_t *cmap, int64_t n,
                               const int32_t *bp, int32_t *bc, double *bv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  int64_t d = bp[i];
  for (int32_t k = ap[rq[i]]; k < ap[rq[i] + 1]; ++k, ++d) {
    bc[d] = cmap[ac[k]];
    bv[d] = av[k];
  }
}

__device__ void heap_sort64(int64_t *keys, int64_t n) {
  auto sift = [&](int64_t root, int64_t end) {
    for (;;) {
      int64_t child = 2 * root + 1;
      if (child >= end) return;
      if (child + 1 < end && keys[child + 1] > keys[child]) ++child;
      if (keys[root] >= keys[child]) return;
      const int64_t t = keys[root];
      keys[root] = keys[child];
      keys[child] = t;
      root = child;
    }
  };
  for (int64_t s = n / 2 - 1; s >= 0; --s) sift(s, n);
  for (int64_t e = n - 1; e > 0; --e) {
    const int64_t t = keys[0];
    keys[0] = keys[e];
    keys[e] = t;
    sift(0, e);
  }
}

// Block columns of block row b (the union over its three scalar rows),
// ordered by (rank, column) and made unique, left at the start of the
// rows' own CSR range of `scratch` (keys rank << 32 | column); cnt[b] their
// number.
__global__ void k_b3_columns(const int32_t *ap, const int32_t *ac, int64_t nb,
                             const int32_t *key, int64_t *scratch, int64_t *cnt) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (b >= nb) return;
  const int64_t k0 = ap[3 * b], k1 = ap[3 * b + 3];
  int64_t *s = scratch + k0;
  for (int64_t k = k0; k < k1; ++k) {
    const int64_t cb = ac[k] / 3;
    const int64_t rank = key ? key[cb] : cb;
    s[k - k0] = (rank << 32) | cb;
  }
  const int64_t m = k1 - k0;
  heap_sort64(s, m);
  int64_t u = 0;
  for (int64_t i = 0; i < m; ++i)
    if (i == 0 || s[i] != s[u - 1]) s[u++] = s[i];
  cnt[b] = u;
}

// Block columns and values of every block row: values start at 0 and the
// scalar entries are added in CSR order (rows 3b, 3b + 1, 3b + 2).
__global__ void k_b3_fill(const int32_t *ap, const int32_t *ac, const double *av, int64_t nb,
                          const int64_t *scratch, const int64_t *bptr, int32_t *bcol,
                          double *bval) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (b >= nb) return;
  const int64_t *s = scratch + ap[3 * b];
  const int64_t p0 = bptr[b], m = bptr[b + 1] - p0;
  for (int64_t i = 0; i < m; ++i) bcol[p0 + i] = static_cast<int32_t>(s[i] & 0xffffffff);
  double *v = bval + 9 * p0;
  for (int64_t i = 0; i < 9 * m; ++i) v[i] = 0.0;
  for (int l = 0; l < 3; ++l)
    for (int32_t k = ap[3 * b + l]; k < ap[3 * b + l + 1]; ++k) {
      const int32_t cb = ac[k] / 3;
      int64_t slot = 0;
      while (bcol[p0 + slot] != cb) ++slot;
      v[9 * slot + 3 * l + ac[k] % 3] += av[k];
    }
}

__global__ void k_b3_below_count(const int64_t *bptr, const int32_t *bcol, int64_t nb,
                                 const int32_t *start, int64_t *cnt) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (b >= nb) return;
  int64_t c = 0;
  for (int64_t k = bptr[b]; k < bptr[b + 1]; ++k) c += bcol[k] < start[b] ? 1 : 0;
  cnt[b] = c;
}

__global__ void k_b3_below_copy(const int64_t *bptr, const int32_t *bcol, const double *bval,
                                int64_t nb, const int32_t *start, const int64_t *optr,
                                int32_t *ocol, double *oval) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (b >= nb) return;
  int64_t d = optr[b];
  for (int64_t k = bptr[b]; k < bptr[b + 1]; ++k)
    if (bcol[k] < start[b]) {
      ocol[d] = bcol[k];
      for (int e = 0; e < 9; ++e) oval[9 * d + e] = bval[9 * k + e];
      ++d;
    }
}

// Width of each 64-position slice (position p holds block row order[p], or
// p without an order).
__global__ void k_b3_slice_width(const int64_t *bptr, const int32_t *order, int64_t nb,
                                 int64_t ns, int32_t *w) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (s >= ns) return;
  int64_t m = 0;
  for (int64_t p = 64 * s; p < 64 * s + 64 && p < nb; ++p) {
    const int64_t r = order ? order[p] : p;
    m = max(m, bptr[r + 1] - bptr[r]);
  }
  w[s] = static_cast<int32_t>(m);
}

// One (slice, lane) per thread: the lane's block row across the slice's
// slots, with the host packer's padding (a zero block on the position's own
// column for square operators, else the row's last block column or 0).
template <class T>
__global__ void k_b3_sell_fill(const int64_t *bptr, const int32_t *bcol, const double *bval,
                               const int32_t *order, int64_t nb, bool square, int64_t ns,
                               const int64_t *slice_ptr, const int32_t *slice_w, bool pairs,
                               int32_t *obcol, T *oval) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t s = t >> 6, l = t & 63;
  if (s >= ns) return;
  const int64_t p = 64 * s + l;
  const bool valid = p < nb;
  const int64_t r = valid ? (order ? order[p] : p) : 0;
  const int64_t k0 = valid ? bptr[r] : 0, len = valid ? bptr[r + 1] - k0 : 0;
  int32_t pc = 0;
  if (square) pc = valid ? static_cast<int32_t>(p) : 0;
  else if (len > 0) pc = bcol[k0 + len - 1];
  for (int32_t j = 0; j < slice_w[s]; ++j) {
    const int64_t slot = slice_ptr[s] + static_cast<int64_t>(j) * 64;
    const bool have = j < len;
    obcol[slot + l] = have ? bcol[k0 + j] : pc;
    T *v = oval + slot * 9;
    for (int e = 0; e < 9; ++e) {
      const int64_t off = pairs ? (e < 8 ? 128 * (e >> 1) + 2 * l + (e & 1) : 512 + l) : 64 * e + l;
      v[off] = have ? static_cast<T>(bval[9 * (k0 + j) + e]) : T(0);
    }
  }
}

template <class T>
__global__ void k_cast(const double *src, int64_t n, T *dst) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) dst[i] = static_cast<T>(src[i]);
}

template <class T>
std::vector<T> fetch(const DevBuf &b, int64_t n) {
  std::vector<T> h(n);
  if (n) GS_HIP(hipMemcpy(h.data(), b.get(), n * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}

// Exclusive prefix of per-row counts (downloaded), as int64 offsets.
std::vector<int64_t> offsets(const DevBuf &cnt, int64_t n) {
  const std::vector<int64_t> c = fetch<int64_t>(cnt, n);
  std::vector<int64_t> ptr(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + c[i];
  return ptr;
}

}  // namespace

DCsr relabel(const DCsr &A, const std::vector<int32_t> &rq, const std::vector<int32_t> &cmap) {
  const int64_t n = A.rows;
  DCsr B;
  B.rows = n;
  B.cols = A.cols;
  B.nnz = A.nnz;
  DevBuf rq_d, cmap_d, len_d(std::max<int64_t>(n, 1) * sizeof(int32_t));
  rq_d.upload(rq);
  cmap_d.upload(cmap);
  k_relabel_len<<<blocks_for(n), kThreads>>>(A.p(), rq_d.as<int32_t>(), n, len_d.as<int32_t>());
  const std::vector<int32_t> len = fetch<int32_t>(len_d, n);
  B.ptr.upload(row_ptr(len, "relabel: nnz exceeds int32."));
  B.col.alloc(A.nnz * sizeof(int32_t));
  B.val.alloc(A.nnz * sizeof(double));
  k_relabel_copy<<<blocks_for(n), kThreads>>>(A.p(), A.c(), A.v(), rq_d.as<int32_t>(),
                                             cmap_d.as<int32_t>(), n, B.ptr.as<int32_t>(),
                                             B.col.as<int32_t>(), B.val.as<double>());
  GS_HIP(hipGetLastError());
  GS_HIP(hipDeviceSynchronize());
  return B;
}

bool bsr3_pack(const DCsr &A, const int32_t *key, DBsr3Csr &out) {
  if (A.rows % 3 != 0 || A.cols % 3 != 0) return false;
  const int64_t nb = A.rows / 3;
  DevBuf key_d, scratch(std::max<int64_t>(A.nnz, 1) * sizeof(int64_t)),
      cnt(std::max<int64_t>(nb, 1) * sizeof(int64_t));
  if (key) key_d.upload(key, static_cast<size_t>(A.cols / 3));
  k_b3_columns<<<blocks_for(nb), kThreads>>>(A.p(), A.c(), nb, key ? key_d.as<int32_t>() : nullptr,
                                            scratch.as<int64_t>(), cnt.as<int64_t>());
  GS_HIP(hipGetLastError());
  out.nb = nb;
  out.ncb = A.cols / 3;
  out.host_ptr = offsets(cnt, nb);
  out.nblocks = out.host_ptr[nb];
  int64_t width = 0;
  for (int64_t b = 0; b < nb; ++b) width = std::max(width, out.host_ptr[b + 1] - out.host_ptr[b]);
  out.width = static_cast<int32_t>(width);
  out.ptr.upload(out.host_ptr);
  out.bcol.alloc(out.nblocks * sizeof(int32_t));
  out.val.alloc(9 * out.nblocks * sizeof(double));
  k_b3_fill<<<blocks_for(nb), kThreads>>>(A.p(), A.c(), A.v(), nb, scratch.as<int64_t>(),
                                         out.ptr.as<int64_t>(), out.bcol.as<int32_t>(),
                                         out.val.as<double>());
  GS_HIP(hipGetLastError());
  GS_HIP(hipDeviceSynchronize());
  return true;
}

DBsr3Csr bsr3_below(const DBsr3Csr &B, const std::vector<int32_t> &start) {
  DBsr3Csr out;
  out.nb = B.nb;
  out.ncb = B.ncb;
  DevBuf start_d, cnt(std::max<int64_t>(B.nb, 1) * sizeof(int64_t));
  start_d.upload(start);
  k_b3_below_count<<<blocks_for(B.nb), kThreads>>>(B.ptr.as<int64_t>(), B.bcol.as<int32_t>(), B.nb,
                                                  start_d.as<int32_t>(), cnt.as<int64_t>());
  out.host_ptr = offsets(cnt, B.nb);
  out.nblocks = out.host_ptr[B.nb];
  int64_t width = 0;
  for (int64_t b = 0; b < B.nb; ++b) width = std::max(width, out.host_ptr[b + 1] - out.host_ptr[b]);
  out.width = static_cast<int32_t>(width);
  out.ptr.upload(out.host_ptr);
  out.bcol.alloc(out.nblocks * sizeof(int32_t));
  out.val.alloc(9 * out.nblocks * sizeof(double));
  k_b3_below_copy<<<blocks_for(B.nb), kThreads>>>(B.ptr.as<int64_t>(), B.bcol.as<int32_t>(),
                                                 B.val.as<double>(), B.nb, start_d.as<int32_t>(),
                                                 out.ptr.as<int64_t>(), out.bcol.as<int32_t>(),
                                                 out.val.as<double>());
  GS_HIP(hipGetLastError());
  GS_HIP(hipDeviceSynchronize());
  return out;
}

template <class T>
void bsr3_sell_pack(const DBsr3Csr &B, const int32_t *order, gsd::DBsr3<T> &out) {
  const int64_t nb = B.nb, ns = (nb + 63) / 64;
  DevBuf order_d, w_d(std::max<int64_t>(ns, 1) * sizeof(int32_t));
  if (order) order_d.upload(order, static_cast<size_t>(nb));
  const int32_t *ord = order ? order_d.as<int32_t>() : nullptr;
  k_b3_slice_width<<<blocks_for(ns), kThreads>>>(B.ptr.as<int64_t>(), ord, nb, ns,
                                                w_d.as<int32_t>());
  const std::vector<int32_t> w = fetch<int32_t>(w_d, ns);
  std::vector<int64_t> sp(ns + 1, 0);
  for (int64_t s = 0; s < ns; ++s) sp[s + 1] = sp[s] + int64_t(w[s]) * 64;
  out.nb = nb;
  out.nblocks = B.nblocks;
  out.slots = sp[ns];
  out.slice_ptr.upload(sp);
  out.slice_w.upload(w);
  out.bcol.alloc(out.slots * sizeof(int32_t));
  out.val.alloc(9 * out.slots * sizeof(T));
  k_b3_sell_fill<T><<<blocks_for(64 * ns), kThreads>>>(
      B.ptr.as<int64_t>(), B.bcol.as<int32_t>(), B.val.as<double>(), ord, nb, B.nb == B.ncb, ns,
      out.slice_ptr.template as<int64_t>(), out.slice_w.template as<int32_t>(), out.pairs,
      out.bcol.template as<int32_t>(), out.val.template as<T>());
  GS_HIP(hipGetLastError());
  if (order) {
    out.row_map = std::move(order_d);
    out.mapped = true;
  } else {
    out.row_map.release();
    out.mapped = false;
  }
  GS_HIP(hipDeviceSynchronize());
}

template <class T>
void bsr3_rows(const DBsr3Csr &B, gsd::DBRows<T> &out) {
  out.nb = B.nb;
  out.nblocks = B.nblocks;
  out.ptr.upload(B.host_ptr);
  out.bcol.alloc(std::max<int64_t>(B.nblocks, 1) * sizeof(int32_t));
  if (B.nblocks)
    GS_HIP(hipMemcpy(out.bcol.get(), B.bcol.get(), B.nblocks * sizeof(int32_t),
                     hipMemcpyDeviceToDevice));
  out.val.alloc(9 * B.nblocks * sizeof(T));
  if (B.nblocks)
    k_cast<T><<<blocks_for(9 * B.nblocks), kThreads>>>(B.val.as<double>(), 9 * B.nblocks,
                                                      out.val.template as<T>());
  GS_HIP(hipGetLastError());
  out.row_map.release();
  out.mapped = false;
  GS_HIP(hipDeviceSynchronize());
}

template void bsr3_sell_pack<float>(const DBsr3Csr &, const int32_t *, gsd::DBsr3<float> &);
template void bsr3_sell_pack<double>(const DBsr3Csr &, const int32_t *, gsd::DBsr3<double> &);
template void bsr3_rows<float>(const DBsr3Csr &, gsd::DBRows<float> &);
template void bsr3_rows<double>(const DBsr3Csr &, gsd::DBRows<double> &);

}  // namespace gsdev
