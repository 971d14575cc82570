// GPU side of the smoothed-aggregation setup: the sparse products of one
// coarsening step (P = M P_tent, R = P^T, A_c = R (A P)), which dominate the
// parallel part of building a hierarchy. The reference obtains its
// hierarchy from amgcl (examples/amgSolver/src/SmoothedAggregationCoarsener.cc:
// 93-97: transfer_operators, then coarse_operator); the host restatement is
// gsx::GalerkinOps. The results here equal the host's bit for bit:
//
//  * SpGEMM C = A B, one thread per output row, replays the host's
//    Gustavson loop exactly: A's row entries in order, each one's B row in
//    order; a column's first product starts its sum and later products are
//    added in that order (a per-row open-addressing table maps columns to
//    their first-touch slot); the row is then sorted by column. A symbolic
//    pass counts each row's distinct columns first. The tables live in one
//    scratch buffer, rows processed in chunks that fit it.
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

struct DCsr {
  int64_t rows = 0, cols = 0, nnz = 0;
  DevBuf ptr, col, val;
  const int32_t *p() const { return ptr.as<int32_t>(); }
  const int32_t *c() const { return col.as<int32_t>(); }
  const double *v() const { return val.as<double>(); }
};

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

gsx::Csr download(const DCsr &d) {
  gsx::Csr A;
  A.rows = d.rows;
  A.cols = d.cols;
  A.ptr.resize(d.rows + 1);
  A.col.resize(d.nnz);
  A.val.resize(d.nnz);
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
  int32_t cnt = 0;
  if (H) {
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

DCsr spgemm(const DCsr &A, const DCsr &B, DevBuf &scratch) {
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
  // Symbolic pass: tables sized for the distinct-column bound.
  std::vector<int32_t> tsize(n);
  for (int64_t r = 0; r < n; ++r) {
    const int64_t bound = std::min<int64_t>(np[r], B.cols);
    tsize[r] = bound ? static_cast<int32_t>(pow2_at_least(2 * bound)) : 0;
  }
  std::vector<int64_t> toff;
  DevBuf toff_d, tsize_d, rowlen_d(n * sizeof(int32_t));
  tsize_d.upload(tsize);
  for (const auto &ch : chunks(tsize, toff, sizeof(int32_t))) {
    toff_d.upload(toff);
    const int64_t m = ch.second - ch.first;
    k_symbolic<<<blocks_for(m), kThreads>>>(A.p(), A.c(), B.p(), B.c(), ch.first, ch.second,
                                           toff_d.as<int64_t>(), tsize_d.as<int32_t>(),
                                           scratch.as<int32_t>(), rowlen_d.as<int32_t>());
    GS_HIP(hipGetLastError());
  }
  std::vector<int32_t> rowlen(n);
  GS_HIP(hipMemcpy(rowlen.data(), rowlen_d.get(), n * sizeof(int32_t), hipMemcpyDeviceToHost));
  const std::vector<int32_t> cptr = row_ptr(rowlen, "spgemm: nnz exceeds int32.");
  C.nnz = cptr[n];
  C.ptr.upload(cptr);
  C.col.alloc(C.nnz * sizeof(int32_t));
  C.val.alloc(C.nnz * sizeof(double));
  // Numeric pass: tables sized for the exact distinct count.
  for (int64_t r = 0; r < n; ++r)
    tsize[r] = rowlen[r] ? static_cast<int32_t>(pow2_at_least(2 * int64_t(rowlen[r]))) : 0;
  tsize_d.upload(tsize);
  for (const auto &ch : chunks(tsize, toff, 2 * sizeof(int32_t))) {
    toff_d.upload(toff);
    const int64_t m = ch.second - ch.first;
    k_numeric<<<blocks_for(m), kThreads>>>(A.p(), A.c(), A.v(), B.p(), B.c(), B.v(), ch.first,
                                          ch.second, toff_d.as<int64_t>(), tsize_d.as<int32_t>(),
                                          scratch.as<int32_t>(), C.ptr.as<int32_t>(),
                                          C.col.as<int32_t>(), C.val.as<double>());
    GS_HIP(hipGetLastError());
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
  DevBuf scratch(kScratchBytes);
  const DCsr dM = upload(M), dPt = upload(Pt);
  const DCsr dP = spgemm(dM, dPt, scratch);
  const DCsr dR = transpose(dP);
  const DCsr dA = upload(A);
  const DCsr dAP = spgemm(dA, dP, scratch);
  const DCsr dAc = spgemm(dR, dAP, scratch);
  P = download(dP);
  R = download(dR);
  Ac = download(dAc);
}

gsx::Csr device_spgemm(const gsx::Csr &A, const gsx::Csr &B) {
  DevBuf scratch(kScratchBytes);
  return download(spgemm(upload(A), upload(B), scratch));
}

gsx::Csr device_transpose(const gsx::Csr &A) { return download(transpose(upload(A))); }

}  // namespace gsdev
