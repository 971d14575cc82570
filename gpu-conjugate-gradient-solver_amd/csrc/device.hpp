// Device-memory helpers shared by the solver engine and the Backend plugin:
// HIP error checking that raises std::runtime_error (the reference's only
// error channel, include/GPUSolver/cuda/CudaHelper.h:11-21), an owning
// device buffer, and the SELL-64 / BSR3-SELL-64 device matrices with their
// host-side packers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "host/setup.h"
#include "kernels.hpp"
#include "kernels_wide.hpp"

namespace gsd {

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

#define GS_HIP(call) ::gsd::hip_check((call), #call)

class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 256;  // keep a valid pointer for empty arrays
    GS_HIP(hipMalloc(&p_, bytes));
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p_); }
  void *get() const { return p_; }
  size_t bytes() const { return n_; }
  template <class T, class A>
  void upload(const std::vector<T, A> &v) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) GS_HIP(hipMemcpy(p_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  template <class T>
  void upload(const T *v, size_t n) {
    alloc(n * sizeof(T));
    if (n) GS_HIP(hipMemcpy(p_, v, n * sizeof(T), hipMemcpyHostToDevice));
  }

 private:
  void *p_ = nullptr;
  size_t n_ = 0;
};

template <class T>
std::vector<T> cast_values(const std::vector<double> &v) {
  std::vector<T> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = static_cast<T>(v[i]);
  return out;
}

// Uploads n FP64 values as T (the FP32 cast runs in parallel on the host).
template <class T>
void upload_values(DevBuf &buf, const double *v, size_t n) {
  if constexpr (std::is_same<T, double>::value) {
    buf.upload(v, n);
  } else {
    gsx::RawVec<float> f(n);
    gsx::to_float(v, n, f.data());
    buf.upload(f.data(), n);
  }
}

// SELL-64 device matrix (layout and validation in gsx::sell64_pack).
template <class T>
struct DSell {
  int64_t rows = 0, cols = 0;
  int64_t nnz = 0;    // stored CSR entries
  int64_t slots = 0;  // including padding
  DevBuf slice_ptr, slice_w, col, val, row_map;
  bool mapped = false;  // rows sorted by length within windows (row_map)

  // sigma > 0 (non-square operators only): SELL-C-sigma row order.
  // order (non-square operators only): stored position p holds row
  // order[p] before any sorting.
  void build(const gsx::Csr &A, int64_t sigma = 0, const int32_t *order = nullptr) {
    gsx::Sell64 P;
    if (order) {
      const gsx::Csr B = gsx::take_rows(A, order, A.rows);
      P = sigma > 0 ? gsx::sell64_pack_sorted(B, sigma) : gsx::sell64_pack(B);
      if (P.row_map.empty()) P.row_map.assign(order, order + A.rows);
      else
        for (auto &r : P.row_map) r = order[r];
    } else {
      P = sigma > 0 ? gsx::sell64_pack_sorted(A, sigma) : gsx::sell64_pack(A);
    }
    rows = P.rows;
    cols = P.cols;
    nnz = P.nnz;
    slots = P.slots;
    slice_ptr.upload(P.slice_ptr);
    slice_w.upload(P.slice_w);
    col.upload(P.col);
    upload_values<T>(val, P.val.data(), P.val.size());
    mapped = !P.row_map.empty();
    if (mapped) row_map.upload(P.row_map);
    else row_map.release();
  }
  gsk::Sell<T> view() const {
    return gsk::Sell<T>{rows, cols, slice_ptr.as<int64_t>(), slice_w.as<int32_t>(),
                        col.as<int32_t>(), val.as<T>(),
                        mapped ? row_map.as<int32_t>() : nullptr};
  }
  size_t device_bytes() const {
    return slice_ptr.bytes() + slice_w.bytes() + col.bytes() + val.bytes() + row_map.bytes();
  }
};

// BSR3-SELL-64 device matrix (layout in gsx::bsr3_sell64_pack).
template <class T>
struct DBsr3 {
  int64_t nb = 0;
  int64_t nblocks = 0;  // stored blocks (without padding)
  int64_t slots = 0;
  DevBuf slice_ptr, slice_w, bcol, val, row_map;
  bool mapped = false;
  // Values as four planes of pairs and one plane per 64-slot group
  // (gsx::bsr3_pair_planes); set before build, read by the kernels that
  // take the PAIRS layout.
  bool pairs = false;

  void build(const gsx::Bsr3Csr &B) {
    gsx::Bsr3Sell64 P = gsx::bsr3_sell64_pack(B);
    if (pairs) gsx::bsr3_pair_planes(P);
    nb = P.nb;
    nblocks = P.nblocks;
    slots = P.slots;
    slice_ptr.upload(P.slice_ptr);
    slice_w.upload(P.slice_w);
    bcol.upload(P.bcol);
    upload_values<T>(val, P.val.data(), P.val.size());
    mapped = false;
    row_map.release();
  }
  // Transfer operators: block rows executed in `order` (null: natural
  // order), then sorted by block count within windows of `sigma` (> 0).
  void build(const gsx::Bsr3Csr &B, int64_t sigma, const int32_t *order) {
    std::vector<int32_t> map;
    if (order) map.assign(order, order + B.nb);
    const gsx::Bsr3Csr B1 = order ? gsx::bsr3_take_rows(B, order, B.nb) : gsx::Bsr3Csr{};
    const gsx::Bsr3Csr &Bo = order ? B1 : B;
    if (sigma > 0) {
      const std::vector<int32_t> srt = gsx::bsr3_sorted_order(Bo, sigma);
      std::vector<int32_t> m2(B.nb);
      for (int64_t p = 0; p < B.nb; ++p) m2[p] = order ? map[srt[p]] : srt[p];
      build(gsx::bsr3_take_rows(Bo, srt.data(), B.nb));
      map = std::move(m2);
    } else {
      build(Bo);
    }
    if (!map.empty()) {
      row_map.upload(map);
      mapped = true;
    }
  }
  gsk::Bsr3<T> view() const {
    return gsk::Bsr3<T>{nb, slice_ptr.as<int64_t>(), slice_w.as<int32_t>(), bcol.as<int32_t>(),
                        val.as<T>(), mapped ? row_map.as<int32_t>() : nullptr};
  }
  int64_t slices() const { return (nb + 63) / 64; }
  size_t device_bytes() const {
    return slice_ptr.bytes() + slice_w.bytes() + bcol.bytes() + val.bytes() + row_map.bytes();
  }
};

// Row-contiguous CSR for the group-per-row kernels.
template <class T>
struct DRows {
  int64_t rows = 0, cols = 0, nnz = 0;
  DevBuf ptr, col, val, row_map;
  bool mapped = false;
  // order: stored position p holds row order[p] (null: natural order).
  void build(const gsx::Csr &A, const int32_t *order = nullptr) {
    if (order) {
      build(gsx::take_rows(A, order, A.rows));
      row_map.upload(order, static_cast<size_t>(A.rows));
      mapped = true;
      return;
    }
    rows = A.rows;
    cols = A.cols;
    nnz = A.nnz();
    std::vector<int64_t> p(A.ptr.begin(), A.ptr.end());
    gsx::check_columns(A, "CSR upload: column index out of range.");
    ptr.upload(p);
    col.upload(A.col);
    upload_values<T>(val, A.val.data(), A.val.size());
    mapped = false;
    row_map.release();
  }
  gsk::Rows<T> view() const {
    return gsk::Rows<T>{rows, cols, ptr.as<int64_t>(), col.as<int32_t>(), val.as<T>(),
                        mapped ? row_map.as<int32_t>() : nullptr};
  }
};

// Row-contiguous 3x3 block CSR (blocks of a row in the packer's order).
template <class T>
struct DBRows {
  int64_t nb = 0, nblocks = 0;
  DevBuf ptr, bcol, val, row_map;
  bool mapped = false;
  // order: stored position p holds block row order[p] (null: natural).
  void build(const gsx::Bsr3Csr &B, const int32_t *order = nullptr) {
    if (order) {
      build(gsx::bsr3_take_rows(B, order, B.nb));
      row_map.upload(order, static_cast<size_t>(B.nb));
      mapped = true;
      return;
    }
    nb = B.nb;
    nblocks = B.nblocks;
    ptr.upload(B.ptr);
    bcol.upload(B.bcol);
    upload_values<T>(val, B.val.data(), B.val.size());
    mapped = false;
    row_map.release();
  }
  gsk::BRows<T> view() const {
    return gsk::BRows<T>{nb, ptr.as<int64_t>(), bcol.as<int32_t>(), val.as<T>(),
                         mapped ? row_map.as<int32_t>() : nullptr};
  }
};

// Lanes per row for the group-per-row kernels given the average row length
// (entries, or blocks for block rows); 0 keeps the thread-per-row SELL
// layout. min_avg < 0 never groups.
inline int pick_group(double avg, double min_avg, int max_g = 64) {
  if (min_avg < 0.0 || avg < min_avg) return 0;
  int g = 64;
  if (avg <= 16.0) g = 16;
  else if (avg <= 32.0) g = 32;
  return g < max_g ? g : max_g;
}
// Block rows: lanes per block row from the average blocks per row (the
// smoothers load two blocks per lane and round; the transfers pick 2-4).
inline int pick_block_group(double avg, double min_avg) {
  if (min_avg < 0.0 || avg < min_avg) return 0;
  if (avg <= 16.0) return 8;
  if (avg <= 32.0) return 16;
  if (avg <= 64.0) return 32;
  return 64;
}

// Transfer operator: scalar (SELL-64 or row-contiguous CSR) or 3x3 blocks
// (BSR3-SELL-64 or row-contiguous block CSR).
template <class T>
struct DMat {
  int G = 0;  // lanes per row; 0: SELL-64, one thread per row
  int K = 2;  // lane groups: entries (scalar: 2, 4, 8; 3x3 blocks: 2, 3, 4)
             // per lane and load round
  bool b3 = false;
  int64_t rows = 0, cols = 0, nnz = 0;
  int64_t nblocks = 0;  // 3x3 blocks (block storage)
  DSell<T> s;
  DRows<T> w;
  DBsr3<T> sb;
  DBRows<T> bw;
  // 3x3-block storage of a transfer whose row and column unknowns both come
  // in node triples: key orders each block row's blocks (the CSR rows'
  // column order), order[p] is the block row at position p (null: natural),
  // group_min: average blocks per block row from which lane groups are used.
  void build_blocks(const gsx::Csr &A, const int32_t *key, double group_min, int64_t sigma,
                    const int32_t *order) {
    gsx::Bsr3Csr B;
    if (!gsx::bsr3_pack(A, B, key))
      throw std::runtime_error("block transfer: dimensions are not multiples of 3.");
    b3 = true;
    rows = A.rows;
    cols = A.cols;
    nblocks = B.nblocks;
    nnz = 9 * B.nblocks;
    const double mean = B.nb ? static_cast<double>(B.nblocks) / B.nb : 0.0;
    G = pick_block_group(mean, group_min);
    // Blocks per lane and round: one round for rows up to 1.25x the mean
    // block count where possible (2, 3 or 4).
    K = 4;
    for (int k : {2, 3})
      if (G && G * k >= 1.25 * mean) {
        K = k;
        break;
      }
    if (G) bw.build(B, order);
    else sb.build(B, sigma, order);
  }
  // sigma: SELL-C-sigma window for the thread-per-row layout (0: none).
  // order (non-square operators only): rows are executed in this order,
  // order[p] being the row at position p (null: natural order).
  void build(const gsx::Csr &A, double group_min, int group_max = 64, int64_t sigma = 0,
             const int32_t *order = nullptr) {
    b3 = false;
    nblocks = 0;
    rows = A.rows;
    cols = A.cols;
    nnz = A.nnz();
    const bool square = A.rows == A.cols;
    const double mean = rows ? static_cast<double>(nnz) / rows : 0.0;
    G = pick_group(mean, group_min, group_max);
    // One load round for rows up to 1.25x the mean length where possible.
    K = 8;
    for (int k : {2, 4})
      if (G && G * k >= 1.25 * mean) {
        K = k;
        break;
      }
    if (G) w.build(A, square ? nullptr : order);
    else s.build(A, square ? 0 : sigma, square ? nullptr : order);
  }
};

// Square operator: scalar or 3x3 blocks, each either SELL-64 (thread per
// row) or row-contiguous (G lanes per row).
template <class T>
struct DOp {
  bool b3 = false;
  int G = 0;
  int64_t n = 0;
  int64_t nnz = 0;      // scalar entries (blocks: 9 per block)
  int64_t nblocks = 0;  // 3x3 blocks (0 for scalar storage)
  DSell<T> s;
  DBsr3<T> b;
  DRows<T> w;
  DBRows<T> bw;
  std::vector<int64_t> bptr;  // block rows: host copy of the block row pointer
  // block_key: optional rank of each block column, the order in which a
  // block row's blocks are stored and summed. group_min: average row length
  // (entries; blocks for block rows) from which rows are processed by lane
  // groups (< 0: never).
  // order (3x3-block SELL storage only): block row order[p] is stored at
  // position p (the CG operator in the level's spatial order).
  void build(const gsx::Csr &A, bool prefer_b3, const int32_t *block_key = nullptr,
             double group_min_scalar = -1.0, double group_min_block = -1.0,
             const int32_t *order = nullptr) {
    n = A.rows;
    b3 = false;
    G = 0;
    nblocks = 0;
    bptr.clear();
    if (prefer_b3) {
      gsx::Bsr3Csr B;
      if (gsx::bsr3_pack(A, B, block_key)) {
        b3 = true;
        nblocks = B.nblocks;
        bptr = B.ptr;
        nnz = 9 * B.nblocks;
        G = pick_block_group(B.nb ? static_cast<double>(B.nblocks) / B.nb : 0.0, group_min_block);
        if (G) bw.build(B);
        else if (order) b.build(B, 0, order);
        else b.build(B);
        return;
      }
    }
    nnz = A.nnz();
    G = pick_group(n ? static_cast<double>(nnz) / n : 0.0, group_min_scalar);
    if (G) w.build(A);
    else s.build(A);
  }
  // Algorithmic bytes of one y = A x (values + indices + x + y), the
  // roofline unit: scalar nnz (s + 4) + 2 n s, blocks nb (9 s + 4) + 2 n s.
  double spmv_bytes() const {
    const double sz = sizeof(T);
    if (b3) return nblocks * (9 * sz + 4) + 2.0 * n * sz;
    return nnz * (sz + 4) + 2.0 * n * sz;
  }
};

}  // namespace gsd
