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
#include <vector>

#include "host/setup.h"
#include "kernels.hpp"

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
  template <class T>
  void upload(const std::vector<T> &v) {
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

// SELL-64 device matrix. Rows keep their CSR entry order.
template <class T>
struct DSell {
  int64_t rows = 0, cols = 0;
  int64_t nnz = 0;    // stored CSR entries
  int64_t slots = 0;  // including padding
  DevBuf slice_ptr, slice_w, col, val;

  void build(const gsx::Csr &A) {
    rows = A.rows;
    cols = A.cols;
    nnz = A.nnz();
    const int64_t ns = (rows + gsk::kSlice - 1) / gsk::kSlice;
    std::vector<int64_t> sp(ns + 1, 0);
    std::vector<int32_t> sw(ns, 0);
    for (int64_t s = 0; s < ns; ++s) {
      int32_t w = 0;
      for (int64_t r = s * gsk::kSlice; r < std::min(rows, (s + 1) * gsk::kSlice); ++r)
        w = std::max(w, A.ptr[r + 1] - A.ptr[r]);
      sw[s] = w;
      sp[s + 1] = sp[s] + static_cast<int64_t>(w) * gsk::kSlice;
    }
    slots = sp[ns];
    std::vector<int32_t> c(slots, 0);
    std::vector<T> v(slots, T(0));
    for (int64_t s = 0; s < ns; ++s)
      for (int64_t l = 0; l < gsk::kSlice; ++l) {
        const int64_t r = s * gsk::kSlice + l;
        int32_t pad = 0;
        int32_t len = 0;
        if (r < rows) {
          len = A.ptr[r + 1] - A.ptr[r];
          pad = len > 0 ? A.col[A.ptr[r + 1] - 1] : static_cast<int32_t>(std::min<int64_t>(r, cols - 1));
          for (int32_t j = 0; j < len; ++j) {
            c[sp[s] + j * gsk::kSlice + l] = A.col[A.ptr[r] + j];
            v[sp[s] + j * gsk::kSlice + l] = static_cast<T>(A.val[A.ptr[r] + j]);
          }
        }
        for (int32_t j = len; j < sw[s]; ++j) c[sp[s] + j * gsk::kSlice + l] = pad;
      }
    slice_ptr.upload(sp);
    slice_w.upload(sw);
    col.upload(c);
    val.upload(v);
  }
  gsk::Sell<T> view() const {
    return gsk::Sell<T>{rows, cols, slice_ptr.as<int64_t>(), slice_w.as<int32_t>(),
                        col.as<int32_t>(), val.as<T>()};
  }
  size_t device_bytes() const {
    return slice_ptr.bytes() + slice_w.bytes() + col.bytes() + val.bytes();
  }
};

// BSR3-SELL-64 device matrix built from a host BSR3-ELL.
template <class T>
struct DBsr3 {
  int64_t nb = 0;
  int64_t nblocks = 0;  // stored blocks (without padding)
  int64_t slots = 0;
  DevBuf slice_ptr, slice_w, bcol, val;

  void build(const gsx::Bsr3Ell &B) {
    nb = B.nb;
    nblocks = B.nblocks;
    const int64_t ns = (nb + gsk::kSlice - 1) / gsk::kSlice;
    std::vector<int64_t> sp(ns + 1, 0);
    std::vector<int32_t> sw(ns, 0);
    for (int64_t s = 0; s < ns; ++s) {
      int32_t w = 0;
      for (int64_t b = s * gsk::kSlice; b < std::min(nb, (s + 1) * gsk::kSlice); ++b)
        w = std::max(w, B.nblk[b]);
      sw[s] = w;
      sp[s + 1] = sp[s] + static_cast<int64_t>(w) * gsk::kSlice;
    }
    slots = sp[ns];
    std::vector<int32_t> c(slots, 0);
    std::vector<T> v(slots * 9, T(0));
    for (int64_t s = 0; s < ns; ++s)
      for (int64_t l = 0; l < gsk::kSlice; ++l) {
        const int64_t b = s * gsk::kSlice + l;
        for (int32_t j = 0; j < sw[s]; ++j) {
          const int64_t slot = sp[s] + static_cast<int64_t>(j) * gsk::kSlice;
          if (b < nb && j < B.nblk[b]) {
            c[slot + l] = B.bcol[static_cast<int64_t>(j) * nb + b];
            for (int e = 0; e < 9; ++e)
              v[slot * 9 + e * gsk::kSlice + l] =
                  static_cast<T>(B.val[(static_cast<int64_t>(j) * 9 + e) * nb + b]);
          } else {
            c[slot + l] = static_cast<int32_t>(b < nb ? b : 0);
          }
        }
      }
    slice_ptr.upload(sp);
    slice_w.upload(sw);
    bcol.upload(c);
    val.upload(v);
  }
  gsk::Bsr3<T> view() const {
    return gsk::Bsr3<T>{nb, slice_ptr.as<int64_t>(), slice_w.as<int32_t>(), bcol.as<int32_t>(),
                        val.as<T>()};
  }
  size_t device_bytes() const {
    return slice_ptr.bytes() + slice_w.bytes() + bcol.bytes() + val.bytes();
  }
};

// Square operator stored either as scalar SELL or as BSR3-SELL.
template <class T>
struct DOp {
  bool b3 = false;
  int64_t n = 0;
  DSell<T> s;
  DBsr3<T> b;
  void build(const gsx::Csr &A, bool prefer_b3) {
    n = A.rows;
    b3 = false;
    if (prefer_b3) {
      gsx::Bsr3Ell B;
      if (gsx::bsr3_pack(A, B)) {
        b.build(B);
        b3 = true;
        return;
      }
    }
    s.build(A);
  }
  // Algorithmic bytes of one y = A x (values + indices + x + y), the
  // roofline unit: scalar nnz (s + 4) + 2 n s, blocks nb (9 s + 4) + 2 n s.
  double spmv_bytes() const {
    const double sz = sizeof(T);
    if (b3) return b.nblocks * (9 * sz + 4) + 2.0 * n * sz;
    return s.nnz * (sz + 4) + 2.0 * n * sz;
  }
};

}  // namespace gsd
