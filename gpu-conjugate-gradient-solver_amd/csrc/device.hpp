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

// SELL-64 device matrix (layout and validation in gsx::sell64_pack).
template <class T>
struct DSell {
  int64_t rows = 0, cols = 0;
  int64_t nnz = 0;    // stored CSR entries
  int64_t slots = 0;  // including padding
  DevBuf slice_ptr, slice_w, col, val;

  void build(const gsx::Csr &A) {
    const gsx::Sell64 P = gsx::sell64_pack(A);
    rows = P.rows;
    cols = P.cols;
    nnz = P.nnz;
    slots = P.slots;
    slice_ptr.upload(P.slice_ptr);
    slice_w.upload(P.slice_w);
    col.upload(P.col);
    val.upload(cast_values<T>(P.val));
  }
  gsk::Sell<T> view() const {
    return gsk::Sell<T>{rows, cols, slice_ptr.as<int64_t>(), slice_w.as<int32_t>(),
                        col.as<int32_t>(), val.as<T>()};
  }
  size_t device_bytes() const {
    return slice_ptr.bytes() + slice_w.bytes() + col.bytes() + val.bytes();
  }
};

// BSR3-SELL-64 device matrix (layout in gsx::bsr3_sell64_pack).
template <class T>
struct DBsr3 {
  int64_t nb = 0;
  int64_t nblocks = 0;  // stored blocks (without padding)
  int64_t slots = 0;
  DevBuf slice_ptr, slice_w, bcol, val;

  void build(const gsx::Bsr3Ell &B) {
    const gsx::Bsr3Sell64 P = gsx::bsr3_sell64_pack(B);
    nb = P.nb;
    nblocks = P.nblocks;
    slots = P.slots;
    slice_ptr.upload(P.slice_ptr);
    slice_w.upload(P.slice_w);
    bcol.upload(P.bcol);
    val.upload(cast_values<T>(P.val));
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
