// HIP implementation of the reference's Backend plugin interface
// (include/GPUSolver/Backend.h:15-159): device objects (DeviceVector,
// DeviceIndexVector, DeviceScalar, DeviceCSRMatrix, DeviceELLMatrix), the
// encoder (one HIP stream, as the CUDA backend's one stream per encoder,
// src/cuda/CudaEncoder.cc:7-17) and every encode* method as a HIP kernel.
//
// Semantics follow the reference's CPU backend (src/CPU/CPUBackend.cc):
// row sums in CSR order, Gauss-Seidel's sum-then-divide, block
// Gauss-Seidel's off-block sum then 3x3 inverse, plain IEEE scalar division
// and square root. Values are FP32 as in the reference, or FP64 when the
// backend is created with GS_DTYPE_F64 (the interface extension that the
// FP64 oracle comparisons use). Dot products are deterministic two-level
// reductions that write the device scalar in the same launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gpusolver.h"
#include "device.hpp"

using gsd::DevBuf;
using gsk::grid_for;
using gsk::kBlock;

namespace gsc {
void set_error(const char *msg);  // shared with the rest of the C ABI
}

namespace {
template <class F>
int bguard(F &&f) {
  try {
    f();
    return GS_OK;
  } catch (const std::exception &e) {
    gsc::set_error(e.what());
    return GS_ERR;
  }
}
}  // namespace

// The backend reports through the library-wide error slot.
extern "C" const char *gs_backend_last_error(void) { return gs_last_error(); }

struct gs_backend_s {
  int dtype = GS_DTYPE_F32;
  size_t esize() const { return dtype == GS_DTYPE_F64 ? 8 : 4; }
};

struct gs_encoder_s {
  gs_backend_s *be = nullptr;
  hipStream_t stream = nullptr;
  DevBuf scratch;  // dot-product partials, then the zero-diagonal flag
  bool submitted = false;
  ~gs_encoder_s() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

struct gs_dvec_s {
  gs_backend_s *be = nullptr;
  int64_t n = 0;
  DevBuf d;
};
struct gs_divec_s {
  int64_t n = 0;
  DevBuf d;
};
struct gs_dscalar_s {
  gs_backend_s *be = nullptr;
  DevBuf d;
};
struct gs_dcsr_s {
  gs_backend_s *be = nullptr;
  int64_t rows = 0, cols = 0, nnz = 0;
  DevBuf ptr, col, val;
};
struct gs_dell_s {
  gs_backend_s *be = nullptr;
  int64_t rows = 0, cols = 0, width = 0, ovf_nnz = 0;
  DevBuf col, val, optr, ocol, oval;
};

namespace gsb {

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_scale(T *x, T a, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) x[i] *= a;
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_scale_dev(T *x, const T *a, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) x[i] *= *a;
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_copy(const T *x, T *y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] = x[i];
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_scale_vector(const T *s, T *o, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) o[i] *= s[i];
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_zero(T *x, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) x[i] = T(0);
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_axpy(T a, const T *x, T *y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_axpy_dev(const T *a, const T *x, T *y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] += (*a) * x[i];
}

// src/CPU/CPUBackend.cc:217-236: one row per thread, CSR order.
template <class T>
__global__ __launch_bounds__(kBlock) void k_csr_spmv(const int32_t *ptr, const int32_t *col,
                                                     const T *val, const T *x, T *y,
                                                     int64_t rows) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= rows) return;
  T s = T(0);
  for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) s += val[k] * x[col[k]];
  y[r] = s;
}

// Slot-major ELL then overflow CSR (kernels/metal/SpMVKernels.metal:21-91
// for the layout; the row sum visits the slots in order, then overflow).
template <class T>
__global__ __launch_bounds__(kBlock) void k_ell_spmv(const int32_t *ecol, const T *eval,
                                                     int64_t width, const int32_t *optr,
                                                     const int32_t *ocol, const T *oval,
                                                     const T *x, T *y, int64_t rows) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= rows) return;
  T s = T(0);
  for (int64_t k = 0; k < width; ++k) s += eval[k * rows + r] * x[ecol[k * rows + r]];
  for (int32_t k = optr[r]; k < optr[r + 1]; ++k) s += oval[k] * x[ocol[k]];
  y[r] = s;
}

// src/CPU/CPUBackend.cc:238-268 with rows taken from colorVertices.
template <class T>
__global__ __launch_bounds__(kBlock) void k_csr_gs_color(const int32_t *ptr, const int32_t *col,
                                                         const T *val, const int32_t *order,
                                                         int64_t start, int64_t count,
                                                         int64_t rows, const T *rhs, T *x,
                                                         T omega, int32_t *zero_diag) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= count) return;
  const int32_t row = order[start + k];
  if (row < 0 || row >= rows) {  // invalid colorVertices entry
    *zero_diag = 2;
    return;
  }
  T diag = T(0), sum = T(0);
  for (int32_t e = ptr[row]; e < ptr[row + 1]; ++e) {
    const int32_t c = col[e];
    if (c == row) {
      diag = val[e];
      continue;
    }
    sum += val[e] * x[c];
  }
  if (fabs(static_cast<double>(diag)) < 1e-12) {
    *zero_diag = 1;
    return;
  }
  const T nv = (rhs[row] - sum) / diag;
  x[row] = (T(1) - omega) * x[row] + omega * nv;
}

// src/CPU/CPUBackend.cc:350-387, blocks colorBlocks[begin .. end).
template <class T>
__global__ __launch_bounds__(kBlock) void k_csr_block_gs_color(
    const int32_t *ptr, const int32_t *col, const T *val, const int32_t *order, const T *dinv,
    int64_t begin, int64_t end, int64_t nb, const T *rhs, T *x, T omega, int32_t *bad) {
  const int64_t k = begin + static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= end) return;
  const int64_t b = order[k];
  if (b < 0 || b >= nb) {  // invalid colorBlocks entry
    *bad = 2;
    return;
  }
  const int64_t r0 = 3 * b;
  T lr[3];
  for (int l = 0; l < 3; ++l) {
    T s = T(0);
    for (int32_t e = ptr[r0 + l]; e < ptr[r0 + l + 1]; ++e) {
      const int32_t c = col[e];
      if (c >= r0 && c < r0 + 3) continue;
      s += val[e] * x[c];
    }
    lr[l] = rhs[r0 + l] - s;
  }
  const T *bi = dinv + 9 * b;
  for (int l = 0; l < 3; ++l) {
    T nv = T(0);
    for (int c = 0; c < 3; ++c) nv += bi[3 * l + c] * lr[c];
    x[r0 + l] = (T(1) - omega) * x[r0 + l] + omega * nv;
  }
}

// src/CPU/CPUBackend.cc:323-348.
template <class T>
__global__ __launch_bounds__(kBlock) void k_block_inv(const T *inv, const T *res, T *corr,
                                                      int64_t nb) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (b >= nb) return;
  for (int r = 0; r < 3; ++r) {
    T s = T(0);
    for (int c = 0; c < 3; ++c) s += inv[9 * b + 3 * r + c] * res[3 * b + c];
    corr[3 * b + r] = s;
  }
}

enum ScalarOp { kDiv, kMul, kSet, kCopy, kNeg, kSqrt };
template <class T>
__global__ void k_scalar(int op, const T *a, const T *b, T v, T *out) {
  switch (op) {
    case kDiv: *out = *a / *b; break;
    case kMul: *out = *a * *b; break;
    case kSet: *out = v; break;
    case kCopy: *out = *a; break;
    case kNeg: *out = -*a; break;
    case kSqrt: *out = sqrt(*a); break;
  }
}

// ---------------------------------------------------------------------------
// Host helpers.
// ---------------------------------------------------------------------------
void need(const void *p, const char *what) {
  if (!p) throw std::runtime_error(std::string("HipBackend: ") + what + " must not be null.");
}

bool f64(const gs_backend_s *b) { return b->dtype == GS_DTYPE_F64; }

void same_size(const gs_dvec_s *a, const gs_dvec_s *b, const char *who) {
  if (a->n != b->n) throw std::runtime_error(std::string("HipBackend::") + who + ": size mismatch.");
}

template <class T>
T *dp(const DevBuf &b) {
  return static_cast<T *>(b.get());
}

int32_t *zero_diag_flag(gs_encoder_s *e) {
  return reinterpret_cast<int32_t *>(static_cast<char *>(e->scratch.get()) +
                                     (gsk::kDotMaxGrid + 32) * sizeof(double));
}

template <class F>
void dispatch(const gs_backend_s *be, F &&f) {
  if (f64(be)) f(double());
  else f(float());
}

}  // namespace gsb

using namespace gsb;

extern "C" {

int gs_backend_create(int32_t dtype, gs_backend_h *out) {
  return bguard([&] {
    need(out, "out");
    if (dtype != GS_DTYPE_F32 && dtype != GS_DTYPE_F64)
      throw std::runtime_error("HipBackend: unknown value type.");
    int dev = 0;
    GS_HIP(hipGetDeviceCount(&dev));
    if (dev <= 0) throw std::runtime_error("HipBackend: no HIP device.");
    auto b = std::make_unique<gs_backend_s>();
    b->dtype = dtype;
    *out = b.release();
  });
}

int gs_backend_destroy(gs_backend_h b) {
  delete b;
  return GS_OK;
}

int gs_backend_create_encoder(gs_backend_h b, gs_encoder_h *out) {
  return bguard([&] {
    need(b, "backend");
    auto e = std::make_unique<gs_encoder_s>();
    e->be = b;
    GS_HIP(hipStreamCreate(&e->stream));
    e->scratch.alloc((gsk::kDotMaxGrid + 64) * sizeof(double));
    // The zero-diagonal flag starts at zero (a fill kernel, so the stream
    // stays capturable).
    k_zero<double><<<1, kBlock, 0, e->stream>>>(
        reinterpret_cast<double *>(static_cast<char *>(e->scratch.get()) + gsk::kDotMaxGrid * sizeof(double)),
        64);
    *out = e.release();
  });
}

int gs_backend_submit(gs_backend_h, gs_encoder_h e) {
  return bguard([&] {
    need(e, "encoder");
    e->submitted = true;  // work is already queued on the encoder's stream
  });
}

int gs_backend_submit_and_wait(gs_backend_h, gs_encoder_h e) {
  return bguard([&] {
    need(e, "encoder");
    e->submitted = true;
    int32_t flag = 0;
    GS_HIP(hipMemcpyAsync(&flag, zero_diag_flag(e), sizeof(flag), hipMemcpyDeviceToHost, e->stream));
    GS_HIP(hipStreamSynchronize(e->stream));
    if (flag) {
      k_zero<int32_t><<<1, kBlock, 0, e->stream>>>(zero_diag_flag(e), 1);
      GS_HIP(hipStreamSynchronize(e->stream));
      if (flag == 2)
        throw std::runtime_error("HipBackend: color vertex index out of range.");
      throw std::runtime_error("HipBackend::encodeGaussSeidelColor: zero diagonal entry.");
    }
  });
}

int gs_encoder_destroy(gs_encoder_h e) {
  delete e;
  return GS_OK;
}

int gs_encoder_stream(gs_encoder_h e, void **stream) {
  return bguard([&] {
    need(e, "encoder");
    need(stream, "output");
    *stream = static_cast<void *>(e->stream);
  });
}

int gs_backend_create_vector(gs_backend_h b, int64_t n, const void *values, gs_dvec_h *out) {
  return bguard([&] {
    need(b, "backend");
    auto v = std::make_unique<gs_dvec_s>();
    v->be = b;
    v->n = n;
    v->d.alloc(n * b->esize());
    if (values) GS_HIP(hipMemcpy(v->d.get(), values, n * b->esize(), hipMemcpyHostToDevice));
    else GS_HIP(hipMemset(v->d.get(), 0, n * b->esize()));
    *out = v.release();
  });
}

int gs_dvec_size(gs_dvec_h v, int64_t *n) {
  return bguard([&] {
    need(v, "vector");
    *n = v->n;
  });
}

int gs_dvec_update(gs_dvec_h v, const void *values, int64_t n) {
  return bguard([&] {
    need(v, "vector");
    need(values, "values");
    if (n != v->n) throw std::runtime_error("DeviceVector::updateValues: size mismatch.");
    GS_HIP(hipMemcpy(v->d.get(), values, n * v->be->esize(), hipMemcpyHostToDevice));
  });
}

int gs_dvec_download(gs_dvec_h v, void *out) {
  return bguard([&] {
    need(v, "vector");
    need(out, "output");
    GS_HIP(hipDeviceSynchronize());
    GS_HIP(hipMemcpy(out, v->d.get(), v->n * v->be->esize(), hipMemcpyDeviceToHost));
  });
}

int gs_dvec_destroy(gs_dvec_h v) {
  delete v;
  return GS_OK;
}

int gs_dvec_data(gs_dvec_h v, void **ptr) {
  return bguard([&] {
    need(v, "vector");
    need(ptr, "output");
    *ptr = v->d.get();
  });
}

int gs_backend_create_index_vector(gs_backend_h b, int64_t n, const int32_t *values,
                                   gs_divec_h *out) {
  return bguard([&] {
    need(b, "backend");
    auto v = std::make_unique<gs_divec_s>();
    v->n = n;
    v->d.alloc(n * sizeof(int32_t));
    if (values) GS_HIP(hipMemcpy(v->d.get(), values, n * sizeof(int32_t), hipMemcpyHostToDevice));
    else GS_HIP(hipMemset(v->d.get(), 0, n * sizeof(int32_t)));
    *out = v.release();
  });
}

int gs_divec_download(gs_divec_h v, int32_t *out) {
  return bguard([&] {
    need(v, "index vector");
    GS_HIP(hipDeviceSynchronize());
    GS_HIP(hipMemcpy(out, v->d.get(), v->n * sizeof(int32_t), hipMemcpyDeviceToHost));
  });
}

int gs_divec_destroy(gs_divec_h v) {
  delete v;
  return GS_OK;
}

int gs_divec_size(gs_divec_h v, int64_t *n) {
  return bguard([&] {
    need(v, "index vector");
    need(n, "output");
    *n = v->n;
  });
}

// src/CPU/CPUDeviceIndexVector.cc:24-31: same size, or an error.
int gs_divec_update(gs_divec_h v, const int32_t *values, int64_t n) {
  return bguard([&] {
    need(v, "index vector");
    need(values, "values");
    if (n != v->n) throw std::runtime_error("DeviceIndexVector::updateValues: size mismatch.");
    GS_HIP(hipMemcpy(v->d.get(), values, n * sizeof(int32_t), hipMemcpyHostToDevice));
  });
}

int gs_backend_create_scalar(gs_backend_h b, double value, gs_dscalar_h *out) {
  return bguard([&] {
    need(b, "backend");
    auto s = std::make_unique<gs_dscalar_s>();
    s->be = b;
    s->d.alloc(8);
    if (f64(b)) GS_HIP(hipMemcpy(s->d.get(), &value, 8, hipMemcpyHostToDevice));
    else {
      const float f = static_cast<float>(value);
      GS_HIP(hipMemcpy(s->d.get(), &f, 4, hipMemcpyHostToDevice));
    }
    *out = s.release();
  });
}

int gs_dscalar_download(gs_dscalar_h s, double *value) {
  return bguard([&] {
    need(s, "scalar");
    GS_HIP(hipDeviceSynchronize());
    if (f64(s->be)) GS_HIP(hipMemcpy(value, s->d.get(), 8, hipMemcpyDeviceToHost));
    else {
      float f;
      GS_HIP(hipMemcpy(&f, s->d.get(), 4, hipMemcpyDeviceToHost));
      *value = f;
    }
  });
}

int gs_dscalar_destroy(gs_dscalar_h s) {
  delete s;
  return GS_OK;
}

int gs_dscalar_data(gs_dscalar_h s, void **ptr) {
  return bguard([&] {
    need(s, "scalar");
    need(ptr, "output");
    *ptr = s->d.get();
  });
}

int gs_backend_create_csr(gs_backend_h b, gs_csr_h host, gs_dcsr_h *out) {
  return bguard([&] {
    need(b, "backend");
    need(host, "matrix");
    int64_t rows, cols, nnz;
    if (gs_csr_info(host, &rows, &cols, &nnz) != GS_OK) throw std::runtime_error(gs_last_error());
    std::vector<int32_t> ptr(rows + 1), col(nnz);
    std::vector<double> val(nnz);
    if (gs_csr_export(host, ptr.data(), col.data(), val.data()) != GS_OK)
      throw std::runtime_error(gs_last_error());
    auto m = std::make_unique<gs_dcsr_s>();
    m->be = b;
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->ptr.upload(ptr);
    m->col.upload(col);
    if (f64(b)) m->val.upload(val);
    else m->val.upload(gsd::cast_values<float>(val));
    *out = m.release();
  });
}

int gs_dcsr_update_values(gs_dcsr_h m, const double *values, int64_t nnz) {
  return bguard([&] {
    need(m, "matrix");
    need(values, "values");
    if (nnz != m->nnz) throw std::runtime_error("DeviceCSRMatrix::updateValues: size mismatch.");
    std::vector<double> v(values, values + nnz);
    if (f64(m->be)) GS_HIP(hipMemcpy(m->val.get(), v.data(), nnz * 8, hipMemcpyHostToDevice));
    else {
      std::vector<float> f = gsd::cast_values<float>(v);
      GS_HIP(hipMemcpy(m->val.get(), f.data(), nnz * 4, hipMemcpyHostToDevice));
    }
  });
}

int gs_dcsr_destroy(gs_dcsr_h m) {
  delete m;
  return GS_OK;
}

int gs_dcsr_info(gs_dcsr_h m, int64_t *rows, int64_t *cols, int64_t *nnz) {
  return bguard([&] {
    need(m, "matrix");
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    if (nnz) *nnz = m->nnz;
  });
}

int gs_backend_create_ell(gs_backend_h b, gs_csr_h host, int32_t width, gs_dell_h *out) {
  return bguard([&] {
    need(b, "backend");
    need(host, "matrix");
    int64_t rows, cols, nnz, ov = 0;
    if (gs_csr_info(host, &rows, &cols, &nnz) != GS_OK) throw std::runtime_error(gs_last_error());
    if (gs_ell_pack(host, width, nullptr, nullptr, nullptr, nullptr, nullptr, &ov) != GS_OK)
      throw std::runtime_error(gs_last_error());
    std::vector<int32_t> ec(rows * width), op(rows + 1), oc(std::max<int64_t>(ov, 1));
    std::vector<double> ev(rows * width), ovv(std::max<int64_t>(ov, 1));
    if (gs_ell_pack(host, width, ec.data(), ev.data(), op.data(), oc.data(), ovv.data(), &ov) !=
        GS_OK)
      throw std::runtime_error(gs_last_error());
    auto m = std::make_unique<gs_dell_s>();
    m->be = b;
    m->rows = rows;
    m->cols = cols;
    m->width = width;
    m->ovf_nnz = ov;
    m->col.upload(ec);
    m->optr.upload(op);
    m->ocol.upload(oc);
    if (f64(b)) {
      m->val.upload(ev);
      m->oval.upload(ovv);
    } else {
      m->val.upload(gsd::cast_values<float>(ev));
      m->oval.upload(gsd::cast_values<float>(ovv));
    }
    *out = m.release();
  });
}

int gs_dell_destroy(gs_dell_h m) {
  delete m;
  return GS_OK;
}

// An ELL matrix already packed on the host (the arrays HostELLMatrix holds,
// include/GPUSolver/HostELLMatrix.h:19-31: slot-major ELL, overflow CSR).
// Every column index and the overflow row pointer are checked here, since
// the SpMV kernel reads x through them unchecked.
int gs_backend_create_ell_packed(gs_backend_h b, int64_t rows, int64_t cols, int64_t width,
                                 const int32_t *ell_col, const void *ell_val, int64_t ovf_nnz,
                                 const int32_t *ovf_ptr, const int32_t *ovf_col,
                                 const void *ovf_val, gs_dell_h *out) {
  return bguard([&] {
    need(b, "backend");
    need(out, "output");
    if (rows <= 0 || cols <= 0)
      throw std::runtime_error("DeviceELLMatrix: matrix dimensions must be nonzero.");
    if (width <= 0) throw std::runtime_error("DeviceELLMatrix: ELL width must be nonzero.");
    if (ovf_nnz < 0) throw std::runtime_error("DeviceELLMatrix: negative overflow count.");
    if (!ell_col || !ell_val || !ovf_ptr || (ovf_nnz > 0 && (!ovf_col || !ovf_val)))
      throw std::runtime_error("DeviceELLMatrix: storage pointers must not be null.");
    const int64_t slots = rows * width;
    for (int64_t i = 0; i < slots; ++i)
      if (ell_col[i] < 0 || ell_col[i] >= cols)
        throw std::runtime_error("DeviceELLMatrix: column index out of bounds.");
    if (ovf_ptr[0] != 0 || ovf_ptr[rows] != ovf_nnz)
      throw std::runtime_error("DeviceELLMatrix: overflow row pointer does not match its count.");
    for (int64_t r = 0; r < rows; ++r)
      if (ovf_ptr[r + 1] < ovf_ptr[r])
        throw std::runtime_error("DeviceELLMatrix: invalid overflow row pointer.");
    for (int64_t k = 0; k < ovf_nnz; ++k)
      if (ovf_col[k] < 0 || ovf_col[k] >= cols)
        throw std::runtime_error("DeviceELLMatrix: column index out of bounds.");
    const size_t es = b->esize();
    auto m = std::make_unique<gs_dell_s>();
    m->be = b;
    m->rows = rows;
    m->cols = cols;
    m->width = width;
    m->ovf_nnz = ovf_nnz;
    m->col.upload(ell_col, static_cast<size_t>(slots));
    m->val.alloc(slots * es);
    GS_HIP(hipMemcpy(m->val.get(), ell_val, slots * es, hipMemcpyHostToDevice));
    m->optr.upload(ovf_ptr, static_cast<size_t>(rows + 1));
    m->ocol.alloc(std::max<int64_t>(ovf_nnz, 1) * sizeof(int32_t));
    m->oval.alloc(std::max<int64_t>(ovf_nnz, 1) * es);
    if (ovf_nnz > 0) {
      GS_HIP(hipMemcpy(m->ocol.get(), ovf_col, ovf_nnz * sizeof(int32_t), hipMemcpyHostToDevice));
      GS_HIP(hipMemcpy(m->oval.get(), ovf_val, ovf_nnz * es, hipMemcpyHostToDevice));
    }
    *out = m.release();
  });
}

int gs_dell_info(gs_dell_h m, int64_t *rows, int64_t *cols, int64_t *width, int64_t *ovf_nnz) {
  return bguard([&] {
    need(m, "matrix");
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    if (width) *width = m->width;
    if (ovf_nnz) *ovf_nnz = m->ovf_nnz;
  });
}

// ---- encode* (include/GPUSolver/Backend.h:47-114) --------------------------
int gs_encode_scale(gs_encoder_h e, gs_dvec_h x, double alpha) {
  return bguard([&] {
    need(e, "encoder");
    need(x, "vector");
    if (x->n == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_scale<T><<<grid_for(x->n), kBlock, 0, e->stream>>>(dp<T>(x->d), static_cast<T>(alpha), x->n);
    });
  });
}

int gs_encode_scale_dev(gs_encoder_h e, gs_dvec_h x, gs_dscalar_h alpha) {
  return bguard([&] {
    need(e, "encoder");
    need(x, "vector");
    need(alpha, "scalar");
    if (x->n == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_scale_dev<T><<<grid_for(x->n), kBlock, 0, e->stream>>>(dp<T>(x->d), dp<T>(alpha->d), x->n);
    });
  });
}

int gs_encode_copy(gs_encoder_h e, gs_dvec_h x, gs_dvec_h y) {
  return bguard([&] {
    need(e, "encoder");
    need(x, "vector");
    need(y, "vector");
    same_size(x, y, "encodeCopy");
    if (x->n == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_copy<T><<<grid_for(x->n), kBlock, 0, e->stream>>>(dp<T>(x->d), dp<T>(y->d), x->n);
    });
  });
}

int gs_encode_scale_vector(gs_encoder_h e, gs_dvec_h s, gs_dvec_h out) {
  return bguard([&] {
    need(e, "encoder");
    need(s, "vector");
    need(out, "vector");
    same_size(s, out, "encodeScaleVector");
    if (s->n == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_scale_vector<T><<<grid_for(s->n), kBlock, 0, e->stream>>>(dp<T>(s->d), dp<T>(out->d), s->n);
    });
  });
}

int gs_encode_set_zero(gs_encoder_h e, gs_dvec_h x) {
  return bguard([&] {
    need(e, "encoder");
    need(x, "vector");
    if (x->n == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_zero<T><<<grid_for(x->n), kBlock, 0, e->stream>>>(dp<T>(x->d), x->n);
    });
  });
}

int gs_encode_axpy(gs_encoder_h e, double alpha, gs_dvec_h x, gs_dvec_h y) {
  return bguard([&] {
    need(e, "encoder");
    need(x, "vector");
    need(y, "vector");
    same_size(x, y, "encodeAxpy");
    if (x->n == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_axpy<T><<<grid_for(x->n), kBlock, 0, e->stream>>>(static_cast<T>(alpha), dp<T>(x->d),
                                                          dp<T>(y->d), x->n);
    });
  });
}

int gs_encode_axpy_dev(gs_encoder_h e, gs_dscalar_h alpha, gs_dvec_h x, gs_dvec_h y) {
  return bguard([&] {
    need(e, "encoder");
    need(alpha, "scalar");
    need(x, "vector");
    need(y, "vector");
    same_size(x, y, "encodeAxpy");
    if (x->n == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_axpy_dev<T><<<grid_for(x->n), kBlock, 0, e->stream>>>(dp<T>(alpha->d), dp<T>(x->d),
                                                              dp<T>(y->d), x->n);
    });
  });
}

int gs_encode_spmv(gs_encoder_h e, gs_dcsr_h A, gs_dvec_h x, gs_dvec_h y) {
  return bguard([&] {
    need(e, "encoder");
    need(A, "matrix");
    need(x, "vector");
    need(y, "vector");
    if (x->n != A->cols || y->n != A->rows)
      throw std::runtime_error("HipBackend::encodeSpmv: size mismatch.");
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_csr_spmv<T><<<grid_for(A->rows), kBlock, 0, e->stream>>>(
          dp<int32_t>(A->ptr), dp<int32_t>(A->col), dp<T>(A->val), dp<T>(x->d), dp<T>(y->d),
          A->rows);
    });
  });
}

int gs_encode_spmv_ell(gs_encoder_h e, gs_dell_h A, gs_dvec_h x, gs_dvec_h y) {
  return bguard([&] {
    need(e, "encoder");
    need(A, "matrix");
    need(x, "vector");
    need(y, "vector");
    if (x->n != A->cols || y->n != A->rows)
      throw std::runtime_error("HipBackend::encodeSpmvELL: size mismatch.");
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_ell_spmv<T><<<grid_for(A->rows), kBlock, 0, e->stream>>>(
          dp<int32_t>(A->col), dp<T>(A->val), A->width, dp<int32_t>(A->optr),
          dp<int32_t>(A->ocol), dp<T>(A->oval), dp<T>(x->d), dp<T>(y->d), A->rows);
    });
  });
}

int gs_encode_gauss_seidel_color(gs_encoder_h e, gs_dcsr_h A, gs_divec_h colors, int64_t start,
                                 int64_t count, gs_dvec_h rhs, gs_dvec_h x, double omega) {
  return bguard([&] {
    need(e, "encoder");
    need(A, "matrix");
    need(colors, "color vertices");
    need(rhs, "vector");
    need(x, "vector");
    if (start < 0 || count < 0 || start + count > colors->n)
      throw std::runtime_error("HipBackend::encodeGaussSeidelColor: color range exceeds colorVertices size.");
    if (rhs->n != A->rows || x->n != A->rows)
      throw std::runtime_error("HipBackend::encodeGaussSeidelColor: size mismatch.");
    if (count == 0) return;
    // A zero diagonal makes the CPU backend throw (src/CPU/CPUBackend.cc:
    // 256-266); the kernel raises the encoder's sticky flag, which
    // submitAndWait turns into the same exception.
    int32_t *flag = zero_diag_flag(e);
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_csr_gs_color<T><<<grid_for(count), kBlock, 0, e->stream>>>(
          dp<int32_t>(A->ptr), dp<int32_t>(A->col), dp<T>(A->val), dp<int32_t>(colors->d), start,
          count, A->rows, dp<T>(rhs->d), dp<T>(x->d), static_cast<T>(omega), flag);
    });
  });
}

int gs_encode_block_gauss_seidel_color3x3(gs_encoder_h e, gs_dcsr_h A, gs_divec_h blocks,
                                          gs_dvec_h dinv, int64_t begin, int64_t end,
                                          gs_dvec_h rhs, gs_dvec_h x, double omega) {
  return bguard([&] {
    need(e, "encoder");
    need(A, "matrix");
    need(blocks, "color blocks");
    need(dinv, "inverse blocks");
    need(rhs, "vector");
    need(x, "vector");
    if (end < begin)
      throw std::runtime_error("HipBackend::encodeBlockGaussSeidelColor3x3: invalid color range.");
    if (end > blocks->n)
      throw std::runtime_error(
          "HipBackend::encodeBlockGaussSeidelColor3x3: color range exceeds colorBlocks size.");
    if (rhs->n != x->n)
      throw std::runtime_error("HipBackend::encodeBlockGaussSeidelColor3x3: rhs/solution size mismatch.");
    if (A->rows != rhs->n || A->cols != rhs->n)
      throw std::runtime_error("HipBackend::encodeBlockGaussSeidelColor3x3: matrix/vector size mismatch.");
    if (rhs->n % 3 != 0)
      throw std::runtime_error(
          "HipBackend::encodeBlockGaussSeidelColor3x3: vector size must be divisible by 3.");
    if (dinv->n != 3 * rhs->n)
      throw std::runtime_error(
          "HipBackend::encodeBlockGaussSeidelColor3x3: wrong diagonal inverse array size.");
    if (end == begin) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_csr_block_gs_color<T><<<grid_for(end - begin), kBlock, 0, e->stream>>>(
          dp<int32_t>(A->ptr), dp<int32_t>(A->col), dp<T>(A->val), dp<int32_t>(blocks->d),
          dp<T>(dinv->d), begin, end, rhs->n / 3, dp<T>(rhs->d), dp<T>(x->d),
          static_cast<T>(omega), zero_diag_flag(e));
    });
  });
}

int gs_encode_apply_block_inverses3x3(gs_encoder_h e, gs_dvec_h inv, gs_dvec_h res, gs_dvec_h corr) {
  return bguard([&] {
    need(e, "encoder");
    need(inv, "inverse blocks");
    need(res, "vector");
    need(corr, "vector");
    same_size(res, corr, "encodeApplyBlockInverses3x3");
    if (inv->n != 3 * res->n)
      throw std::runtime_error("HipBackend::encodeApplyBlockInverses3x3: wrong inverse array size.");
    const int64_t nb = res->n / 3;
    if (nb == 0) return;
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_block_inv<T><<<grid_for(nb), kBlock, 0, e->stream>>>(dp<T>(inv->d), dp<T>(res->d),
                                                             dp<T>(corr->d), nb);
    });
  });
}

int gs_encode_dot(gs_encoder_h e, gs_dvec_h x, gs_dvec_h y, gs_dscalar_h out) {
  return bguard([&] {
    need(e, "encoder");
    need(x, "vector");
    need(y, "vector");
    need(out, "scalar");
    same_size(x, y, "encodeDot");
    char *base = static_cast<char *>(e->scratch.get());
    const uint32_t G = gsk::dot_grid(x->n);
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      gsk::k_dot<T><<<G, kBlock, 0, e->stream>>>(dp<T>(x->d), dp<T>(y->d), x->n,
                                            reinterpret_cast<T *>(base));
      gsk::k_dot_final<T><<<1, kBlock, 0, e->stream>>>(reinterpret_cast<const T *>(base), G,
                                                  dp<T>(out->d));
    });
  });
}

static int scalar_op(gs_encoder_h e, int op, gs_dscalar_h a, gs_dscalar_h b, double v,
                     gs_dscalar_h out) {
  return bguard([&] {
    need(e, "encoder");
    need(out, "scalar");
    if (op != kSet) need(a, "scalar");
    if (op == kDiv || op == kMul) need(b, "scalar");
    dispatch(e->be, [&](auto t) {
      using T = decltype(t);
      k_scalar<T><<<1, 1, 0, e->stream>>>(op, a ? dp<T>(a->d) : nullptr, b ? dp<T>(b->d) : nullptr,
                                          static_cast<T>(v), dp<T>(out->d));
    });
  });
}

int gs_encode_scalar_divide(gs_encoder_h e, gs_dscalar_h num, gs_dscalar_h den, gs_dscalar_h out) {
  return scalar_op(e, kDiv, num, den, 0.0, out);
}
int gs_encode_scalar_multiply(gs_encoder_h e, gs_dscalar_h a, gs_dscalar_h b, gs_dscalar_h out) {
  return scalar_op(e, kMul, a, b, 0.0, out);
}
int gs_encode_scalar_set(gs_encoder_h e, double value, gs_dscalar_h out) {
  return scalar_op(e, kSet, nullptr, nullptr, value, out);
}
int gs_encode_scalar_copy(gs_encoder_h e, gs_dscalar_h in, gs_dscalar_h out) {
  return scalar_op(e, kCopy, in, nullptr, 0.0, out);
}
int gs_encode_scalar_negate(gs_encoder_h e, gs_dscalar_h in, gs_dscalar_h out) {
  return scalar_op(e, kNeg, in, nullptr, 0.0, out);
}
int gs_encode_scalar_sqrt(gs_encoder_h e, gs_dscalar_h in, gs_dscalar_h out) {
  return scalar_op(e, kSqrt, in, nullptr, 0.0, out);
}

}  // extern "C"
