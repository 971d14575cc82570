// gpuSolver::HipBackend over the C ABI: every Backend method
// (include/GPUSolver/Backend.h:15-159) forwards to the gs_* call that
// implements it; this file needs only the reference's Eigen-free interface
// headers. Objects created by another backend are rejected (dynamic_cast),
// where the reference's backends would cast blindly.
#include "HipBackend.h"

#include <memory>
#include <stdexcept>
#include <string>
#include <typeinfo>

namespace gpuSolver {
namespace {

void ok(int rc) {
  if (rc != GS_OK) throw std::runtime_error(gs_last_error());
}

template <class To, class From>
To &as(From &obj, const char *what) {
  auto *p = dynamic_cast<To *>(&obj);
  if (!p) throw std::runtime_error(std::string("HipBackend: ") + what + " was not created by HipBackend.");
  return *p;
}

gs_encoder_h E(BackendEncoder &e) { return as<HipEncoder>(e, "encoder").handle(); }
gs_dvec_h V(const DeviceVector &v) { return as<const HipDeviceVector>(v, "vector").handle(); }
gs_divec_h IV(const DeviceIndexVector &v) {
  return as<const HipDeviceIndexVector>(v, "index vector").handle();
}
gs_dscalar_h S(const DeviceScalar &s) { return as<const HipDeviceScalar>(s, "scalar").handle(); }
gs_dcsr_h M(const DeviceCSRMatrix &m) { return as<const HipDeviceCSRMatrix>(m, "matrix").handle(); }
gs_dell_h ELL(const DeviceELLMatrix &m) {
  return as<const HipDeviceELLMatrix>(m, "ELL matrix").handle();
}

}  // namespace

// ---- device objects ----------------------------------------------------------
HipEncoder::~HipEncoder() { gs_encoder_destroy(h_); }

void *HipEncoder::stream() const {
  void *s = nullptr;
  ok(gs_encoder_stream(h_, &s));
  return s;
}

HipDeviceVector::~HipDeviceVector() { gs_dvec_destroy(h_); }

void HipDeviceVector::updateValues(const float *values, std::size_t sizeOfValues) {
  ok(gs_dvec_update(h_, values, static_cast<int64_t>(sizeOfValues)));
}

float *HipDeviceVector::download() const {
  host_.resize(n_);
  if (n_) ok(gs_dvec_download(h_, host_.data()));
  return host_.data();
}

float *HipDeviceVector::data() const {
  void *p = nullptr;
  ok(gs_dvec_data(h_, &p));
  return static_cast<float *>(p);
}

HipDeviceIndexVector::~HipDeviceIndexVector() { gs_divec_destroy(h_); }

void HipDeviceIndexVector::updateValues(const int *values, const size_t sizeOfValues) {
  ok(gs_divec_update(h_, values, static_cast<int64_t>(sizeOfValues)));
}

int *HipDeviceIndexVector::download() const {
  host_.resize(n_);
  if (n_) ok(gs_divec_download(h_, host_.data()));
  return host_.data();
}

HipDeviceScalar::~HipDeviceScalar() { gs_dscalar_destroy(h_); }

float HipDeviceScalar::download() const {
  double v = 0.0;
  ok(gs_dscalar_download(h_, &v));
  return static_cast<float>(v);  // the FP32 value, widened exactly on the way
}

HipDeviceCSRMatrix::HipDeviceCSRMatrix(gs_dcsr_h h) : h_(h) {
  int64_t r = 0, c = 0, z = 0;
  ok(gs_dcsr_info(h, &r, &c, &z));
  rows_ = static_cast<std::size_t>(r);
  cols_ = static_cast<std::size_t>(c);
  nnz_ = static_cast<std::size_t>(z);
}

HipDeviceCSRMatrix::~HipDeviceCSRMatrix() { gs_dcsr_destroy(h_); }

void HipDeviceCSRMatrix::updateValues(const float *values, std::size_t numberOfValues) {
  if (!values) throw std::runtime_error("DeviceCSRMatrix::updateValues: values must not be null.");
  const std::vector<double> v(values, values + numberOfValues);  // exact widening
  ok(gs_dcsr_update_values(h_, v.data(), static_cast<int64_t>(numberOfValues)));
}

HipDeviceELLMatrix::HipDeviceELLMatrix(gs_dell_h h) : h_(h) {
  int64_t r = 0, c = 0, w = 0, o = 0;
  ok(gs_dell_info(h, &r, &c, &w, &o));
  rows_ = static_cast<std::size_t>(r);
  cols_ = static_cast<std::size_t>(c);
  width_ = static_cast<std::size_t>(w);
  ovf_ = static_cast<std::size_t>(o);
}

HipDeviceELLMatrix::~HipDeviceELLMatrix() { gs_dell_destroy(h_); }

// ---- backend -----------------------------------------------------------------
HipBackend::HipBackend() { ok(gs_backend_create(GS_DTYPE_F32, &h_)); }

HipBackend::~HipBackend() { gs_backend_destroy(h_); }

BackendEncoder *HipBackend::createEncoder() {
  gs_encoder_h e = nullptr;
  ok(gs_backend_create_encoder(h_, &e));
  return new HipEncoder(e);
}

void HipBackend::submit(BackendEncoder &encoder) { ok(gs_backend_submit(h_, E(encoder))); }

void HipBackend::submitAndWait(BackendEncoder &encoder) {
  ok(gs_backend_submit_and_wait(h_, E(encoder)));
}

DeviceVector *HipBackend::createVector(std::size_t size) {
  gs_dvec_h v = nullptr;
  ok(gs_backend_create_vector(h_, static_cast<int64_t>(size), nullptr, &v));
  return new HipDeviceVector(v, size);
}

DeviceVector *HipBackend::createVector(std::size_t size, const float *values) {
  if (values == nullptr && size != 0)  // src/CPU/CPUDeviceVector.cc:10-16
    throw std::runtime_error("HipDeviceVector: nullptr values with nonzero size.");
  gs_dvec_h v = nullptr;
  ok(gs_backend_create_vector(h_, static_cast<int64_t>(size), values, &v));
  return new HipDeviceVector(v, size);
}

DeviceIndexVector *HipBackend::createIndexVector(std::size_t size) {
  gs_divec_h v = nullptr;
  ok(gs_backend_create_index_vector(h_, static_cast<int64_t>(size), nullptr, &v));
  return new HipDeviceIndexVector(v, size);
}

DeviceIndexVector *HipBackend::createIndexVector(std::size_t size, const int *values) {
  if (values == nullptr && size != 0)
    throw std::runtime_error("HipDeviceIndexVector: nullptr values with nonzero size.");
  gs_divec_h v = nullptr;
  ok(gs_backend_create_index_vector(h_, static_cast<int64_t>(size), values, &v));
  return new HipDeviceIndexVector(v, size);
}

DeviceScalar *HipBackend::createScalar() { return createScalar(0.0f); }

DeviceScalar *HipBackend::createScalar(float value) {
  gs_dscalar_h s = nullptr;
  ok(gs_backend_create_scalar(h_, value, &s));
  return new HipDeviceScalar(s);
}

DeviceCSRMatrix *HipBackend::createCSRMatrix(std::size_t rows, std::size_t cols, std::size_t nnz,
                                             const int *rowPtr, const int *colIdx,
                                             const float *values) {
  if (nnz > 0 && !values) throw std::runtime_error("HipDeviceCSRMatrix: nullptr passed to constructor.");
  // The host matrix checks dimensions, pointers and indices (the CPU
  // backend's checks, src/CPU/CPUDeviceMatrix.cc:10-31, plus column
  // bounds); FP32 values are widened exactly and narrowed back on upload.
  const std::vector<double> v(values, values + nnz);
  gs_csr_h host = nullptr;
  ok(gs_csr_create(static_cast<int64_t>(rows), static_cast<int64_t>(cols),
                   static_cast<int64_t>(nnz), rowPtr, colIdx, v.data(), &host));
  std::unique_ptr<gs_csr_s, int (*)(gs_csr_h)> hold(host, gs_csr_destroy);
  gs_dcsr_h m = nullptr;
  ok(gs_backend_create_csr(h_, host, &m));
  return new HipDeviceCSRMatrix(m);
}

DeviceELLMatrix *HipBackend::createELLMatrix(std::size_t rows, std::size_t cols,
                                             std::size_t ellWidth, const int *ellColIdx,
                                             const float *ellValues, std::size_t overflowNnz,
                                             const int *overflowRowPtr,
                                             const int *overflowColIdx,
                                             const float *overflowValues) {
  gs_dell_h m = nullptr;
  ok(gs_backend_create_ell_packed(h_, static_cast<int64_t>(rows), static_cast<int64_t>(cols),
                                  static_cast<int64_t>(ellWidth), ellColIdx, ellValues,
                                  static_cast<int64_t>(overflowNnz), overflowRowPtr,
                                  overflowColIdx, overflowValues, &m));
  return new HipDeviceELLMatrix(m);
}

void HipBackend::encodeScale(BackendEncoder &encoder, DeviceVector &x, float alpha) {
  ok(gs_encode_scale(E(encoder), V(x), alpha));
}

void HipBackend::encodeScale(BackendEncoder &encoder, DeviceVector &x, const DeviceScalar &alpha) {
  ok(gs_encode_scale_dev(E(encoder), V(x), S(alpha)));
}

void HipBackend::encodeCopy(BackendEncoder &encoder, const DeviceVector &x, DeviceVector &xCopy) {
  ok(gs_encode_copy(E(encoder), V(x), V(xCopy)));
}

void HipBackend::encodeScaleVector(BackendEncoder &encoder, const DeviceVector &scaleVector,
                                   DeviceVector &output) {
  ok(gs_encode_scale_vector(E(encoder), V(scaleVector), V(output)));
}

void HipBackend::encodeSetZero(BackendEncoder &encoder, DeviceVector &input) {
  ok(gs_encode_set_zero(E(encoder), V(input)));
}

void HipBackend::encodeAxpy(BackendEncoder &encoder, float alpha, const DeviceVector &x,
                            DeviceVector &y) {
  ok(gs_encode_axpy(E(encoder), alpha, V(x), V(y)));
}

void HipBackend::encodeAxpy(BackendEncoder &encoder, const DeviceScalar &alpha,
                            const DeviceVector &x, DeviceVector &y) {
  ok(gs_encode_axpy_dev(E(encoder), S(alpha), V(x), V(y)));
}

void HipBackend::encodeSpmv(BackendEncoder &encoder, const DeviceCSRMatrix &A,
                            const DeviceVector &x, DeviceVector &y) {
  ok(gs_encode_spmv(E(encoder), M(A), V(x), V(y)));
}

void HipBackend::encodeSpmvELL(BackendEncoder &encoder, const DeviceELLMatrix &A,
                               const DeviceVector &x, DeviceVector &y) {
  ok(gs_encode_spmv_ell(E(encoder), ELL(A), V(x), V(y)));
}

void HipBackend::encodeGaussSeidelColor(BackendEncoder &encoder, const DeviceCSRMatrix &matrix,
                                        const DeviceIndexVector &colorVertices,
                                        std::size_t colorStart, std::size_t colorCount,
                                        const DeviceVector &rhs, DeviceVector &solution,
                                        float omega) {
  ok(gs_encode_gauss_seidel_color(E(encoder), M(matrix), IV(colorVertices),
                                  static_cast<int64_t>(colorStart),
                                  static_cast<int64_t>(colorCount), V(rhs), V(solution), omega));
}

void HipBackend::encodeDot(BackendEncoder &encoder, const DeviceVector &x, const DeviceVector &y,
                           DeviceScalar &result) {
  ok(gs_encode_dot(E(encoder), V(x), V(y), S(result)));
}

void HipBackend::encodeScalarDivide(BackendEncoder &encoder, const DeviceScalar &numerator,
                                    const DeviceScalar &denominator, DeviceScalar &result) {
  ok(gs_encode_scalar_divide(E(encoder), S(numerator), S(denominator), S(result)));
}

void HipBackend::encodeScalarMultiply(BackendEncoder &encoder, const DeviceScalar &factor1,
                                      const DeviceScalar &factor2, DeviceScalar &result) {
  ok(gs_encode_scalar_multiply(E(encoder), S(factor1), S(factor2), S(result)));
}

void HipBackend::encodeScalarSet(BackendEncoder &encoder, float value, DeviceScalar &output) {
  ok(gs_encode_scalar_set(E(encoder), value, S(output)));
}

void HipBackend::encodeScalarCopy(BackendEncoder &encoder, const DeviceScalar &input,
                                  DeviceScalar &output) {
  ok(gs_encode_scalar_copy(E(encoder), S(input), S(output)));
}

void HipBackend::encodeScalarNegate(BackendEncoder &encoder, const DeviceScalar &input,
                                    DeviceScalar &output) {
  ok(gs_encode_scalar_negate(E(encoder), S(input), S(output)));
}

void HipBackend::encodeScalarSqrt(BackendEncoder &encoder, const DeviceScalar &input,
                                  DeviceScalar &output) {
  ok(gs_encode_scalar_sqrt(E(encoder), S(input), S(output)));
}

void HipBackend::encodeApplyBlockInverses3x3(BackendEncoder &encoder,
                                             const DeviceVector &inverseDiagBlocks,
                                             const DeviceVector &residual,
                                             DeviceVector &correction) {
  ok(gs_encode_apply_block_inverses3x3(E(encoder), V(inverseDiagBlocks), V(residual),
                                       V(correction)));
}

void HipBackend::encodeBlockGaussSeidelColor3x3(BackendEncoder &encoder,
                                                const DeviceCSRMatrix &matrix,
                                                const DeviceIndexVector &colorBlocks,
                                                const DeviceVector &diagonalInverses,
                                                std::size_t begin, std::size_t end,
                                                const DeviceVector &rhs, DeviceVector &solution,
                                                float omega) {
  ok(gs_encode_block_gauss_seidel_color3x3(E(encoder), M(matrix), IV(colorBlocks),
                                           V(diagonalInverses), static_cast<int64_t>(begin),
                                           static_cast<int64_t>(end), V(rhs), V(solution), omega));
}

// ---- standalone API: own encoder, submit, wait (Backend.h:116-157) ----------
namespace {
template <class F>
void blocking(HipBackend &b, F &&f) {
  std::unique_ptr<BackendEncoder> e(b.createEncoder());
  f(*e);
  b.submitAndWait(*e);
}
}  // namespace

void HipBackend::scale(DeviceVector &x, float alpha) {
  blocking(*this, [&](BackendEncoder &e) { encodeScale(e, x, alpha); });
}
void HipBackend::scale(DeviceVector &x, const DeviceScalar &alpha) {
  blocking(*this, [&](BackendEncoder &e) { encodeScale(e, x, alpha); });
}
void HipBackend::scaleVectorByVector(const DeviceVector &scaleVector, DeviceVector &output) {
  blocking(*this, [&](BackendEncoder &e) { encodeScaleVector(e, scaleVector, output); });
}
void HipBackend::axpy(float alpha, const DeviceVector &x, DeviceVector &y) {
  blocking(*this, [&](BackendEncoder &e) { encodeAxpy(e, alpha, x, y); });
}
void HipBackend::axpy(const DeviceScalar &alpha, const DeviceVector &x, DeviceVector &y) {
  blocking(*this, [&](BackendEncoder &e) { encodeAxpy(e, alpha, x, y); });
}
void HipBackend::spmv(const DeviceCSRMatrix &A, const DeviceVector &x, DeviceVector &y) {
  blocking(*this, [&](BackendEncoder &e) { encodeSpmv(e, A, x, y); });
}
void HipBackend::dot(const DeviceVector &x, const DeviceVector &y, DeviceScalar &result) {
  blocking(*this, [&](BackendEncoder &e) { encodeDot(e, x, y, result); });
}
void HipBackend::scalarSet(float value, DeviceScalar &result) {
  blocking(*this, [&](BackendEncoder &e) { encodeScalarSet(e, value, result); });
}
void HipBackend::scalarCopy(const DeviceScalar &input, DeviceScalar &output) {
  blocking(*this, [&](BackendEncoder &e) { encodeScalarCopy(e, input, output); });
}
void HipBackend::scalarDivide(const DeviceScalar &numerator, const DeviceScalar &denominator,
                              DeviceScalar &result) {
  blocking(*this, [&](BackendEncoder &e) { encodeScalarDivide(e, numerator, denominator, result); });
}
void HipBackend::scalarMultiply(const DeviceScalar &factor1, const DeviceScalar &factor2,
                                DeviceScalar &result) {
  blocking(*this, [&](BackendEncoder &e) { encodeScalarMultiply(e, factor1, factor2, result); });
}
void HipBackend::scalarNegate(const DeviceScalar &input, DeviceScalar &output) {
  blocking(*this, [&](BackendEncoder &e) { encodeScalarNegate(e, input, output); });
}
void HipBackend::scalarSqrt(const DeviceScalar &input, DeviceScalar &output) {
  blocking(*this, [&](BackendEncoder &e) { encodeScalarSqrt(e, input, output); });
}

}  // namespace gpuSolver
