// gpuSolver::HipBackend: the reference's Backend plugin interface
// (include/GPUSolver/Backend.h:15-159) implemented over this library's C ABI
// (include/gpusolver.h, gs_backend_* / gs_encode_*), so the reference's own
// CGSolver, preconditioners and smoothers run on the HIP kernels unchanged:
//
//   gpuSolver::HipBackend backend;
//   gpuSolver::MultigridGaussSeidelPreconditioner M(...);
//   gpuSolver::CGSolver cg(backend, M, A);   // in place of CPUBackend / CudaBackend
//
// Build: compile HipBackend.cc (needs only the reference's Eigen-free
// interface headers: Backend.h, BackendEncoder.h, Device*.h) and
// HipBackendHost.cc (createCSRMatrix / createELLMatrix from the host
// matrices, whose headers include Eigen) with -I <reference>/include, and
// link libgpusolver.so. Values are FP32 as in the reference.
//
// Semantics follow the reference's CPU and CUDA backends: size checks and
// their error texts, (begin, end) for the block Gauss-Seidel colour range
// (src/cuda/CudaBackend.cc:702-780), download() returning a host buffer
// owned by the object (src/cuda/CudaDeviceVector.cc:90-100), submit()
// queueing and submitAndWait() blocking. Errors from the library are thrown
// as std::runtime_error with the library's message.
#pragma once

#include <GPUSolver/Backend.h>
#include <GPUSolver/BackendEncoder.h>
#include <GPUSolver/DeviceELLMatrix.h>
#include <GPUSolver/DeviceIndexVector.h>
#include <GPUSolver/DeviceScalar.h>
#include <GPUSolver/DeviceSparseMatrix.h>
#include <GPUSolver/DeviceVector.h>

#include <cstddef>
#include <vector>

#include "../../include/gpusolver.h"

namespace gpuSolver {

// One HIP stream (CudaEncoder's one stream, src/cuda/CudaEncoder.cc:7-17).
class HipEncoder final : public BackendEncoder {
 public:
  explicit HipEncoder(gs_encoder_h h) : h_(h) {}
  ~HipEncoder() override;
  gs_encoder_h handle() const { return h_; }
  // The encoder's hipStream_t, for work enqueued in order with it.
  void *stream() const;

 private:
  gs_encoder_h h_;
};

class HipDeviceVector final : public DeviceVector {
 public:
  explicit HipDeviceVector(gs_dvec_h h, std::size_t n) : h_(h), n_(n) {}
  ~HipDeviceVector() override;
  std::size_t size() const override { return n_; }
  void updateValues(const float *values, std::size_t sizeOfValues) override;
  // Synchronous copy into a buffer the vector owns (valid until the next
  // download or the vector's destruction).
  float *download() const override;
  gs_dvec_h handle() const { return h_; }
  float *data() const;  // device address

 private:
  gs_dvec_h h_;
  std::size_t n_;
  mutable std::vector<float> host_;
};

class HipDeviceIndexVector final : public DeviceIndexVector {
 public:
  explicit HipDeviceIndexVector(gs_divec_h h, std::size_t n) : h_(h), n_(n) {}
  ~HipDeviceIndexVector() override;
  std::size_t size() const override { return n_; }
  std::size_t size() override { return n_; }
  void updateValues(const int *values, const size_t sizeOfValues) override;
  int *download() const override;
  gs_divec_h handle() const { return h_; }

 private:
  gs_divec_h h_;
  std::size_t n_;
  mutable std::vector<int> host_;
};

class HipDeviceScalar final : public DeviceScalar {
 public:
  explicit HipDeviceScalar(gs_dscalar_h h) : h_(h) {}
  ~HipDeviceScalar() override;
  float download() const override;
  gs_dscalar_h handle() const { return h_; }

 private:
  gs_dscalar_h h_;
};

class HipDeviceCSRMatrix final : public DeviceCSRMatrix {
 public:
  explicit HipDeviceCSRMatrix(gs_dcsr_h h);
  ~HipDeviceCSRMatrix() override;
  std::size_t rows() const override { return rows_; }
  std::size_t cols() const override { return cols_; }
  std::size_t nnz() const override { return nnz_; }
  void updateValues(const float *values, std::size_t numberOfValues) override;
  gs_dcsr_h handle() const { return h_; }

 private:
  gs_dcsr_h h_;
  std::size_t rows_ = 0, cols_ = 0, nnz_ = 0;
};

class HipDeviceELLMatrix final : public DeviceELLMatrix {
 public:
  explicit HipDeviceELLMatrix(gs_dell_h h);
  ~HipDeviceELLMatrix() override;
  std::size_t rows() const override { return rows_; }
  std::size_t cols() const override { return cols_; }
  std::size_t ellWidth() const override { return width_; }
  std::size_t overflowNnz() const override { return ovf_; }
  gs_dell_h handle() const { return h_; }

 private:
  gs_dell_h h_;
  std::size_t rows_ = 0, cols_ = 0, width_ = 0, ovf_ = 0;
};

class HipBackend final : public Backend {
 public:
  HipBackend();
  ~HipBackend() override;
  HipBackend(const HipBackend &) = delete;
  HipBackend &operator=(const HipBackend &) = delete;

  BackendEncoder *createEncoder() override;
  void submit(BackendEncoder &encoder) override;
  void submitAndWait(BackendEncoder &encoder) override;

  DeviceVector *createVector(std::size_t size) override;
  DeviceVector *createVector(std::size_t size, const float *values) override;
  DeviceIndexVector *createIndexVector(std::size_t size) override;
  DeviceIndexVector *createIndexVector(std::size_t size, const int *values) override;
  DeviceScalar *createScalar() override;
  DeviceScalar *createScalar(float value) override;
  // From the host matrix's active (permuted, if a permutation was applied)
  // arrays, as src/cuda/CudaBackend.cc:175-187; in HipBackendHost.cc.
  DeviceCSRMatrix *createCSRMatrix(const HostCSRMatrix &matrix) override;
  DeviceELLMatrix *createELLMatrix(const HostELLMatrix &matrix) override;
  // The same from raw arrays (what the two overloads above forward).
  DeviceCSRMatrix *createCSRMatrix(std::size_t rows, std::size_t cols, std::size_t nnz,
                                   const int *rowPtr, const int *colIdx, const float *values);
  DeviceELLMatrix *createELLMatrix(std::size_t rows, std::size_t cols, std::size_t ellWidth,
                                   const int *ellColIdx, const float *ellValues,
                                   std::size_t overflowNnz, const int *overflowRowPtr,
                                   const int *overflowColIdx, const float *overflowValues);

  void encodeScale(BackendEncoder &encoder, DeviceVector &x, float alpha) override;
  void encodeScale(BackendEncoder &encoder, DeviceVector &x, const DeviceScalar &alpha) override;
  void encodeCopy(BackendEncoder &encoder, const DeviceVector &x, DeviceVector &xCopy) override;
  void encodeScaleVector(BackendEncoder &encoder, const DeviceVector &scaleVector,
                         DeviceVector &output) override;
  void encodeSetZero(BackendEncoder &encoder, DeviceVector &input) override;
  void encodeAxpy(BackendEncoder &encoder, float alpha, const DeviceVector &x,
                  DeviceVector &y) override;
  void encodeAxpy(BackendEncoder &encoder, const DeviceScalar &alpha, const DeviceVector &x,
                  DeviceVector &y) override;
  void encodeSpmv(BackendEncoder &encoder, const DeviceCSRMatrix &A, const DeviceVector &x,
                  DeviceVector &y) override;
  // ELL SpMV (the Metal backend's encodeSpmvELLMetal,
  // include/GPUSolver/metal/MetalBackend.h:229-236); not in Backend.
  void encodeSpmvELL(BackendEncoder &encoder, const DeviceELLMatrix &A, const DeviceVector &x,
                     DeviceVector &y);
  void encodeGaussSeidelColor(BackendEncoder &encoder, const DeviceCSRMatrix &matrix,
                              const DeviceIndexVector &colorVertices, std::size_t colorStart,
                              std::size_t colorCount, const DeviceVector &rhs,
                              DeviceVector &solution, float omega) override;
  void encodeDot(BackendEncoder &encoder, const DeviceVector &x, const DeviceVector &y,
                 DeviceScalar &result) override;
  void encodeScalarDivide(BackendEncoder &encoder, const DeviceScalar &numerator,
                          const DeviceScalar &denominator, DeviceScalar &result) override;
  void encodeScalarMultiply(BackendEncoder &encoder, const DeviceScalar &factor1,
                            const DeviceScalar &factor2, DeviceScalar &result) override;
  void encodeScalarSet(BackendEncoder &encoder, float value, DeviceScalar &output) override;
  void encodeScalarCopy(BackendEncoder &encoder, const DeviceScalar &input,
                        DeviceScalar &output) override;
  void encodeScalarNegate(BackendEncoder &encoder, const DeviceScalar &input,
                          DeviceScalar &output) override;
  void encodeScalarSqrt(BackendEncoder &encoder, const DeviceScalar &input,
                        DeviceScalar &output) override;
  void encodeApplyBlockInverses3x3(BackendEncoder &encoder, const DeviceVector &inverseDiagBlocks,
                                   const DeviceVector &residual,
                                   DeviceVector &correction) override;
  void encodeBlockGaussSeidelColor3x3(BackendEncoder &encoder, const DeviceCSRMatrix &matrix,
                                      const DeviceIndexVector &colorBlocks,
                                      const DeviceVector &diagonalInverses, std::size_t begin,
                                      std::size_t end, const DeviceVector &rhs,
                                      DeviceVector &solution, float omega) override;

  void scale(DeviceVector &x, float alpha) override;
  void scale(DeviceVector &x, const DeviceScalar &alpha) override;
  void scaleVectorByVector(const DeviceVector &scaleVector, DeviceVector &output) override;
  void axpy(float alpha, const DeviceVector &x, DeviceVector &y) override;
  void axpy(const DeviceScalar &alpha, const DeviceVector &x, DeviceVector &y) override;
  void spmv(const DeviceCSRMatrix &A, const DeviceVector &x, DeviceVector &y) override;
  void dot(const DeviceVector &x, const DeviceVector &y, DeviceScalar &result) override;
  void scalarSet(float value, DeviceScalar &result) override;
  void scalarCopy(const DeviceScalar &input, DeviceScalar &output) override;
  void scalarDivide(const DeviceScalar &numerator, const DeviceScalar &denominator,
                    DeviceScalar &result) override;
  void scalarMultiply(const DeviceScalar &factor1, const DeviceScalar &factor2,
                      DeviceScalar &result) override;
  void scalarNegate(const DeviceScalar &input, DeviceScalar &output) override;
  void scalarSqrt(const DeviceScalar &input, DeviceScalar &output) override;

  gs_backend_h handle() const { return h_; }

 private:
  gs_backend_h h_ = nullptr;
};

}  // namespace gpuSolver
