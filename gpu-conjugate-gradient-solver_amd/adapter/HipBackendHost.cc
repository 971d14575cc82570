// The two HipBackend methods that take the reference's host matrices
// (include/GPUSolver/HostSparseMatrix.h, HostELLMatrix.h; the first
// includes Eigen), forwarding their arrays to the raw-array overloads.
#include <GPUSolver/HostELLMatrix.h>
#include <GPUSolver/HostSparseMatrix.h>

#include "HipBackend.h"

namespace gpuSolver {

// The active arrays: permuted, if a permutation was applied
// (src/cuda/CudaBackend.cc:175-187).
DeviceCSRMatrix *HipBackend::createCSRMatrix(const HostCSRMatrix &matrix) {
  return createCSRMatrix(matrix.rows(), matrix.cols(), matrix.nnz(), matrix.activeRowPtr(),
                         matrix.activeColPtr(), matrix.activeValPtr());
}

DeviceELLMatrix *HipBackend::createELLMatrix(const HostELLMatrix &matrix) {
  return createELLMatrix(matrix.rows(), matrix.cols(), matrix.ellWidth(), matrix.ellColIdx(),
                         matrix.ellValues(), matrix.overflowNnz(), matrix.overflowRowPtr(),
                         matrix.overflowColIdx(), matrix.overflowValues());
}

}  // namespace gpuSolver
