// HipBackend built without the reference's host-matrix headers (which need
// Eigen): the two overloads taking HostCSRMatrix / HostELLMatrix report that
// they are not available; the raw-array overloads do the same work. Link
// HipBackendHost.cc instead where Eigen is present.
#include <stdexcept>

#include "HipBackend.h"

namespace gpuSolver {

DeviceCSRMatrix *HipBackend::createCSRMatrix(const HostCSRMatrix &) {
  throw std::logic_error(
      "HipBackend::createCSRMatrix(const HostCSRMatrix&): built without HipBackendHost.cc "
      "(Eigen); use the raw-array overload.");
}

DeviceELLMatrix *HipBackend::createELLMatrix(const HostELLMatrix &) {
  throw std::logic_error(
      "HipBackend::createELLMatrix(const HostELLMatrix&): built without HipBackendHost.cc "
      "(Eigen); use the raw-array overload.");
}

}  // namespace gpuSolver
