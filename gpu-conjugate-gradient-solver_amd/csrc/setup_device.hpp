// GPU side of the smoothed-aggregation setup (setup_device.hip): the sparse
// products of a coarsening step, bit for bit the host's.
#pragma once

#include "host/setup.h"

namespace gsdev {

// P = M P_tent, R = P^T, A_c = R (A P) on the GPU (the matrices are
// uploaded once per coarsening step and the results downloaded).
struct DeviceGalerkinOps : gsx::GalerkinOps {
  void run(const gsx::Csr &A, const gsx::Csr &M, const gsx::Csr &Pt, gsx::Csr &P, gsx::Csr &R,
           gsx::Csr &Ac) const override;
};

// C = A B and A^T on the GPU, equal to gsx::spgemm / gsx::transpose.
gsx::Csr device_spgemm(const gsx::Csr &A, const gsx::Csr &B);
gsx::Csr device_transpose(const gsx::Csr &A);

}  // namespace gsdev
