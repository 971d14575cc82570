// GPU side of the smoothed-aggregation setup (setup_device.hip): the sparse
// products of a coarsening step, bit for bit the host's.
#pragma once

#include "device.hpp"
#include "host/setup.h"

namespace gsdev {

// CSR matrix in device memory (gsx::Csr layout: int32 row pointers and
// columns, FP64 values).
struct DCsr {
  int64_t rows = 0, cols = 0, nnz = 0;
  gsd::DevBuf ptr, col, val;
  const int32_t *p() const { return ptr.as<int32_t>(); }
  const int32_t *c() const { return col.as<int32_t>(); }
  const double *v() const { return val.as<double>(); }
};
DCsr upload(const gsx::Csr &A);

// P = M P_tent, R = P^T, A_c = R (A P) on the GPU (the matrices are
// uploaded once per coarsening step and the results downloaded).
struct DeviceGalerkinOps : gsx::GalerkinOps {
  void run(const gsx::Csr &A, const gsx::Csr &M, const gsx::Csr &Pt, gsx::Csr &P, gsx::Csr &R,
           gsx::Csr &Ac) const override;
};

// Block CSR in device memory (gsx::Bsr3Csr layout: int64 block-row
// pointers, int32 block columns, nine FP64 values per block row-major);
// host_ptr is a host copy of the pointers.
struct DBsr3Csr {
  int64_t nb = 0, ncb = 0, nblocks = 0;
  int32_t width = 0;
  gsd::DevBuf ptr, bcol, val;
  std::vector<int64_t> host_ptr;
};

// Row i of the result is row rq[i] of A with column c renamed cmap[c]
// (gsx::relabel).
DCsr relabel(const DCsr &A, const std::vector<int32_t> &rq, const std::vector<int32_t> &cmap);
// gsx::bsr3_pack on the device (key: host array of a rank per block column,
// or null); false unless A's dimensions are multiples of 3.
bool bsr3_pack(const DCsr &A, const int32_t *key, DBsr3Csr &out);
// The blocks of each block row b whose block column is below start[b], in
// their order (the earlier-colour part of a colour-ordered operator).
DBsr3Csr bsr3_below(const DBsr3Csr &B, const std::vector<int32_t> &start);
// BSR3-SELL-64 storage (gsx::bsr3_sell64_pack of the block rows taken in
// `order`, or in their own order when null; value pairs when out.pairs is
// set, as gsx::bsr3_pair_planes), values rounded to T.
template <class T>
void bsr3_sell_pack(const DBsr3Csr &B, const int32_t *order, gsd::DBsr3<T> &out);
// Row-contiguous block storage for the lane-group kernels.
template <class T>
void bsr3_rows(const DBsr3Csr &B, gsd::DBRows<T> &out);

// C = A B and A^T on the GPU, equal to gsx::spgemm / gsx::transpose.
gsx::Csr device_spgemm(const gsx::Csr &A, const gsx::Csr &B);
gsx::Csr device_transpose(const gsx::Csr &A);

}  // namespace gsdev
