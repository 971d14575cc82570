// Device-built dense FP64 SPD inverse (coarse-level direct solve).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace gsdense {

// Row stride of the device inverse: n rounded up to the 64-row pivot block.
int64_t padded_size(int64_t n);

// Writes A^-1 (row-major, row stride padded_size(n), padding rows/columns
// identity) to d_inv, which must hold padded_size(n)^2 doubles. `a` is the
// dense row-major n x n SPD matrix. Throws when a pivot is not positive.
void spd_inverse_device(int64_t n, const std::vector<double> &a, double *d_inv,
                        hipStream_t stream);
// The same from the n x n matrix in CSR form (host arrays; duplicates are
// summed in stored order): the padded matrix is assembled on the device.
void spd_inverse_device_csr(int64_t n, const int32_t *ptr, const int32_t *col,
                            const double *val, double *d_inv, hipStream_t stream);

// Doubles held by the lower 64 x 64 tiles of the padded inverse.
int64_t lower_tiles_size(int64_t n);
// Copies the lower tiles (I >= J) of the padded row-major inverse d_full
// (row stride padded_size(n)) to d_tiles: tile (I, J) at
// (I (I + 1) / 2 + J) * 4096, row-major.
void pack_lower_tiles(int64_t n, const double *d_full, double *d_tiles, hipStream_t stream);

}  // namespace gsdense
