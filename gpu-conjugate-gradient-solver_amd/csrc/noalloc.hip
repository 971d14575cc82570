// The allocation-free variant of the Backend plugin's dot product
// (gs_encode_dot, Backend::encodeDot): for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops). It runs
// the kernels and grid of gs_encode_dot on caller-owned device arrays and
// takes its partial sums from a file-scope device array instead of an
// encoder's buffer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/gpusolver.h"
#include "device.hpp"

namespace gsc {
void set_error(const char *msg);
}

// Partial sums: one per workgroup of the dot's grid, which is capped at
// gsk::kDotMaxGrid (2048) workgroups, so 16 KiB serve every size. Each GPU
// holds its own copy until the process ends. Host code never uses this name
// as an address: it takes the device address hipGetSymbolAddress returns
// for the current device (scratch_address).
static __device__ __align__(256) double scratch[gsk::kDotMaxGrid];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device, with that device current, and kept.
void *scratch_address() {
  int dev = 0;
  GS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    throw std::runtime_error("gs_dot_noalloc: device index out of range.");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_addr[dev]) {
    void *p = nullptr;
    GS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("gs_dot_noalloc: no device address for the scratch array.");
    g_addr[dev] = p;
  }
  return g_addr[dev];
}

template <class T>
void launch(const void *x, const void *y, int64_t n, void *out, void *parts, uint32_t G,
            hipStream_t st) {
  gsk::k_dot<T><<<G, gsk::kBlock, 0, st>>>(static_cast<const T *>(x), static_cast<const T *>(y),
                                           n, static_cast<T *>(parts));
  gsk::k_dot_final<T><<<1, gsk::kBlock, 0, st>>>(static_cast<const T *>(parts), G,
                                                 static_cast<T *>(out));
}

}  // namespace

extern "C" int gs_dot_noalloc(int32_t dtype, const void *x, const void *y, int64_t n, void *out,
                              void *stream) {
  try {
    if (dtype != GS_DTYPE_F32 && dtype != GS_DTYPE_F64)
      throw std::runtime_error("gs_dot_noalloc: unknown value type.");
    if (n < 0) throw std::runtime_error("gs_dot_noalloc: negative size.");
    if (!out) throw std::runtime_error("gs_dot_noalloc: output must not be null.");
    if (n > 0 && (!x || !y)) throw std::runtime_error("gs_dot_noalloc: vectors must not be null.");
    const uint32_t G = gsk::dot_grid(n);
    if (G > gsk::kDotMaxGrid)
      throw std::runtime_error("gs_dot_noalloc: more partial sums than the scratch array holds.");
    void *parts = scratch_address();
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dtype == GS_DTYPE_F64) launch<double>(x, y, n, out, parts, G, st);
    else launch<float>(x, y, n, out, parts, G, st);
    GS_HIP(hipGetLastError());
    return GS_OK;
  } catch (const std::exception &e) {
    gsc::set_error(e.what());
    return GS_ERR;
  }
}
