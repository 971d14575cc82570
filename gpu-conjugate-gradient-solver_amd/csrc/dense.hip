// Dense FP64 SPD inverse on the device: the coarse-level direct solve.
//
// The coarse level is solved as x = A_c^-1 rhs with an explicit FP64
// inverse (one bandwidth-bound product per V-cycle over the inverse's lower
// 64 x 64 tiles, see k_dense_sym_strip). Building that inverse on the host
// costs O(n^3) serial
// flops, minutes at n ~ 10^4; here it is blocked Gauss-Jordan elimination in
// place on the device, 64 x 64 pivot blocks:
//
//   for each pivot block K:
//     D^-1    = inverse of A[K,K]                    (one workgroup, LDS)
//     C       = A[:,K]                               (saved column panel)
//     R[:,j]  = D^-1 A[K,j]   (j not in K),  R[:,K] = D^-1
//     A[K,:]  = R
//     A[i,:]  = A[i,:] - C[i,:] R   (i not in K; A[i,K] taken as 0)
//
// which leaves A^-1 in place. The rank-64 updates, 2 n^3 of the flops, run
// on the FP64 matrix cores. Without pivoting this is stable for SPD
// matrices: every pivot block is a Schur complement of an SPD matrix, hence
// SPD; a non-positive pivot stops the elimination with an error. The matrix
// is padded to a multiple of 64 with an identity block (still SPD; the
// padding does not couple to the real rows). 2 n^3 flops, 2 n^2 * 8 B of
// traffic per pivot block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "dense.hpp"

namespace gsdense {

namespace {

constexpr int NB = 64;      // pivot block
constexpr int kThreads = 256;

#define GS_DHIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

// Inverse of the 64 x 64 block A[K,K] by Gauss-Jordan in LDS. flag[0] is set
// to 1 when a pivot is not positive (matrix not SPD).
__global__ __launch_bounds__(kThreads) void k_pivot_inverse(const double *__restrict__ A, int64_t N,
                                                            int64_t k0, double *__restrict__ Dinv,
                                                            int *flag) {
  __shared__ double a[NB][NB + 1];
  __shared__ double prow[NB];
  for (int e = threadIdx.x; e < NB * NB; e += kThreads) {
    const int i = e / NB, j = e % NB;
    a[i][j] = A[(k0 + i) * N + k0 + j];
  }
  __syncthreads();
  for (int p = 0; p < NB; ++p) {
    const double piv = a[p][p];
    if (!(piv > 0.0)) {
      if (threadIdx.x == 0) flag[0] = 1;
      return;  // uniform: every thread read the same pivot
    }
    const double inv = 1.0 / piv;
    __syncthreads();
    // Row p scaled; its pivot entry becomes 1 / piv.
    if (threadIdx.x < NB) {
      const int j = threadIdx.x;
      const double v = (j == p) ? inv : a[p][j] * inv;
      prow[j] = v;
    }
    __syncthreads();
    if (threadIdx.x < NB) a[p][threadIdx.x] = prow[threadIdx.x];
    // Other rows: a[i][j] -= a[i][p] * prow[j]; a[i][p] = -a[i][p] * inv.
    for (int e = threadIdx.x; e < NB * NB; e += kThreads) {
      const int i = e / NB, j = e % NB;
      if (i == p) continue;
      const double f = a[i][p];
      if (j == p) continue;
      a[i][j] = fma(-f, prow[j], a[i][j]);
    }
    __syncthreads();
    if (threadIdx.x < NB && threadIdx.x != p) {
      const int i = threadIdx.x;
      a[i][p] = -a[i][p] * inv;
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < NB * NB; e += kThreads) {
    const int i = e / NB, j = e % NB;
    Dinv[i * NB + j] = a[i][j];
  }
}

// C[i][c] = A[i][k0 + c] for all rows i (the column panel before update).
__global__ __launch_bounds__(kThreads) void k_save_column(const double *__restrict__ A, int64_t N,
                                                          int64_t k0, double *__restrict__ C) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (e >= N * NB) return;
  const int64_t i = e / NB, c = e % NB;
  C[e] = A[i * N + k0 + c];
}

// R[:, tile jt] = D^-1 A[K, jt] (jt != K) or D^-1 (jt == K); R is NB x N.
__global__ __launch_bounds__(kThreads) void k_row_panel(const double *__restrict__ A, int64_t N,
                                                        int64_t k0, const double *__restrict__ Dinv,
                                                        double *__restrict__ R) {
  __shared__ double d[NB][NB + 1];
  __shared__ double b[NB][NB + 1];
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * NB;
  const bool diag = j0 == k0;
  for (int e = threadIdx.x; e < NB * NB; e += kThreads) {
    const int i = e / NB, j = e % NB;
    d[i][j] = Dinv[i * NB + j];
    b[i][j] = diag ? 0.0 : A[(k0 + i) * N + j0 + j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NB * NB; e += kThreads) {
    const int i = e / NB, j = e % NB;
    double s;
    if (diag) {
      s = d[i][j];
    } else {
      s = 0.0;
      for (int q = 0; q < NB; ++q) s = fma(d[i][q], b[q][j], s);
    }
    R[static_cast<int64_t>(i) * N + j0 + j] = s;
  }
}

// A[it, jt] = (it == K) ? R[:, jt] : A[it, jt]|_{A[i,K] = 0} - C[it, :] R[:, jt].
// One 64 x 64 output tile per workgroup; each wavefront forms a 32 x 32
// quarter of C[it, :] R[:, jt] on the FP64 matrix cores: 2 x 2 tiles of
// v_mfma_f64_16x16x4f64, sixteen steps of 4 over the 64-wide panel. Operand
// and result lanes of that instruction (lane l, result register v):
// A[l % 16][l / 16], B[l / 16][l % 16], D[4 v + l / 16][l % 16].
// The operands are read straight from the panels C and R (cache-resident,
// shared by a row / column of tiles), without LDS staging; the update is
// bound by the tile's own read and write (measured per update at n = 13 047:
// FP64 FMAs from LDS 1.44 ms, MFMA from LDS 1.51 ms, with the tile loaded
// first 1.23 ms, operands from the panels 0.975 ms; a fully unrolled or
// 8-step unrolled panel loop 1.18 / 1.15 ms; each lane's A operands as one
// contiguous 16-byte-vector run, with the panel columns reordered, 1.24 ms).
typedef double dbl4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kThreads) void k_rank_update(double *__restrict__ A, int64_t N,
                                                          int64_t k0, const double *__restrict__ C,
                                                          const double *__restrict__ R) {
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * NB;
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * NB;
  if (i0 == k0) {
    for (int e = threadIdx.x; e < NB * NB; e += kThreads) {
      const int i = e / NB, j = e % NB;
      A[(i0 + i) * N + j0 + j] = R[static_cast<int64_t>(i) * N + j0 + j];
    }
    return;
  }
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int rb = (w / 2) * 32, cb = (w % 2) * 32;
  const int lr = l % 16, lq = l / 16;
  // The tile's old values are loaded first: their latency overlaps the
  // panel loads and the products.
  const bool kcol = j0 == k0;
  double old[2][2][4];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + rb + 16 * ti + 4 * v + lq;
        old[ti][tj][v] = kcol ? 0.0 : A[i * N + j0 + cb + 16 * tj + lr];
      }
  dbl4 acc[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int q0 = 0; q0 < NB; q0 += 4) {
    double a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      a[t] = C[(i0 + rb + 16 * t + lr) * NB + q0 + lq];
      b[t] = R[static_cast<int64_t>(q0 + lq) * N + j0 + cb + 16 * t + lr];
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + rb + 16 * ti + 4 * v + lq;
        A[i * N + j0 + cb + 16 * tj + lr] = old[ti][tj][v] - acc[ti][tj][v];
      }
}

// A = 0 with ones on the padding diagonal (rows and columns n..N-1).
__global__ __launch_bounds__(kThreads) void k_fill_padded(double *__restrict__ A, int64_t N,
                                                          int64_t n) {
  const int64_t total = N * N;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int64_t i = e / N, j = e % N;
    A[e] = (i == j && i >= n) ? 1.0 : 0.0;
  }
}

// A[r][col[k]] += val[k] over row r's entries in CSR order (one thread per
// row, so duplicates are summed in stored order).
__global__ __launch_bounds__(kThreads) void k_scatter_rows(const int32_t *__restrict__ ptr,
                                                           const int32_t *__restrict__ col,
                                                           const double *__restrict__ val,
                                                           int64_t n, double *__restrict__ A,
                                                           int64_t N) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= n) return;
  for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) A[r * N + col[k]] += val[k];
}

// Blocked Gauss-Jordan on the padded N x N matrix in d_inv (see above).
void eliminate(int64_t N, double *d_inv, hipStream_t stream) {
  double *dinv = nullptr, *C = nullptr, *R = nullptr;
  int *flag = nullptr;
  GS_DHIP(hipMalloc(&dinv, NB * NB * sizeof(double)));
  GS_DHIP(hipMalloc(&C, static_cast<size_t>(N) * NB * sizeof(double)));
  GS_DHIP(hipMalloc(&R, static_cast<size_t>(N) * NB * sizeof(double)));
  GS_DHIP(hipMalloc(&flag, sizeof(int)));
  struct Free {
    void *p[4];
    ~Free() {
      for (void *q : p)
        if (q) (void)hipFree(q);
    }
  } guard{{dinv, C, R, flag}};
  GS_DHIP(hipMemsetAsync(flag, 0, sizeof(int), stream));
  const uint32_t T = static_cast<uint32_t>(N / NB);
  for (int64_t k0 = 0; k0 < N; k0 += NB) {
    k_pivot_inverse<<<1, kThreads, 0, stream>>>(d_inv, N, k0, dinv, flag);
    k_save_column<<<static_cast<uint32_t>((N * NB + kThreads - 1) / kThreads), kThreads, 0,
                    stream>>>(d_inv, N, k0, C);
    k_row_panel<<<T, kThreads, 0, stream>>>(d_inv, N, k0, dinv, R);
    k_rank_update<<<dim3(T, T), kThreads, 0, stream>>>(d_inv, N, k0, C, R);
  }
  int hflag = 0;
  GS_DHIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, stream));
  GS_DHIP(hipStreamSynchronize(stream));
  if (hflag) throw std::runtime_error("coarse level matrix is not SPD.");
}

// tiles[(I (I + 1) / 2 + J) 4096 + 64 r + c] = A[64 I + r][64 J + c], I >= J;
// one workgroup per tile.
__global__ __launch_bounds__(kThreads) void k_pack_lower_tiles(const double *__restrict__ A,
                                                               int64_t N,
                                                               double *__restrict__ tiles) {
  const int64_t t = blockIdx.x;
  // Tile (I, J) of linear index t: the largest I with I (I + 1) / 2 <= t.
  int64_t I = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(t) + 1.0) - 1.0) / 2.0);
  while (I * (I + 1) / 2 > t) --I;
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  const int64_t J = t - I * (I + 1) / 2;
  for (int e = threadIdx.x; e < NB * NB; e += kThreads) {
    const int r = e / NB, c = e % NB;
    tiles[t * NB * NB + e] = A[(I * NB + r) * N + J * NB + c];
  }
}

}  // namespace

int64_t padded_size(int64_t n) { return (n + NB - 1) / NB * NB; }

int64_t lower_tiles_size(int64_t n) {
  const int64_t nt = padded_size(n) / NB;
  return nt * (nt + 1) / 2 * NB * NB;
}

void pack_lower_tiles(int64_t n, const double *d_full, double *d_tiles, hipStream_t stream) {
  const int64_t N = padded_size(n), nt = N / NB;
  const int64_t ntiles = nt * (nt + 1) / 2;
  if (ntiles > 0)
    k_pack_lower_tiles<<<static_cast<uint32_t>(ntiles), kThreads, 0, stream>>>(d_full, N, d_tiles);
  GS_DHIP(hipGetLastError());
}

void spd_inverse_device(int64_t n, const std::vector<double> &a, double *d_inv,
                        hipStream_t stream) {
  if (static_cast<int64_t>(a.size()) != n * n)
    throw std::runtime_error("device SPD inverse: size mismatch.");
  const int64_t N = padded_size(n);
  // Padded host copy: identity on the padding.
  std::vector<double> h(static_cast<size_t>(N) * N, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) h[i * N + j] = a[i * n + j];
  for (int64_t i = n; i < N; ++i) h[i * N + i] = 1.0;
  GS_DHIP(hipMemcpyAsync(d_inv, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                         stream));
  eliminate(N, d_inv, stream);
}

void spd_inverse_device_csr(int64_t n, const int32_t *ptr, const int32_t *col,
                            const double *val, double *d_inv, hipStream_t stream) {
  if (n <= 0) throw std::runtime_error("device SPD inverse: empty matrix.");
  const int64_t nnz = ptr[n];
  for (int64_t r = 0; r < n; ++r)
    if (ptr[r + 1] < ptr[r]) throw std::runtime_error("device SPD inverse: invalid row pointer.");
  for (int64_t k = 0; k < nnz; ++k)
    if (col[k] < 0 || col[k] >= n)
      throw std::runtime_error("device SPD inverse: column index out of range.");
  const int64_t N = padded_size(n);
  int32_t *dptr = nullptr, *dcol = nullptr;
  double *dval = nullptr;
  struct Free {
    void *p[3] = {nullptr, nullptr, nullptr};
    ~Free() {
      for (void *q : p)
        if (q) (void)hipFree(q);
    }
  } guard;
  GS_DHIP(hipMalloc(&dptr, (n + 1) * sizeof(int32_t)));
  guard.p[0] = dptr;
  GS_DHIP(hipMalloc(&dcol, (nnz > 0 ? nnz : 1) * sizeof(int32_t)));
  guard.p[1] = dcol;
  GS_DHIP(hipMalloc(&dval, (nnz > 0 ? nnz : 1) * sizeof(double)));
  guard.p[2] = dval;
  GS_DHIP(hipMemcpyAsync(dptr, ptr, (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  if (nnz > 0) {
    GS_DHIP(hipMemcpyAsync(dcol, col, nnz * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    GS_DHIP(hipMemcpyAsync(dval, val, nnz * sizeof(double), hipMemcpyHostToDevice, stream));
  }
  const int64_t fill_blocks = std::min<int64_t>((N * N + kThreads - 1) / kThreads, 16384);
  k_fill_padded<<<static_cast<uint32_t>(fill_blocks), kThreads, 0, stream>>>(d_inv, N, n);
  k_scatter_rows<<<static_cast<uint32_t>((n + kThreads - 1) / kThreads), kThreads, 0, stream>>>(
      dptr, dcol, dval, n, d_inv, N);
  eliminate(N, d_inv, stream);
}

}  // namespace gsdense
