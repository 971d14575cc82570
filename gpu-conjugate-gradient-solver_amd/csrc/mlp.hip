// Learned smoother weights: a small MLP (in -> H -> H -> out, tanh hidden
// layers, linear output) that maps per-level features to per-level damping
// and sweep targets. Forward, backward (MSE loss) and Adam run as HIP
// kernels in FP64. Per-sample gradients are written to a [B x P] buffer and
// summed over the batch in a fixed order, so training is deterministic.
//
// Parameter layout (flat): W1 [H x in], b1 [H], W2 [H x H], b2 [H],
// W3 [out x H], b3 [out], all row-major.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gpusolver.h"
#include "device.hpp"
#include "host/setup.h"

using gsd::DevBuf;

namespace gsc {
void set_error(const char *msg);
}

struct gs_mlp_s {
  int in = 5, hid = 64, out = 2;
  int64_t nparams = 0;
  int64_t cap = 0;  // batch capacity of the work buffers
  DevBuf params, m, v, grad, pgrad, X, T, H1, H2, Y, loss;
  int64_t step = 0;
  hipStream_t stream = nullptr;
  ~gs_mlp_s() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace gsm {

constexpr int kMaxHid = 128;

struct Layout {
  int in, hid, out;
  __host__ __device__ int64_t w1() const { return 0; }
  __host__ __device__ int64_t b1() const { return static_cast<int64_t>(hid) * in; }
  __host__ __device__ int64_t w2() const { return b1() + hid; }
  __host__ __device__ int64_t b2() const { return w2() + static_cast<int64_t>(hid) * hid; }
  __host__ __device__ int64_t w3() const { return b2() + hid; }
  __host__ __device__ int64_t b3() const { return w3() + static_cast<int64_t>(out) * hid; }
  __host__ __device__ int64_t total() const { return b3() + out; }
};

// One workgroup per sample; thread j owns hidden unit j.
__global__ void k_forward(Layout L, const double *__restrict__ p, const double *__restrict__ X,
                          double *__restrict__ H1, double *__restrict__ H2,
                          double *__restrict__ Y, int64_t B) {
  const int64_t s = blockIdx.x;
  if (s >= B) return;
  __shared__ double h1[kMaxHid], h2[kMaxHid];
  const int j = threadIdx.x;
  const double *x = X + s * L.in;
  if (j < L.hid) {
    double a = p[L.b1() + j];
    for (int i = 0; i < L.in; ++i) a += p[L.w1() + static_cast<int64_t>(j) * L.in + i] * x[i];
    h1[j] = tanh(a);
    H1[s * L.hid + j] = h1[j];
  }
  __syncthreads();
  if (j < L.hid) {
    double a = p[L.b2() + j];
    for (int i = 0; i < L.hid; ++i) a += p[L.w2() + static_cast<int64_t>(j) * L.hid + i] * h1[i];
    h2[j] = tanh(a);
    H2[s * L.hid + j] = h2[j];
  }
  __syncthreads();
  if (j < L.out) {
    double a = p[L.b3() + j];
    for (int i = 0; i < L.hid; ++i) a += p[L.w3() + static_cast<int64_t>(j) * L.hid + i] * h2[i];
    Y[s * L.out + j] = a;
  }
}

// Per-sample gradient of 0.5 / B * sum_o (y_o - t_o)^2 written to
// G[s * P ...]; per-sample loss to Ls[s].
__global__ void k_backward(Layout L, const double *__restrict__ p, const double *__restrict__ X,
                           const double *__restrict__ T, const double *__restrict__ H1,
                           const double *__restrict__ H2, const double *__restrict__ Y,
                           double *__restrict__ G, double *__restrict__ Ls, int64_t B) {
  const int64_t s = blockIdx.x;
  if (s >= B) return;
  __shared__ double dy[8], d2[kMaxHid], d1[kMaxHid];
  const int j = threadIdx.x;
  const double invB = 1.0 / static_cast<double>(B);
  const double *x = X + s * L.in;
  const double *h1 = H1 + s * L.hid;
  const double *h2 = H2 + s * L.hid;
  double *g = G + s * L.total();
  if (j < L.out) dy[j] = (Y[s * L.out + j] - T[s * L.out + j]) * invB;
  __syncthreads();
  if (j == 0) {
    double l = 0.0;
    for (int o = 0; o < L.out; ++o) {
      const double e = Y[s * L.out + o] - T[s * L.out + o];
      l += 0.5 * e * e * invB;
    }
    Ls[s] = l;
  }
  if (j < L.hid) {
    double a = 0.0;
    for (int o = 0; o < L.out; ++o) a += p[L.w3() + static_cast<int64_t>(o) * L.hid + j] * dy[o];
    d2[j] = a * (1.0 - h2[j] * h2[j]);
    for (int o = 0; o < L.out; ++o) g[L.w3() + static_cast<int64_t>(o) * L.hid + j] = dy[o] * h2[j];
  }
  if (j < L.out) g[L.b3() + j] = dy[j];
  __syncthreads();
  if (j < L.hid) {
    double a = 0.0;
    for (int k = 0; k < L.hid; ++k) a += p[L.w2() + static_cast<int64_t>(k) * L.hid + j] * d2[k];
    d1[j] = a * (1.0 - h1[j] * h1[j]);
    g[L.b2() + j] = d2[j];
    for (int i = 0; i < L.hid; ++i) g[L.w2() + static_cast<int64_t>(j) * L.hid + i] = d2[j] * h1[i];
  }
  __syncthreads();
  if (j < L.hid) {
    g[L.b1() + j] = d1[j];
    for (int i = 0; i < L.in; ++i) g[L.w1() + static_cast<int64_t>(j) * L.in + i] = d1[j] * x[i];
  }
}

// grad[q] = sum_s G[s, q] in sample order; loss = sum_s Ls[s].
__global__ void k_reduce_grad(const double *__restrict__ G, const double *__restrict__ Ls,
                              int64_t B, int64_t P, double *__restrict__ grad,
                              double *__restrict__ loss) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < P) {
    double a = 0.0;
    for (int64_t s = 0; s < B; ++s) a += G[s * P + q];
    grad[q] = a;
  }
  if (q == 0) {
    double l = 0.0;
    for (int64_t s = 0; s < B; ++s) l += Ls[s];
    *loss = l;
  }
}

__global__ void k_adam(double *__restrict__ p, const double *__restrict__ g, double *__restrict__ m,
                       double *__restrict__ v, int64_t P, double lr, double b1, double b2,
                       double eps, double c1, double c2) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= P) return;
  const double gq = g[q];
  const double mq = b1 * m[q] + (1.0 - b1) * gq;
  const double vq = b2 * v[q] + (1.0 - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  const double mh = mq / c1;
  const double vh = vq / c2;
  p[q] -= lr * mh / (sqrt(vh) + eps);
}

template <class F>
int mguard(F &&f) {
  try {
    f();
    return GS_OK;
  } catch (const std::exception &e) {
    gsc::set_error(e.what());
    return GS_ERR;
  }
}

Layout layout(const gs_mlp_s *h) { return Layout{h->in, h->hid, h->out}; }

void ensure_batch(gs_mlp_s *h, int64_t B) {
  if (B <= h->cap) return;
  h->X.alloc(B * h->in * sizeof(double));
  h->T.alloc(B * h->out * sizeof(double));
  h->H1.alloc(B * h->hid * sizeof(double));
  h->H2.alloc(B * h->hid * sizeof(double));
  h->Y.alloc(B * h->out * sizeof(double));
  h->pgrad.alloc(B * h->nparams * sizeof(double) + B * sizeof(double));
  h->cap = B;
}

double *loss_parts(gs_mlp_s *h) {
  return reinterpret_cast<double *>(static_cast<char *>(h->pgrad.get()) +
                                    h->cap * h->nparams * sizeof(double));
}

void forward_backward(gs_mlp_s *h, int64_t B, bool backward) {
  const Layout L = layout(h);
  const int threads = 128;
  hipLaunchKernelGGL(k_forward, dim3(static_cast<uint32_t>(B)), dim3(threads), 0, h->stream, L,
                     h->params.as<double>(), h->X.as<double>(), h->H1.as<double>(),
                     h->H2.as<double>(), h->Y.as<double>(), B);
  if (!backward) return;
  hipLaunchKernelGGL(k_backward, dim3(static_cast<uint32_t>(B)), dim3(threads), 0, h->stream, L,
                     h->params.as<double>(), h->X.as<double>(), h->T.as<double>(),
                     h->H1.as<double>(), h->H2.as<double>(), h->Y.as<double>(),
                     h->pgrad.as<double>(), loss_parts(h), B);
  const uint32_t g = static_cast<uint32_t>((h->nparams + 255) / 256);
  hipLaunchKernelGGL(k_reduce_grad, dim3(g), dim3(256), 0, h->stream, h->pgrad.as<double>(),
                     loss_parts(h), B, h->nparams, h->grad.as<double>(), h->loss.as<double>());
}

void adam(gs_mlp_s *h, double lr, double b1, double b2, double eps) {
  h->step += 1;
  const double c1 = 1.0 - std::pow(b1, static_cast<double>(h->step));
  const double c2 = 1.0 - std::pow(b2, static_cast<double>(h->step));
  const uint32_t g = static_cast<uint32_t>((h->nparams + 255) / 256);
  hipLaunchKernelGGL(k_adam, dim3(g), dim3(256), 0, h->stream, h->params.as<double>(),
                     h->grad.as<double>(), h->m.as<double>(), h->v.as<double>(), h->nparams, lr,
                     b1, b2, eps, c1, c2);
}

}  // namespace gsm

using namespace gsm;

extern "C" {

int gs_mlp_create(int32_t in, int32_t hidden, int32_t out, uint64_t seed, gs_mlp_h *res) {
  return mguard([&] {
    if (in <= 0 || out <= 0 || out > 8 || hidden <= 0 || hidden > kMaxHid)
      throw std::runtime_error("gs_mlp_create: unsupported layer sizes.");
    auto h = std::make_unique<gs_mlp_s>();
    h->in = in;
    h->hid = hidden;
    h->out = out;
    const Layout L{in, hidden, out};
    h->nparams = L.total();
    // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, from the
    // SplitMix64 stream.
    std::vector<double> p(h->nparams, 0.0), u(h->nparams);
    gsx::uniform_pm1(seed, h->nparams, u.data());
    for (int64_t q = L.w1(); q < L.b1(); ++q) p[q] = u[q] / std::sqrt(static_cast<double>(in));
    for (int64_t q = L.w2(); q < L.b2(); ++q) p[q] = u[q] / std::sqrt(static_cast<double>(hidden));
    for (int64_t q = L.w3(); q < L.b3(); ++q) p[q] = u[q] / std::sqrt(static_cast<double>(hidden));
    GS_HIP(hipStreamCreate(&h->stream));
    h->params.upload(p);
    std::vector<double> z(h->nparams, 0.0);
    h->m.upload(z);
    h->v.upload(z);
    h->grad.upload(z);
    h->loss.alloc(sizeof(double));
    *res = h.release();
  });
}

int gs_mlp_num_params(gs_mlp_h h, int64_t *n) {
  return mguard([&] { *n = h->nparams; });
}

int gs_mlp_get_params(gs_mlp_h h, double *out) {
  return mguard([&] {
    GS_HIP(hipStreamSynchronize(h->stream));
    GS_HIP(hipMemcpy(out, h->params.get(), h->nparams * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int gs_mlp_set_params(gs_mlp_h h, const double *in) {
  return mguard([&] {
    GS_HIP(hipStreamSynchronize(h->stream));
    GS_HIP(hipMemcpy(h->params.get(), in, h->nparams * sizeof(double), hipMemcpyHostToDevice));
    std::vector<double> z(h->nparams, 0.0);
    GS_HIP(hipMemcpy(h->m.get(), z.data(), h->nparams * sizeof(double), hipMemcpyHostToDevice));
    GS_HIP(hipMemcpy(h->v.get(), z.data(), h->nparams * sizeof(double), hipMemcpyHostToDevice));
    h->step = 0;
  });
}

int gs_mlp_forward(gs_mlp_h h, const double *X, int64_t B, double *Y) {
  return mguard([&] {
    if (B <= 0) throw std::runtime_error("gs_mlp_forward: empty batch.");
    ensure_batch(h, B);
    GS_HIP(hipMemcpyAsync(h->X.get(), X, B * h->in * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    forward_backward(h, B, false);
    GS_HIP(hipMemcpyAsync(Y, h->Y.get(), B * h->out * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    GS_HIP(hipStreamSynchronize(h->stream));
  });
}

int gs_mlp_loss_grad(gs_mlp_h h, const double *X, const double *T, int64_t B, double *loss,
                     double *grad) {
  return mguard([&] {
    if (B <= 0) throw std::runtime_error("gs_mlp_loss_grad: empty batch.");
    ensure_batch(h, B);
    GS_HIP(hipMemcpyAsync(h->X.get(), X, B * h->in * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    GS_HIP(hipMemcpyAsync(h->T.get(), T, B * h->out * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    forward_backward(h, B, true);
    if (loss)
      GS_HIP(hipMemcpyAsync(loss, h->loss.get(), sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (grad)
      GS_HIP(hipMemcpyAsync(grad, h->grad.get(), h->nparams * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
    GS_HIP(hipStreamSynchronize(h->stream));
  });
}

int gs_mlp_adam_step(gs_mlp_h h, double lr, double beta1, double beta2, double eps) {
  return mguard([&] {
    adam(h, lr, beta1, beta2, eps);
    GS_HIP(hipStreamSynchronize(h->stream));
  });
}

// Full-batch Adam for `epochs` steps on device; *final_loss is the loss
// before the last update.
int gs_mlp_train(gs_mlp_h h, const double *X, const double *T, int64_t B, int64_t epochs,
                 double lr, double *final_loss) {
  return mguard([&] {
    if (B <= 0 || epochs <= 0) throw std::runtime_error("gs_mlp_train: empty batch or epochs.");
    ensure_batch(h, B);
    GS_HIP(hipMemcpyAsync(h->X.get(), X, B * h->in * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    GS_HIP(hipMemcpyAsync(h->T.get(), T, B * h->out * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    for (int64_t e = 0; e < epochs; ++e) {
      forward_backward(h, B, true);
      adam(h, lr, 0.9, 0.999, 1e-8);
    }
    if (final_loss)
      GS_HIP(hipMemcpyAsync(final_loss, h->loss.get(), sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
    GS_HIP(hipStreamSynchronize(h->stream));
  });
}

int gs_mlp_destroy(gs_mlp_h h) {
  delete h;
  return GS_OK;
}

}  // extern "C"
