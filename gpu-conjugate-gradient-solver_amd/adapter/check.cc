// Checks gpuSolver::HipBackend through the reference's interface only
// (gpuSolver::Backend &, include/GPUSolver/Backend.h:15-159): every method
// runs once through the adapter and once through the C ABI it forwards to,
// on the same FP32 inputs, and the results must agree bit for bit (the C
// ABI itself is checked against the oracle elsewhere). Then the error
// behaviour of the interface (size checks, objects of another backend) and
// a conjugate-gradient loop written against Backend & alone, in the order
// of src/CGSolver.cc:87-306 with the identity preconditioner. Prints one
// line per group and exits non-zero on the first failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "HipBackend.h"

using namespace gpuSolver;

namespace {

int checks = 0;

void expect(bool c, const std::string &what) {
  ++checks;
  if (!c) throw std::runtime_error("check failed: " + what);
}

void ok(int rc) {
  if (rc != GS_OK) throw std::runtime_error(gs_last_error());
}

bool same_bits(const float *a, const float *b, size_t n) {
  return n == 0 || std::memcmp(a, b, n * sizeof(float)) == 0;
}

// 5-point Laplacian on an nx x ny grid, rows in natural order, columns
// sorted; FP32 values (4 on the diagonal, -1 off it) scaled per row so the
// sums are not all integers.
struct Csr {
  int rows = 0;
  std::vector<int> ptr, col;
  std::vector<float> val;
};

Csr laplacian(int nx, int ny) {
  Csr A;
  A.rows = nx * ny;
  A.ptr.push_back(0);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int r = j * nx + i;
      const float s = 1.0f + 0.001f * static_cast<float>(r % 7);
      auto add = [&](int c, float v) {
        A.col.push_back(c);
        A.val.push_back(v * s);
      };
      if (j > 0) add(r - nx, -1.0f);
      if (i > 0) add(r - 1, -1.0f);
      add(r, 4.0f);
      if (i + 1 < nx) add(r + 1, -1.0f);
      if (j + 1 < ny) add(r + nx, -1.0f);
      A.ptr.push_back(static_cast<int>(A.col.size()));
    }
  return A;
}

std::vector<float> values(size_t n, uint32_t seed) {
  std::vector<float> v(n);
  uint32_t s = seed * 2654435761u + 1u;
  for (auto &x : v) {
    s = s * 1664525u + 1013904223u;
    x = static_cast<float>(static_cast<int32_t>(s >> 8) % 2001 - 1000) / 997.0f;
  }
  return v;
}

// C-ABI twins of the adapter's objects.
struct Raw {
  gs_backend_h b = nullptr;
  gs_encoder_h e = nullptr;
  Raw() {
    ok(gs_backend_create(GS_DTYPE_F32, &b));
    ok(gs_backend_create_encoder(b, &e));
  }
  ~Raw() {
    gs_encoder_destroy(e);
    gs_backend_destroy(b);
  }
  gs_dvec_h vec(const std::vector<float> &v) {
    gs_dvec_h h = nullptr;
    ok(gs_backend_create_vector(b, static_cast<int64_t>(v.size()), v.data(), &h));
    return h;
  }
  gs_dscalar_h scalar(float x) {
    gs_dscalar_h h = nullptr;
    ok(gs_backend_create_scalar(b, x, &h));
    return h;
  }
  std::vector<float> get(gs_dvec_h h, size_t n) {
    std::vector<float> out(n);
    ok(gs_backend_submit_and_wait(b, e));
    ok(gs_dvec_download(h, out.data()));
    return out;
  }
  float get(gs_dscalar_h h) {
    double v = 0;
    ok(gs_backend_submit_and_wait(b, e));
    ok(gs_dscalar_download(h, &v));
    return static_cast<float>(v);
  }
};

void vector_ops(Backend &be) {
  Raw raw;
  const size_t n = 100003;
  const auto x0 = values(n, 1), y0 = values(n, 2), s0 = values(n, 3);
  std::unique_ptr<BackendEncoder> enc(be.createEncoder());
  std::unique_ptr<DeviceVector> x(be.createVector(n, x0.data())), y(be.createVector(n, y0.data())),
      s(be.createVector(n, s0.data())), z(be.createVector(n));
  std::unique_ptr<DeviceScalar> a(be.createScalar(0.625f)), d(be.createScalar());
  gs_dvec_h rx = raw.vec(x0), ry = raw.vec(y0), rs = raw.vec(s0), rz = raw.vec(std::vector<float>(n));
  gs_dscalar_h ra = raw.scalar(0.625f), rd = raw.scalar(0.0f);

  be.encodeScale(*enc, *x, 1.75f);
  ok(gs_encode_scale(raw.e, rx, 1.75f));
  be.encodeAxpy(*enc, -0.375f, *x, *y);
  ok(gs_encode_axpy(raw.e, -0.375f, rx, ry));
  be.encodeAxpy(*enc, *a, *y, *x);
  ok(gs_encode_axpy_dev(raw.e, ra, ry, rx));
  be.encodeScale(*enc, *y, *a);
  ok(gs_encode_scale_dev(raw.e, ry, ra));
  be.encodeScaleVector(*enc, *s, *y);
  ok(gs_encode_scale_vector(raw.e, rs, ry));
  be.encodeCopy(*enc, *y, *z);
  ok(gs_encode_copy(raw.e, ry, rz));
  be.encodeDot(*enc, *x, *z, *d);
  ok(gs_encode_dot(raw.e, rx, rz, rd));
  be.submitAndWait(*enc);
  const auto gx = raw.get(rx, n), gy = raw.get(ry, n), gz = raw.get(rz, n);
  expect(same_bits(x->download(), gx.data(), n), "scale / axpy: x");
  expect(same_bits(y->download(), gy.data(), n), "scaleVector: y");
  expect(same_bits(z->download(), gz.data(), n), "copy: z");
  const float dd = d->download(), rdd = raw.get(rd);
  expect(std::memcmp(&dd, &rdd, 4) == 0, "dot");
  // Host restatement of the elementwise chain (same FP32 operations).
  std::vector<float> hx(x0), hy(y0);
  for (size_t i = 0; i < n; ++i) {
    hx[i] *= 1.75f;
    hy[i] += -0.375f * hx[i];
    hx[i] += 0.625f * hy[i];
    hy[i] *= 0.625f;
    hy[i] *= s0[i];
  }
  expect(same_bits(x->download(), hx.data(), n), "x against the host");
  expect(same_bits(y->download(), hy.data(), n), "y against the host");
  double hd = 0;
  for (size_t i = 0; i < n; ++i) hd += static_cast<double>(hx[i]) * hy[i];
  expect(std::fabs(dd - hd) <= 1e-5 * std::fabs(hd) + 1e-6, "dot against the host");

  be.encodeSetZero(*enc, *z);
  ok(gs_encode_set_zero(raw.e, rz));
  be.submitAndWait(*enc);
  const auto gz0 = raw.get(rz, n);
  expect(same_bits(z->download(), gz0.data(), n), "setZero");

  // Standalone API.
  be.scale(*x, 0.5f);
  be.axpy(2.0f, *y, *x);
  be.axpy(*a, *y, *x);
  be.scale(*x, *a);
  be.scaleVectorByVector(*s, *x);
  be.dot(*x, *y, *d);
  ok(gs_encode_scale(raw.e, rx, 0.5f));
  ok(gs_encode_axpy(raw.e, 2.0f, ry, rx));
  ok(gs_encode_axpy_dev(raw.e, ra, ry, rx));
  ok(gs_encode_scale_dev(raw.e, rx, ra));
  ok(gs_encode_scale_vector(raw.e, rs, rx));
  ok(gs_encode_dot(raw.e, rx, ry, rd));
  const auto gx2 = raw.get(rx, n);
  expect(same_bits(x->download(), gx2.data(), n), "standalone vector API");
  const float d2 = d->download(), rd2 = raw.get(rd);
  expect(std::memcmp(&d2, &rd2, 4) == 0, "standalone dot");

  for (auto h : {rx, ry, rs, rz}) gs_dvec_destroy(h);
  gs_dscalar_destroy(ra);
  gs_dscalar_destroy(rd);
  std::printf("adapter_check: vector operations ok\n");
}

void scalar_ops(Backend &be) {
  std::unique_ptr<BackendEncoder> enc(be.createEncoder());
  std::unique_ptr<DeviceScalar> a(be.createScalar(7.5f)), b(be.createScalar(-2.25f)),
      c(be.createScalar());
  be.encodeScalarDivide(*enc, *a, *b, *c);
  be.submitAndWait(*enc);
  expect(c->download() == 7.5f / -2.25f, "scalarDivide");
  be.encodeScalarMultiply(*enc, *a, *b, *c);
  be.submitAndWait(*enc);
  expect(c->download() == 7.5f * -2.25f, "scalarMultiply");
  be.encodeScalarNegate(*enc, *b, *c);
  be.submitAndWait(*enc);
  expect(c->download() == 2.25f, "scalarNegate");
  be.encodeScalarSqrt(*enc, *c, *c);
  be.submitAndWait(*enc);
  expect(c->download() == std::sqrt(2.25f), "scalarSqrt");
  be.encodeScalarSet(*enc, 3.0f, *c);
  be.encodeScalarCopy(*enc, *c, *a);
  be.submitAndWait(*enc);
  expect(a->download() == 3.0f, "scalarSet / scalarCopy");
  be.scalarSet(0.1f, *a);
  be.scalarCopy(*a, *b);
  be.scalarMultiply(*a, *b, *c);
  expect(c->download() == 0.1f * 0.1f, "standalone scalarMultiply");
  be.scalarDivide(*c, *a, *b);
  expect(b->download() == (0.1f * 0.1f) / 0.1f, "standalone scalarDivide");
  be.scalarNegate(*b, *c);
  be.scalarSqrt(*a, *b);
  expect(c->download() == -((0.1f * 0.1f) / 0.1f) && b->download() == std::sqrt(0.1f),
         "standalone scalarNegate / scalarSqrt");
  std::printf("adapter_check: scalar operations ok\n");
}

// Red-black colouring of the grid (colour (i + j) % 2): vertex lists and
// the start of the second colour.
std::vector<int> red_black(int nx, int ny, size_t *nred) {
  std::vector<int> v;
  for (int c = 0; c < 2; ++c) {
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        if ((i + j) % 2 == c) v.push_back(j * nx + i);
    if (c == 0) *nred = v.size();
  }
  return v;
}

void matrix_ops(HipBackend &hb) {
  Backend &be = hb;
  Raw raw;
  const int nx = 61, ny = 47;
  const Csr A = laplacian(nx, ny);
  const size_t n = A.rows, nnz = A.col.size();
  std::unique_ptr<DeviceCSRMatrix> M(
      hb.createCSRMatrix(n, n, nnz, A.ptr.data(), A.col.data(), A.val.data()));
  expect(M->rows() == n && M->cols() == n && M->nnz() == nnz, "matrix dimensions");
  std::vector<double> vd(A.val.begin(), A.val.end());
  gs_csr_h hc = nullptr;
  ok(gs_csr_create(n, n, nnz, A.ptr.data(), A.col.data(), vd.data(), &hc));
  gs_dcsr_h rm = nullptr;
  ok(gs_backend_create_csr(raw.b, hc, &rm));

  // SpMV.
  const auto x0 = values(n, 11), b0 = values(n, 12);
  std::unique_ptr<BackendEncoder> enc(be.createEncoder());
  std::unique_ptr<DeviceVector> x(be.createVector(n, x0.data())), y(be.createVector(n)),
      rhs(be.createVector(n, b0.data()));
  gs_dvec_h rx = raw.vec(x0), ry = raw.vec(std::vector<float>(n)), rb = raw.vec(b0);
  be.encodeSpmv(*enc, *M, *x, *y);
  be.submitAndWait(*enc);
  ok(gs_encode_spmv(raw.e, rm, rx, ry));
  std::vector<float> hy(n);
  for (size_t r = 0; r < n; ++r) {
    float acc = 0.0f;
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k) acc += A.val[k] * x0[A.col[k]];
    hy[r] = acc;
  }
  expect(same_bits(y->download(), raw.get(ry, n).data(), n), "spmv against the C ABI");
  expect(same_bits(y->download(), hy.data(), n), "spmv against the host (CSR order)");

  // ELL from packed arrays (slot-major, padding column 0 / value 0, the
  // rest of each row in an overflow CSR), against the C ABI's own packing.
  const size_t w = 4;
  std::vector<int> ecol(n * w, 0), optr(n + 1, 0), ocol;
  std::vector<float> eval(n * w, 0.0f), oval;
  for (size_t r = 0; r < n; ++r) {
    size_t s = 0;
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k, ++s) {
      if (s < w) {
        ecol[s * n + r] = A.col[k];
        eval[s * n + r] = A.val[k];
      } else {
        ocol.push_back(A.col[k]);
        oval.push_back(A.val[k]);
      }
    }
    optr[r + 1] = static_cast<int>(ocol.size());
  }
  std::unique_ptr<DeviceELLMatrix> E(hb.createELLMatrix(n, n, w, ecol.data(), eval.data(),
                                                        ocol.size(), optr.data(), ocol.data(),
                                                        oval.data()));
  expect(E->rows() == n && E->cols() == n && E->ellWidth() == w && E->overflowNnz() == ocol.size(),
         "ELL dimensions");
  gs_dell_h re = nullptr;
  ok(gs_backend_create_ell(raw.b, hc, static_cast<int32_t>(w), &re));
  hb.encodeSpmvELL(*enc, *E, *x, *y);
  be.submitAndWait(*enc);
  ok(gs_encode_spmv_ell(raw.e, re, rx, ry));
  expect(same_bits(y->download(), raw.get(ry, n).data(), n), "ELL spmv against the C ABI");
  expect(same_bits(y->download(), hy.data(), n), "ELL spmv against the CSR-order host sum");

  // One symmetric multicolour Gauss-Seidel sweep (red, black, black, red).
  size_t nred = 0;
  const auto colors = red_black(nx, ny, &nred);
  std::unique_ptr<DeviceIndexVector> cv(be.createIndexVector(colors.size(), colors.data()));
  expect(cv->size() == n, "index vector size");
  gs_divec_h rc = nullptr;
  ok(gs_backend_create_index_vector(raw.b, n, colors.data(), &rc));
  const std::pair<size_t, size_t> order[] = {{0, nred}, {nred, n - nred}, {nred, n - nred}, {0, nred}};
  for (auto [st, ct] : order) {
    be.encodeGaussSeidelColor(*enc, *M, *cv, st, ct, *rhs, *x, 0.9f);
    ok(gs_encode_gauss_seidel_color(raw.e, rm, rc, st, ct, rb, rx, 0.9f));
  }
  be.submitAndWait(*enc);
  std::vector<float> hx(x0);
  for (auto [st, ct] : order)
    for (size_t k = st; k < st + ct; ++k) {
      const int r = colors[k];
      float diag = 0.0f, off = 0.0f;
      for (int e = A.ptr[r]; e < A.ptr[r + 1]; ++e) {
        if (A.col[e] == r) {
          diag = A.val[e];
          continue;
        }
        off += A.val[e] * hx[A.col[e]];
      }
      const float nv = (b0[r] - off) / diag;
      hx[r] = (1.0f - 0.9f) * hx[r] + 0.9f * nv;
    }
  expect(same_bits(x->download(), raw.get(rx, n).data(), n), "Gauss-Seidel against the C ABI");
  expect(same_bits(x->download(), hx.data(), n), "Gauss-Seidel against the host");

  // updateValues on the matrix and the index vector.
  std::vector<float> v2(A.val);
  for (auto &v : v2) v *= 0.5f;
  M->updateValues(v2.data(), v2.size());
  std::vector<int> rev(colors.rbegin(), colors.rend());
  cv->updateValues(rev.data(), rev.size());
  expect(std::memcmp(cv->download(), rev.data(), n * sizeof(int)) == 0, "index updateValues");
  be.spmv(*M, *rhs, *y);
  for (size_t r = 0; r < n; ++r) {
    float acc = 0.0f;
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k) acc += v2[k] * b0[A.col[k]];
    hy[r] = acc;
  }
  expect(same_bits(y->download(), hy.data(), n), "matrix updateValues");

  for (auto h : {rx, ry, rb}) gs_dvec_destroy(h);
  gs_divec_destroy(rc);
  gs_dell_destroy(re);
  gs_dcsr_destroy(rm);
  gs_csr_destroy(hc);
  std::printf("adapter_check: matrix operations ok\n");
}

// 3x3 blocks: the grid Laplacian's pattern with a dense SPD diagonal block
// per node and -0.25 * I coupling blocks, three unknowns per node.
Csr block_laplacian(int nx, int ny) {
  const Csr G = laplacian(nx, ny);
  Csr A;
  A.rows = 3 * G.rows;
  A.ptr.push_back(0);
  const float D[9] = {4.0f, 1.0f, 0.5f, 1.0f, 5.0f, 1.0f, 0.5f, 1.0f, 6.0f};
  for (int node = 0; node < G.rows; ++node)
    for (int l = 0; l < 3; ++l) {
      for (int k = G.ptr[node]; k < G.ptr[node + 1]; ++k) {
        const int c = G.col[k];
        for (int m = 0; m < 3; ++m) {
          const float v = c == node ? D[3 * l + m] * (1.0f + 0.01f * (node % 5))
                                    : (l == m ? -0.25f : 0.125f * ((l + m) % 2));
          if (v == 0.0f) continue;
          A.col.push_back(3 * c + m);
          A.val.push_back(v);
        }
      }
      A.ptr.push_back(static_cast<int>(A.col.size()));
    }
  return A;
}

void block_ops(HipBackend &hb) {
  Backend &be = hb;
  Raw raw;
  const int nx = 23, ny = 19;
  const Csr A = block_laplacian(nx, ny);
  const size_t n = A.rows, nb = n / 3, nnz = A.col.size();
  std::unique_ptr<DeviceCSRMatrix> M(
      hb.createCSRMatrix(n, n, nnz, A.ptr.data(), A.col.data(), A.val.data()));
  std::vector<double> vd(A.val.begin(), A.val.end());
  gs_csr_h hc = nullptr;
  ok(gs_csr_create(n, n, nnz, A.ptr.data(), A.col.data(), vd.data(), &hc));
  gs_dcsr_h rm = nullptr;
  ok(gs_backend_create_csr(raw.b, hc, &rm));
  std::vector<double> invd(9 * nb);
  ok(gs_inverse_diag_blocks3(hc, invd.data()));
  const std::vector<float> inv(invd.begin(), invd.end());

  const auto x0 = values(n, 21), b0 = values(n, 22);
  std::unique_ptr<BackendEncoder> enc(be.createEncoder());
  std::unique_ptr<DeviceVector> x(be.createVector(n, x0.data())), rhs(be.createVector(n, b0.data())),
      dinv(be.createVector(inv.size(), inv.data())), corr(be.createVector(n));
  gs_dvec_h rx = raw.vec(x0), rb = raw.vec(b0), rd = raw.vec(inv), rcorr = raw.vec(std::vector<float>(n));
  size_t nred = 0;
  const auto colors = red_black(nx, ny, &nred);
  std::unique_ptr<DeviceIndexVector> cb(be.createIndexVector(colors.size(), colors.data()));
  gs_divec_h rc = nullptr;
  ok(gs_backend_create_index_vector(raw.b, nb, colors.data(), &rc));
  // (begin, end) as the CPU and CUDA backends read the pair.
  const std::pair<size_t, size_t> ranges[] = {{0, nred}, {nred, nb}, {nred, nb}, {0, nred}};
  for (auto [b, e] : ranges) {
    be.encodeBlockGaussSeidelColor3x3(*enc, *M, *cb, *dinv, b, e, *rhs, *x, 0.8f);
    ok(gs_encode_block_gauss_seidel_color3x3(raw.e, rm, rc, rd, b, e, rb, rx, 0.8f));
  }
  be.encodeApplyBlockInverses3x3(*enc, *dinv, *rhs, *corr);
  ok(gs_encode_apply_block_inverses3x3(raw.e, rd, rb, rcorr));
  be.submitAndWait(*enc);
  expect(same_bits(x->download(), raw.get(rx, n).data(), n), "block Gauss-Seidel against the C ABI");
  expect(same_bits(corr->download(), raw.get(rcorr, n).data(), n),
         "block inverses against the C ABI");
  std::vector<float> hc3(n);
  for (size_t b = 0; b < nb; ++b)
    for (int l = 0; l < 3; ++l) {
      float acc = 0.0f;
      for (int m = 0; m < 3; ++m) acc += inv[9 * b + 3 * l + m] * b0[3 * b + m];
      hc3[3 * b + l] = acc;
    }
  expect(same_bits(corr->download(), hc3.data(), n), "block inverses against the host");

  for (auto h : {rx, rb, rd, rcorr}) gs_dvec_destroy(h);
  gs_divec_destroy(rc);
  gs_dcsr_destroy(rm);
  gs_csr_destroy(hc);
  std::printf("adapter_check: block operations ok\n");
}

// A vector of some other backend.
class ForeignVector final : public DeviceVector {
 public:
  std::size_t size() const override { return 4; }
  void updateValues(const float *, std::size_t) override {}
  float *download() const override { return nullptr; }
};

template <class F>
bool throws(F &&f, const char *needle) {
  try {
    f();
  } catch (const std::runtime_error &e) {
    return std::strstr(e.what(), needle) != nullptr;
  }
  return false;
}

void errors(HipBackend &hb) {
  Backend &be = hb;
  std::unique_ptr<BackendEncoder> enc(be.createEncoder());
  std::unique_ptr<DeviceVector> a(be.createVector(4)), b(be.createVector(5));
  const float v[5] = {1, 2, 3, 4, 5};
  expect(throws([&] { a->updateValues(v, 5); }, "size mismatch"), "vector updateValues size");
  expect(throws([&] { be.encodeCopy(*enc, *a, *b); }, "size mismatch"), "copy size");
  expect(throws([&] { be.createVector(3, nullptr); }, "nullptr"), "null values");
  ForeignVector f;
  expect(throws([&] { be.encodeCopy(*enc, f, *a); }, "not created by HipBackend"),
         "object of another backend");
  const int ptr[3] = {0, 1, 2}, col[2] = {0, 7};
  const float val[2] = {1.0f, 1.0f};
  expect(throws([&] { hb.createCSRMatrix(2, 2, 2, ptr, col, val); }, "out of bounds"),
         "column index out of bounds");
  const int ecol[2] = {0, 2}, optr[3] = {0, 0, 0};
  expect(throws([&] { hb.createELLMatrix(2, 2, 1, ecol, val, 0, optr, nullptr, nullptr); },
                "out of bounds"),
         "ELL column index out of bounds");
  std::unique_ptr<DeviceIndexVector> iv(be.createIndexVector(3));
  const int iv4[4] = {0, 1, 2, 3};
  expect(throws([&] { iv->updateValues(iv4, 4); }, "size mismatch"), "index updateValues size");
  std::printf("adapter_check: error behaviour ok\n");
}

// src/CGSolver.cc:87-306 through Backend & with the identity
// preconditioner (IdentityPreconditioner::apply is encodeCopy).
int cg(Backend &be, const DeviceCSRMatrix &A, const std::vector<float> &bh, std::vector<float> &xh,
       float tol, int maxit) {
  const size_t n = bh.size();
  std::unique_ptr<DeviceVector> x(be.createVector(n, xh.data())), b(be.createVector(n, bh.data())),
      res(be.createVector(n)), Ax(be.createVector(n)), z(be.createVector(n)),
      z1(be.createVector(n)), p(be.createVector(n)), Ap(be.createVector(n)),
      res1(be.createVector(n));
  std::unique_ptr<DeviceScalar> pAp(be.createScalar()), alpha(be.createScalar()),
      nalpha(be.createScalar()), beta(be.createScalar()), rr(be.createScalar()),
      bb(be.createScalar()), rel(be.createScalar()), rho(be.createScalar()),
      rho1(be.createScalar());
  std::unique_ptr<BackendEncoder> e(be.createEncoder());
  be.encodeSpmv(*e, A, *x, *Ax);
  be.encodeCopy(*e, *b, *res);
  be.encodeAxpy(*e, -1.0f, *Ax, *res);
  be.encodeCopy(*e, *res, *z);
  be.encodeCopy(*e, *z, *p);
  be.encodeDot(*e, *b, *b, *bb);
  be.encodeDot(*e, *res, *z, *rho);
  be.encodeDot(*e, *res, *res, *rr);
  be.encodeScalarDivide(*e, *rr, *bb, *rel);
  be.submitAndWait(*e);
  int its = maxit;
  for (int it = 0; it < maxit; ++it) {
    std::unique_ptr<BackendEncoder> k(be.createEncoder());
    be.encodeSpmv(*k, A, *p, *Ap);
    be.encodeDot(*k, *p, *Ap, *pAp);
    be.encodeScalarDivide(*k, *rho, *pAp, *alpha);
    be.encodeAxpy(*k, *alpha, *p, *x);
    be.encodeScalarNegate(*k, *alpha, *nalpha);
    be.encodeCopy(*k, *res, *res1);
    be.encodeAxpy(*k, *nalpha, *Ap, *res1);
    be.encodeDot(*k, *res1, *res1, *rr);
    be.encodeScalarDivide(*k, *rr, *bb, *rel);
    be.encodeCopy(*k, *res1, *z1);
    be.encodeDot(*k, *res1, *z1, *rho1);
    be.encodeScalarDivide(*k, *rho1, *rho, *beta);
    be.encodeScale(*k, *p, *beta);
    be.encodeAxpy(*k, 1.0f, *z1, *p);
    be.submitAndWait(*k);
    const float r = rel->download();
    if (!std::isfinite(r)) throw std::runtime_error("CG diverged");
    if (r < tol * tol) {
      its = it;
      break;
    }
    std::swap(rho, rho1);
    std::swap(res, res1);
    std::swap(z, z1);
  }
  const float *xd = x->download();
  xh.assign(xd, xd + n);
  return its;
}

void cg_loop(HipBackend &hb) {
  const Csr A = laplacian(61, 47);
  const size_t n = A.rows;
  std::unique_ptr<DeviceCSRMatrix> M(
      hb.createCSRMatrix(n, n, A.col.size(), A.ptr.data(), A.col.data(), A.val.data()));
  const auto b = values(n, 31);
  std::vector<float> x(n, 0.0f);
  const int its = cg(hb, *M, b, x, 1e-5f, 1000);
  double rn = 0, bn = 0;
  for (size_t r = 0; r < n; ++r) {
    double acc = 0;
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k) acc += static_cast<double>(A.val[k]) * x[A.col[k]];
    rn += (b[r] - acc) * (b[r] - acc);
    bn += static_cast<double>(b[r]) * b[r];
  }
  const double rel = std::sqrt(rn / bn);
  expect(its < 1000 && rel < 3e-5, "CG through Backend & converges");
  std::printf("adapter_check: CG through Backend & ok (%d iterations, relres %.2e)\n", its, rel);
}

}  // namespace

int main() {
  try {
    HipBackend hb;
    Backend &be = hb;
    vector_ops(be);
    scalar_ops(be);
    matrix_ops(hb);
    block_ops(hb);
    errors(hb);
    cg_loop(hb);
    std::printf("adapter_check: %d checks passed\n", checks);
    return 0;
  } catch (const std::exception &e) {
    std::printf("adapter_check: FAILED after %d checks: %s\n", checks, e.what());
    return 1;
  }
}
