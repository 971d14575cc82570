// Problem generators: structured Poisson operators, a finite-volume
// variable-coefficient operator over a Gaussian random field, and P1-tet
// linear elasticity on a Kuhn-split cube. They replace the meshes the
// reference reads from its (absent) data/ directory.
#include "setup.h"

#include <algorithm>
#include <array>
#include <cmath>

namespace gsx {

uint64_t splitmix_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline double u01(uint64_t seed, uint64_t i) {
  return static_cast<double>(splitmix_at(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

void uniform_pm1(uint64_t seed, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) out[i] = 2.0 * u01(seed, i) - 1.0;
}

std::vector<double> gaussian_random_field(int nx, int ny, int nz, uint64_t seed,
                                          double sigma) {
  const int64_t n = static_cast<int64_t>(nx) * ny * nz;
  std::vector<double> g(n), t(n);
  const uint64_t s2 = seed ^ 0x5DEECE66DULL;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double a = 1.0 - u01(s2, 2 * i);  // (0, 1]
    const double b = u01(s2, 2 * i + 1);
    g[i] = std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
  }
  if (sigma > 0.0) {
    const int rad = static_cast<int>(std::ceil(4.0 * sigma));
    std::vector<double> w(2 * rad + 1);
    for (int k = -rad; k <= rad; ++k) w[k + rad] = std::exp(-0.5 * k * k / (sigma * sigma));
    auto idx = [&](int x, int y, int z) { return (static_cast<int64_t>(z) * ny + y) * nx + x; };
    // x pass g -> t, y pass t -> g, z pass g -> t, copy back.
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          double s = 0.0;
          for (int k = -rad; k <= rad; ++k) {
            const int xx = x + k;
            if (xx >= 0 && xx < nx) s += w[k + rad] * g[idx(xx, y, z)];
          }
          t[idx(x, y, z)] = s;
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          double s = 0.0;
          for (int k = -rad; k <= rad; ++k) {
            const int yy = y + k;
            if (yy >= 0 && yy < ny) s += w[k + rad] * t[idx(x, yy, z)];
          }
          g[idx(x, y, z)] = s;
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          double s = 0.0;
          for (int k = -rad; k <= rad; ++k) {
            const int zz = z + k;
            if (zz >= 0 && zz < nz) s += w[k + rad] * g[idx(x, y, zz)];
          }
          t[idx(x, y, z)] = s;
        }
    g.swap(t);
  }
  double mean = 0.0;
  for (double v : g) mean += v;
  mean /= static_cast<double>(n);
  double var = 0.0;
  for (double v : g) var += (v - mean) * (v - mean);
  var /= static_cast<double>(n);
  const double inv = var > 0.0 ? 1.0 / std::sqrt(var) : 1.0;
  for (double &v : g) v = (v - mean) * inv;
  return g;
}

// Rows in natural order (x fastest); each row lists its columns ascending,
// as Eigen::setFromTriplets leaves them.
Csr poisson2d(int nx, int ny) {
  Csr A;
  A.rows = A.cols = static_cast<int64_t>(nx) * ny;
  A.ptr.assign(A.rows + 1, 0);
  A.col.reserve(5 * A.rows);
  A.val.reserve(5 * A.rows);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const int32_t r = y * nx + x;
      auto put = [&](int32_t c, double v) { A.col.push_back(c); A.val.push_back(v); };
      if (y > 0) put(r - nx, -1.0);
      if (x > 0) put(r - 1, -1.0);
      put(r, 4.0);
      if (x + 1 < nx) put(r + 1, -1.0);
      if (y + 1 < ny) put(r + nx, -1.0);
      A.ptr[r + 1] = static_cast<int32_t>(A.col.size());
    }
  return A;
}

Csr poisson3d(int nx, int ny, int nz) {
  Csr A;
  A.rows = A.cols = static_cast<int64_t>(nx) * ny * nz;
  A.ptr.assign(A.rows + 1, 0);
  // Row lengths first so the fill can run in parallel.
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t r = (static_cast<int64_t>(z) * ny + y) * nx + x;
        A.ptr[r + 1] = 1 + (x > 0) + (x + 1 < nx) + (y > 0) + (y + 1 < ny) + (z > 0) + (z + 1 < nz);
      }
  for (int64_t r = 0; r < A.rows; ++r) A.ptr[r + 1] += A.ptr[r];
  A.col.resize(A.ptr[A.rows]);
  A.val.resize(A.ptr[A.rows]);
  const int64_t sxy = static_cast<int64_t>(nx) * ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t r = (static_cast<int64_t>(z) * ny + y) * nx + x;
        int64_t d = A.ptr[r];
        auto put = [&](int64_t c, double v) { A.col[d] = static_cast<int32_t>(c); A.val[d] = v; ++d; };
        if (z > 0) put(r - sxy, -1.0);
        if (y > 0) put(r - nx, -1.0);
        if (x > 0) put(r - 1, -1.0);
        put(r, 6.0);
        if (x + 1 < nx) put(r + 1, -1.0);
        if (y + 1 < ny) put(r + nx, -1.0);
        if (z + 1 < nz) put(r + sxy, -1.0);
      }
  return A;
}

Csr varcoef3d(int n, uint64_t seed, double sigma, double scale,
              std::vector<double> *kfield) {
  std::vector<double> g = gaussian_random_field(n, n, n, seed, sigma);
  std::vector<double> k(g.size());
  for (size_t i = 0; i < g.size(); ++i) k[i] = std::exp(scale * g[i]);
  Csr A = poisson3d(n, n, n);  // same pattern, values replaced below
  const int64_t sxy = static_cast<int64_t>(n) * n;
  auto harm = [](double a, double b) { return 2.0 * a * b / (a + b); };
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int64_t r = (static_cast<int64_t>(z) * n + y) * n + x;
        const double kc = k[r];
        double diag = 0.0;
        int64_t d = A.ptr[r];
        auto face = [&](bool inside, int64_t c) {
          if (inside) {
            const double t = harm(kc, k[c]);
            diag += t;
            return -t;
          }
          diag += 2.0 * kc;  // Dirichlet face at half a cell
          return 0.0;
        };
        // Same neighbour order as poisson3d's fill.
        double vzm = face(z > 0, r - sxy), vym = face(y > 0, r - n), vxm = face(x > 0, r - 1);
        double vxp = face(x + 1 < n, r + 1), vyp = face(y + 1 < n, r + n), vzp = face(z + 1 < n, r + sxy);
        if (z > 0) A.val[d++] = vzm;
        if (y > 0) A.val[d++] = vym;
        if (x > 0) A.val[d++] = vxm;
        A.val[d++] = diag;
        if (x + 1 < n) A.val[d++] = vxp;
        if (y + 1 < n) A.val[d++] = vyp;
        if (z + 1 < n) A.val[d++] = vzp;
      }
  if (kfield) kfield->swap(k);
  return A;
}

// ---------------------------------------------------------------------------
// Elasticity.
// ---------------------------------------------------------------------------
namespace {

using Mat12 = std::array<double, 144>;

// Element matrix of one P1 tet (examples/elasticity/LinearElasticity.ipp:
// 239-266): H_ij = V (mu (gNj . gNi) I + mu gNj gNi^T + lambda gNi gNj^T).
Mat12 tet_matrix(const double X[4][3], double lambda, double mu) {
  double D[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) D[r][c] = X[c + 1][r] - X[0][r];
  const double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                     D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                     D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
  const double V = std::abs(det) / 6.0;
  double inv[3][3];
  inv[0][0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) / det;
  inv[0][1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) / det;
  inv[0][2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) / det;
  inv[1][0] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) / det;
  inv[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) / det;
  inv[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) / det;
  inv[2][0] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) / det;
  inv[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) / det;
  inv[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) / det;
  // grad N_{i} for i = 1..3 is row i-1 of D^-1; grad N_0 = -sum.
  double gN[4][3];
  for (int d = 0; d < 3; ++d) {
    gN[1][d] = inv[0][d];
    gN[2][d] = inv[1][d];
    gN[3][d] = inv[2][d];
    gN[0][d] = -(gN[1][d] + gN[2][d] + gN[3][d]);
  }
  Mat12 H{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const double dot = gN[j][0] * gN[i][0] + gN[j][1] * gN[i][1] + gN[j][2] * gN[i][2];
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double h = mu * gN[j][k] * gN[i][l] + lambda * gN[i][k] * gN[j][l];
          if (k == l) h += mu * dot;
          H[(3 * i + k) * 12 + 3 * j + l] = V * h;
        }
    }
  return H;
}

}  // namespace

ElasticityProblem elasticity(int nv, double E0, double nu, double hetero_sigma,
                             double hetero_scale, uint64_t seed) {
  if (nv < 2) throw std::runtime_error("elasticity: need at least 2 vertices per side.");
  const double h = 1.0 / (nv - 1);
  const int nc = nv - 1;  // hexes per side
  const int64_t nverts = static_cast<int64_t>(nv) * nv * nv;
  auto vid = [&](int x, int y, int z) { return (static_cast<int64_t>(z) * nv + y) * nv + x; };
  // Free vertex numbering: vertices with x > 0, in vertex id order.
  std::vector<int64_t> fidx(nverts, -1);
  int64_t nfree = 0;
  for (int z = 0; z < nv; ++z)
    for (int y = 0; y < nv; ++y)
      for (int x = 0; x < nv; ++x)
        if (x > 0) fidx[vid(x, y, z)] = nfree++;
  // The 15-point vertex stencil of the Kuhn split, ordered by linear offset
  // (z slowest) so that rows come out with ascending columns.
  static const int offs[15][3] = {
      {-1, -1, -1}, {0, -1, -1}, {-1, 0, -1}, {0, 0, -1},  // z-1
      {-1, -1, 0},  {0, -1, 0},  {-1, 0, 0},  {0, 0, 0},   {1, 0, 0},
      {0, 1, 0},    {1, 1, 0},                               // z
      {0, 0, 1},    {1, 0, 1},   {0, 1, 1},   {1, 1, 1}};    // z+1
  // Kuhn tets: v0 = corner 000, then walk the axes in permutation order.
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                  {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::array<std::array<int, 4>, 6> tet_corner{};  // corner code bits x|y<<1|z<<2
  for (int t = 0; t < 6; ++t) {
    int c = 0;
    tet_corner[t][0] = 0;
    for (int s = 0; s < 3; ++s) {
      c |= 1 << perms[t][s];
      tet_corner[t][s + 1] = c;
    }
  }
  const double lam1 = nu / ((1 + nu) * (1 - 2 * nu));  // per unit E
  const double mu1 = 1.0 / (2 * (1 + nu));
  std::array<Mat12, 6> Kbase;
  for (int t = 0; t < 6; ++t) {
    double X[4][3];
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) X[a][d] = ((tet_corner[t][a] >> d) & 1) * h;
    Kbase[t] = tet_matrix(X, lam1, mu1);
  }
  std::vector<double> Ecell;
  if (hetero_sigma > 0.0) {
    std::vector<double> g = gaussian_random_field(nc, nc, nc, seed, hetero_sigma);
    Ecell.resize(g.size());
    for (size_t i = 0; i < g.size(); ++i) Ecell[i] = E0 * std::exp(hetero_scale * g[i]);
  }
  ElasticityProblem P;
  P.free_vertices = nfree;
  P.tets = static_cast<int64_t>(nc) * nc * nc * 6;
  // Row structure: for each free vertex, the free neighbours among the 15
  // stencil points (all of them couple through at least one tet).
  std::vector<int16_t> slot_of(static_cast<size_t>(nfree) * 15, -1);
  std::vector<int32_t> bptr(nfree + 1, 0);
  for (int z = 0; z < nv; ++z)
    for (int y = 0; y < nv; ++y)
      for (int x = 1; x < nv; ++x) {
        const int64_t f = fidx[vid(x, y, z)];
        int cnt = 0;
        for (int o = 0; o < 15; ++o) {
          const int xx = x + offs[o][0], yy = y + offs[o][1], zz = z + offs[o][2];
          if (xx < 1 || yy < 0 || zz < 0 || xx >= nv || yy >= nv || zz >= nv) continue;
          slot_of[f * 15 + o] = static_cast<int16_t>(cnt++);
        }
        bptr[f + 1] = cnt;
      }
  for (int64_t f = 0; f < nfree; ++f) bptr[f + 1] += bptr[f];
  const int64_t nblocks = bptr[nfree];
  std::vector<double> bval(static_cast<size_t>(nblocks) * 9, 0.0);
  std::vector<int32_t> bcol(nblocks);
  for (int z = 0; z < nv; ++z)
    for (int y = 0; y < nv; ++y)
      for (int x = 1; x < nv; ++x) {
        const int64_t f = fidx[vid(x, y, z)];
        for (int o = 0; o < 15; ++o) {
          const int s = slot_of[f * 15 + o];
          if (s < 0) continue;
          bcol[bptr[f] + s] = static_cast<int32_t>(
              fidx[vid(x + offs[o][0], y + offs[o][1], z + offs[o][2])]);
        }
      }
  auto offset_index = [&](int dx, int dy, int dz) {
    for (int o = 0; o < 15; ++o)
      if (offs[o][0] == dx && offs[o][1] == dy && offs[o][2] == dz) return o;
    return -1;
  };
  int pair_off[6][4][4];
  for (int t = 0; t < 6; ++t)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        const int ca = tet_corner[t][a], cb = tet_corner[t][b];
        pair_off[t][a][b] = offset_index(((cb >> 0) & 1) - ((ca >> 0) & 1),
                                         ((cb >> 1) & 1) - ((ca >> 1) & 1),
                                         ((cb >> 2) & 1) - ((ca >> 2) & 1));
      }
  std::vector<double> mass(nverts, 0.0);
  const double Vtet = h * h * h / 6.0;
  // Assemble by z-slab of hexes; two passes over even / odd slabs keep
  // threads from writing the same vertex rows.
  for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int cz = parity; cz < nc; cz += 2) {
      for (int cy = 0; cy < nc; ++cy)
        for (int cx = 0; cx < nc; ++cx) {
          const double E = Ecell.empty() ? E0 : Ecell[(static_cast<int64_t>(cz) * nc + cy) * nc + cx];
          for (int t = 0; t < 6; ++t) {
            int64_t v[4];
            for (int a = 0; a < 4; ++a) {
              const int c = tet_corner[t][a];
              v[a] = vid(cx + (c & 1), cy + ((c >> 1) & 1), cz + ((c >> 2) & 1));
              mass[v[a]] += Vtet / 4.0;
            }
            const Mat12 &K = Kbase[t];
            for (int a = 0; a < 4; ++a) {
              const int64_t fa = fidx[v[a]];
              if (fa < 0) continue;
              for (int b = 0; b < 4; ++b) {
                if (fidx[v[b]] < 0) continue;
                const int s = slot_of[fa * 15 + pair_off[t][a][b]];
                double *blk = &bval[(static_cast<size_t>(bptr[fa]) + s) * 9];
                for (int k = 0; k < 3; ++k)
                  for (int l = 0; l < 3; ++l) blk[3 * k + l] += E * K[(3 * a + k) * 12 + 3 * b + l];
              }
            }
          }
        }
    }
  }
  // Scalar CSR: row 3f+k lists blocks in order, 3 columns each.
  Csr &A = P.A;
  A.rows = A.cols = 3 * nfree;
  A.ptr.assign(A.rows + 1, 0);
  for (int64_t f = 0; f < nfree; ++f)
    for (int k = 0; k < 3; ++k)
      A.ptr[3 * f + k + 1] = 3 * (bptr[f + 1] - bptr[f]);
  for (int64_t r = 0; r < A.rows; ++r) A.ptr[r + 1] += A.ptr[r];
  A.col.resize(A.ptr[A.rows]);
  A.val.resize(A.ptr[A.rows]);
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nfree; ++f)
    for (int k = 0; k < 3; ++k) {
      int64_t d = A.ptr[3 * f + k];
      for (int32_t b = bptr[f]; b < bptr[f + 1]; ++b)
        for (int l = 0; l < 3; ++l) {
          A.col[d] = 3 * bcol[b] + l;
          A.val[d] = bval[static_cast<size_t>(b) * 9 + 3 * k + l];
          ++d;
        }
    }
  // Body force: lumped mass times (0, 0, -1) plus U[-0.1, 0.1] per DOF.
  P.rhs.assign(A.rows, 0.0);
  uniform_pm1(seed ^ 0xB0D1F0ACEULL, A.rows, P.rhs.data());
  for (auto &v : P.rhs) v *= 0.1;
  P.nullspace.assign(static_cast<size_t>(A.rows) * 6, 0.0);
  for (int z = 0; z < nv; ++z)
    for (int y = 0; y < nv; ++y)
      for (int x = 1; x < nv; ++x) {
        const int64_t vv = vid(x, y, z);
        const int64_t f = fidx[vv];
        P.rhs[3 * f + 2] -= mass[vv];
        const double X = x * h - 0.5, Y = y * h - 0.5, Z = z * h - 0.5;
        double *B = &P.nullspace[static_cast<size_t>(3 * f) * 6];
        // rows 3f, 3f+1, 3f+2; columns: tx, ty, tz, rx, ry, rz
        B[0 * 6 + 0] = 1.0;
        B[1 * 6 + 1] = 1.0;
        B[2 * 6 + 2] = 1.0;
        B[1 * 6 + 3] = -Z; B[2 * 6 + 3] = Y;
        B[0 * 6 + 4] = Z;  B[2 * 6 + 4] = -X;
        B[0 * 6 + 5] = -Y; B[1 * 6 + 5] = X;
      }
  return P;
}

// examples/elasticity/cantileverBeam.cc:275-395 on a caller's tet mesh:
// lambda and mu from E and nu as there, element matrices as
// LinearElasticity.ipp:239-266 (a tet's matrix does not depend on its
// vertex order, so the reference's orientation fix at :53-56 changes
// nothing), vertices with x <= x_min + clamp_rel (x_max - x_min) clamped
// and eliminated (constrainSimulation, :379-413; the free vertices keep
// their input order), duplicates summed in element order (:325-348), and
// rhs = density * lumped mass * (0, 0, gravity) per free vertex (:20-79,
// cantileverBeam.cc:341-348). The rigid-body modes are taken about the
// bounding-box centre. A zero-volume tet is an error here (the reference
// would divide by its zero determinant).
ElasticityProblem elasticity_mesh(int64_t nverts, const double *V, int64_t ntets,
                                  const int32_t *T, double young, double nu,
                                  const double *young_per_tet, double clamp_rel,
                                  double density, double gravity) {
  if (nverts <= 0 || ntets <= 0) throw std::runtime_error("elasticity mesh: empty mesh.");
  if (nverts > INT32_MAX / 3) throw std::runtime_error("elasticity mesh: too many vertices.");
  for (int64_t i = 0; i < 4 * ntets; ++i)
    if (T[i] < 0 || T[i] >= nverts)
      throw std::runtime_error("elasticity mesh: tetrahedron vertex index out of range.");
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) lo[d] = hi[d] = V[d];
  for (int64_t v = 0; v < nverts; ++v)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], V[3 * v + d]);
      hi[d] = std::max(hi[d], V[3 * v + d]);
    }
  const double clamp_x = lo[0] + clamp_rel * (hi[0] - lo[0]);
  std::vector<int64_t> fidx(nverts, -1);
  int64_t nfree = 0;
  for (int64_t v = 0; v < nverts; ++v)
    if (!(V[3 * v] <= clamp_x)) fidx[v] = nfree++;
  if (nfree == nverts) throw std::runtime_error("No vertices found on clamped side.");
  if (nfree == 0) throw std::runtime_error("elasticity mesh: every vertex is clamped.");
  // Block pattern: every pair of free vertices sharing a tet.
  std::vector<uint64_t> keys;
  keys.reserve(static_cast<size_t>(ntets) * 16);
  for (int64_t t = 0; t < ntets; ++t)
    for (int a = 0; a < 4; ++a) {
      const int64_t fa = fidx[T[4 * t + a]];
      if (fa < 0) continue;
      for (int b = 0; b < 4; ++b) {
        const int64_t fb = fidx[T[4 * t + b]];
        if (fb >= 0) keys.push_back((static_cast<uint64_t>(fa) << 32) | static_cast<uint64_t>(fb));
      }
    }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  if (keys.size() * 9 > static_cast<size_t>(INT32_MAX))
    throw std::runtime_error("elasticity mesh: more than 2^31 nonzeros.");
  std::vector<int32_t> bptr(nfree + 1, 0), bcol(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) {
    ++bptr[(keys[i] >> 32) + 1];
    bcol[i] = static_cast<int32_t>(keys[i] & 0xFFFFFFFFu);
  }
  for (int64_t f = 0; f < nfree; ++f) bptr[f + 1] += bptr[f];
  std::vector<double> bval(keys.size() * 9, 0.0), mass(nverts, 0.0);
  const double mu0 = young / (2.0 * (1.0 + nu));
  const double lam0 = young * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  for (int64_t t = 0; t < ntets; ++t) {
    double X[4][3];
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) X[a][d] = V[3 * static_cast<int64_t>(T[4 * t + a]) + d];
    double lam = lam0, mu = mu0;
    if (young_per_tet) {
      const double E = young_per_tet[t];
      mu = E / (2.0 * (1.0 + nu));
      lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    }
    double e[3][3];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) e[c][d] = X[c + 1][d] - X[0][d];
    const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                       e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                       e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
    if (det == 0.0 || !std::isfinite(det))
      throw std::runtime_error("elasticity mesh: degenerate tetrahedron.");
    // Lumped mass: a quarter of the signed volume per vertex; inverted tets
    // (negative volume) add none, as the reference skips them
    // (LinearElasticity.ipp:31-47).
    if (det > 0.0)
      for (int a = 0; a < 4; ++a) mass[T[4 * t + a]] += det / 6.0 / 4.0;
    const Mat12 H = tet_matrix(X, lam, mu);
    for (int a = 0; a < 4; ++a) {
      const int64_t fa = fidx[T[4 * t + a]];
      if (fa < 0) continue;
      for (int b = 0; b < 4; ++b) {
        const int64_t fb = fidx[T[4 * t + b]];
        if (fb < 0) continue;
        const int32_t *lo_c = bcol.data() + bptr[fa], *hi_c = bcol.data() + bptr[fa + 1];
        const int64_t slot = std::lower_bound(lo_c, hi_c, static_cast<int32_t>(fb)) - bcol.data();
        double *blk = &bval[static_cast<size_t>(slot) * 9];
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) blk[3 * k + l] += H[(3 * a + k) * 12 + 3 * b + l];
      }
    }
  }
  ElasticityProblem P;
  P.free_vertices = nfree;
  P.tets = ntets;
  Csr &A = P.A;
  A.rows = A.cols = 3 * nfree;
  A.ptr.assign(A.rows + 1, 0);
  for (int64_t f = 0; f < nfree; ++f)
    for (int k = 0; k < 3; ++k) A.ptr[3 * f + k + 1] = 3 * (bptr[f + 1] - bptr[f]);
  for (int64_t r = 0; r < A.rows; ++r) A.ptr[r + 1] += A.ptr[r];
  A.col.resize(A.ptr[A.rows]);
  A.val.resize(A.ptr[A.rows]);
  for (int64_t f = 0; f < nfree; ++f)
    for (int k = 0; k < 3; ++k) {
      int64_t d = A.ptr[3 * f + k];
      for (int32_t b = bptr[f]; b < bptr[f + 1]; ++b)
        for (int l = 0; l < 3; ++l) {
          A.col[d] = 3 * bcol[b] + l;
          A.val[d] = bval[static_cast<size_t>(b) * 9 + 3 * k + l];
          ++d;
        }
    }
  P.rhs.assign(A.rows, 0.0);
  P.nullspace.assign(static_cast<size_t>(A.rows) * 6, 0.0);
  const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
  for (int64_t v = 0; v < nverts; ++v) {
    const int64_t f = fidx[v];
    if (f < 0) continue;
    P.rhs[3 * f + 2] = density * mass[v] * gravity;
    const double X = V[3 * v] - cx, Y = V[3 * v + 1] - cy, Z = V[3 * v + 2] - cz;
    double *B = &P.nullspace[static_cast<size_t>(3 * f) * 6];
    B[0 * 6 + 0] = 1.0;
    B[1 * 6 + 1] = 1.0;
    B[2 * 6 + 2] = 1.0;
    B[1 * 6 + 3] = -Z; B[2 * 6 + 3] = Y;
    B[0 * 6 + 4] = Z;  B[2 * 6 + 4] = -X;
    B[0 * 6 + 5] = -Y; B[1 * 6 + 5] = X;
  }
  return P;
}

}  // namespace gsx
