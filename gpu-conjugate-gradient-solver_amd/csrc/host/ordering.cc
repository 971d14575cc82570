// Greedy multicolouring, reverse Cuthill-McKee, hierarchy-aware
// reordering, block graphs and block permutation expansion.
#include "setup.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <queue>
#include <stdexcept>

namespace gsx {

// src/GraphColoring.cc:9-72: first-fit in vertex order; a colour is
// forbidden for vertex i when an already coloured neighbour holds it. The
// forbidden table is stamped with the vertex index, so it never needs
// clearing.
Coloring greedy_coloring(int64_t n, const int32_t *ptr, const int32_t *col) {
  Coloring out;
  out.colors.assign(n, -1);
  std::vector<int64_t> stamp(n + 1, -1);
  int32_t nc = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int32_t j = col[k];
      if (j != i && out.colors[j] >= 0) stamp[out.colors[j]] = i;
    }
    int32_t c = 0;
    while (stamp[c] == i) ++c;
    out.colors[i] = c;
    nc = std::max(nc, c + 1);
  }
  out.ncolors = nc;
  out.offsets.assign(nc + 1, 0);
  for (int64_t v = 0; v < n; ++v) out.offsets[out.colors[v] + 1]++;
  for (int32_t c = 0; c < nc; ++c) out.offsets[c + 1] += out.offsets[c];
  out.vertices.assign(n, -1);
  std::vector<int32_t> next(out.offsets.begin(), out.offsets.end() - 1);
  for (int64_t v = 0; v < n; ++v)
    out.vertices[next[out.colors[v]]++] = static_cast<int32_t>(v);
  return out;
}

// src/ReorderingStrategies/RCMReordering.cc:11-110. Degree counts the
// off-diagonal entries of a row; each connected component starts at the
// lowest-index unvisited vertex of minimum degree; neighbours are visited in
// std::sort order of degree (the same comparator and input order as the
// reference, so libstdc++ breaks ties the same way); the Cuthill-McKee
// order is reversed at the end.
Perm rcm(int64_t n, const int32_t *ptr, const int32_t *col) {
  std::vector<int> degree(n);
  std::vector<char> visited(n, 0);
  for (int64_t r = 0; r < n; ++r) {
    int d = 0;
    for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k)
      if (col[k] != r) ++d;
    degree[r] = d;
  }
  std::vector<int32_t> cm;
  cm.reserve(n);
  std::queue<int32_t> q;
  std::vector<int> nb;
  const int *deg = degree.data();
  while (static_cast<int64_t>(cm.size()) < n) {
    int32_t start = -1;
    int best = std::numeric_limits<int>::max();
    for (int64_t r = 0; r < n; ++r)
      if (!visited[r] && degree[r] < best) {
        best = degree[r];
        start = static_cast<int32_t>(r);
      }
    q.push(start);
    visited[start] = 1;
    while (!q.empty()) {
      const int32_t v = q.front();
      q.pop();
      cm.push_back(v);
      nb.clear();
      for (int32_t k = ptr[v]; k < ptr[v + 1]; ++k)
        if (col[k] != v) nb.push_back(col[k]);
      std::sort(nb.begin(), nb.end(), [deg](int a, int b) { return deg[a] < deg[b]; });
      for (int w : nb)
        if (!visited[w]) {
          visited[w] = 1;
          q.push(w);
        }
    }
  }
  Perm p;
  p.new_to_old.resize(n);
  p.old_to_new.resize(n);
  for (int64_t i = 0; i < n; ++i) p.new_to_old[i] = cm[n - 1 - i];
  for (int64_t i = 0; i < n; ++i) p.old_to_new[p.new_to_old[i]] = static_cast<int32_t>(i);
  return p;
}

// Strongest coarse parent of every fine DOF
// (src/ReorderingStrategies/HierarchicalReordering.cc:13-52): the coarse row
// whose entry in the DOF's column has the largest |value| in FP32 (the
// reference's hierarchy type), the first such entry in row-major order on a
// tie, and -1 for an empty or all-NaN column. Every entry bids a 64-bit key,
// |value|'s bits above the complement of its position, and one atomic max
// per column keeps the winner whatever order the threads run in; the
// winning position then maps back to its row.
static std::vector<int32_t> strongest_parents(const Csr &R) {
  const int64_t nf = R.cols;
  std::unique_ptr<std::atomic<uint64_t>[]> key(new std::atomic<uint64_t>[nf]);
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nf; ++f) key[f].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t c = 0; c < R.rows; ++c)
    for (int32_t e = R.ptr[c]; e < R.ptr[c + 1]; ++e) {
      const float w = std::fabs(static_cast<float>(R.val[e]));
      if (std::isnan(w)) continue;
      uint32_t bits;
      std::memcpy(&bits, &w, sizeof bits);
      const uint64_t k = (static_cast<uint64_t>(bits) << 32) | (0xFFFFFFFFu - static_cast<uint32_t>(e));
      std::atomic<uint64_t> &slot = key[R.col[e]];
      uint64_t cur = slot.load(std::memory_order_relaxed);
      while (k > cur && !slot.compare_exchange_weak(cur, k, std::memory_order_relaxed)) {
      }
    }
  std::vector<int32_t> parent(nf);
  const int32_t *pb = R.ptr.data(), *pe = pb + R.rows + 1;
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nf; ++f) {
    const uint64_t k = key[f].load(std::memory_order_relaxed);
    if (k == 0) {
      parent[f] = -1;
      continue;
    }
    const int32_t e = static_cast<int32_t>(0xFFFFFFFFu - static_cast<uint32_t>(k));
    parent[f] = static_cast<int32_t>(std::upper_bound(pb, pe, e) - pb - 1);
  }
  return parent;
}

// src/ReorderingStrategies/HierarchicalReordering.cc:217-278 (driver) and
// :74-200 (one level). The coarsest level takes its RCM order; each finer
// level is grouped by the final rank of its DOFs' strongest parents, DOFs
// without a parent last, and ordered within a group by the level's own RCM
// rank. The reference sorts on that (parent rank, RCM rank) key, which is a
// total order; a counting sort over the parent ranks, fed in RCM order, gives
// the same permutation in linear time.
std::vector<Perm> hierarchical_reordering(const Hierarchy &H, std::vector<int64_t> *orphans) {
  const int64_t L = static_cast<int64_t>(H.A.size());
  if (L == 0) throw std::runtime_error("HierarchicalReordering: empty hierarchy.");
  if (static_cast<int64_t>(H.R.size()) + 1 != L)
    throw std::runtime_error(
        "HierarchicalReordering: expected one restriction matrix between consecutive levels.");
  for (int64_t k = L - 2; k >= 0; --k) {
    if (H.A[k].rows != H.A[k].cols)
      throw std::runtime_error("HierarchicalReordering: fine matrix must be square.");
    if (H.R[k].cols != H.A[k].rows)
      throw std::runtime_error("HierarchicalReordering: restriction dimensions do not match fine matrix.");
    // No counterpart in the reference, which indexes the coarse permutation
    // with R's row numbers unchecked.
    if (H.R[k].rows != H.A[k + 1].rows)
      throw std::runtime_error("HierarchicalReordering: restriction rows do not match the coarse level.");
  }
  // No counterpart either: the reference runs RCM on a non-square coarsest
  // matrix and reads past its degree table.
  if (H.A[L - 1].rows != H.A[L - 1].cols)
    throw std::runtime_error("HierarchicalReordering: coarsest matrix must be square.");
  // Every level's own RCM order; the levels are independent.
  std::vector<Perm> own(L);
  std::vector<std::string> err(L);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t k = 0; k < L; ++k) {
    try {
      own[k] = rcm(H.A[k].rows, H.A[k].ptr.data(), H.A[k].col.data());
      own[k].check();
    } catch (const std::exception &e) {
      err[k] = e.what();
    }
  }
  for (int64_t k = L - 1; k >= 0; --k)
    if (!err[k].empty()) throw std::runtime_error(err[k]);
  std::vector<Perm> out(L);
  out[L - 1] = own[L - 1];
  if (orphans) orphans->assign(L - 1, 0);
  for (int64_t k = L - 2; k >= 0; --k) {
    const int64_t nf = H.A[k].rows, nc = H.A[k + 1].rows;
    const std::vector<int32_t> parent = strongest_parents(H.R[k]);
    const int32_t *crank = out[k + 1].old_to_new.data();
    std::vector<int64_t> start(nc + 2, 0);
    for (int64_t f = 0; f < nf; ++f) ++start[(parent[f] >= 0 ? crank[parent[f]] : nc) + 1];
    if (orphans) (*orphans)[k] = start[nc + 1];
    for (int64_t r = 0; r <= nc; ++r) start[r + 1] += start[r];
    Perm &p = out[k];
    p.new_to_old.resize(nf);
    p.old_to_new.resize(nf);
    for (int64_t i = 0; i < nf; ++i) {
      const int32_t f = own[k].new_to_old[i];
      const int64_t r = parent[f] >= 0 ? crank[parent[f]] : nc;
      const int64_t pos = start[r]++;
      p.new_to_old[pos] = f;
      p.old_to_new[f] = static_cast<int32_t>(pos);
    }
  }
  return out;
}

// src/BlockUtils.cc:126-216: every scalar entry coupling two different
// blocks adds the edge in both directions; lists are sorted and
// de-duplicated.
Graph block_graph(const Csr &A, int32_t bs) {
  if (A.rows != A.cols) throw std::runtime_error("buildBlockGraph: matrix must be square.");
  if (A.rows % bs != 0)
    throw std::runtime_error("buildBlockGraph: matrix size is not divisible by block size.");
  const int64_t nb = A.rows / bs;
  // The symmetrised block adjacency without self loops, each row sorted and
  // unique: out-neighbours per block row (R), their transpose (T), and the
  // merge of both.
  auto gather = [&](int64_t br, std::vector<int32_t> &tmp) {
    tmp.clear();
    for (int64_t r = br * bs; r < (br + 1) * bs; ++r)
      for (int32_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
        const int32_t bc = A.col[k] / bs;
        if (bc != br) tmp.push_back(bc);
      }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
  };
  std::vector<int64_t> rp(nb + 1, 0);
#pragma omp parallel
  {
    std::vector<int32_t> tmp;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t br = 0; br < nb; ++br) {
      gather(br, tmp);
      rp[br + 1] = static_cast<int64_t>(tmp.size());
    }
  }
  for (int64_t b = 0; b < nb; ++b) rp[b + 1] += rp[b];
  std::vector<int32_t> rc(rp[nb]);
#pragma omp parallel
  {
    std::vector<int32_t> tmp;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t br = 0; br < nb; ++br) {
      gather(br, tmp);
      std::copy(tmp.begin(), tmp.end(), rc.begin() + rp[br]);
    }
  }
  // Transpose (rows come out sorted: sources visited in increasing order).
  std::vector<int64_t> tp(nb + 1, 0);
  for (int32_t c : rc) tp[c + 1]++;
  for (int64_t b = 0; b < nb; ++b) tp[b + 1] += tp[b];
  std::vector<int32_t> tc(rp[nb]);
  {
    std::vector<int64_t> next(tp.begin(), tp.end() - 1);
    for (int64_t br = 0; br < nb; ++br)
      for (int64_t k = rp[br]; k < rp[br + 1]; ++k) tc[next[rc[k]]++] = static_cast<int32_t>(br);
  }
  // Merge R and T row by row.
  auto merged = [&](int64_t b, int32_t *out) {
    int64_t i = rp[b], j = tp[b], n = 0;
    int32_t last = -1;
    while (i < rp[b + 1] || j < tp[b + 1]) {
      int32_t v;
      if (j >= tp[b + 1] || (i < rp[b + 1] && rc[i] <= tc[j])) v = rc[i++];
      else v = tc[j++];
      if (v != last) {
        if (out) out[n] = v;
        ++n;
        last = v;
      }
    }
    return n;
  };
  Graph g;
  std::vector<int64_t> gp(nb + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t b = 0; b < nb; ++b) gp[b + 1] = merged(b, nullptr);
  for (int64_t b = 0; b < nb; ++b) gp[b + 1] += gp[b];
  if (gp[nb] > INT32_MAX) throw std::runtime_error("buildBlockGraph: too many edges.");
  g.ptr.resize(nb + 1);
  for (int64_t b = 0; b <= nb; ++b) g.ptr[b] = static_cast<int32_t>(gp[b]);
  g.col.resize(gp[nb]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t b = 0; b < nb; ++b) merged(b, g.col.data() + gp[b]);
  return g;
}

// src/BlockUtils.cc:7-55.
Perm expand_block_perm(const Perm &np, int64_t nodes, int32_t bs) {
  Perm p;
  p.old_to_new.resize(nodes * bs);
  p.new_to_old.resize(nodes * bs);
  for (int64_t o = 0; o < nodes; ++o)
    for (int32_t c = 0; c < bs; ++c)
      p.old_to_new[bs * o + c] = bs * np.old_to_new[o] + c;
  for (int64_t nw = 0; nw < nodes; ++nw)
    for (int32_t c = 0; c < bs; ++c)
      p.new_to_old[bs * nw + c] = bs * np.new_to_old[nw] + c;
  p.check();
  return p;
}

}  // namespace gsx
