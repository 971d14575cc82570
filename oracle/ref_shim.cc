// C entry points over the reference's own Eigen-free setup sources
// (src/GraphColoring.cc, src/Permutation.cc,
// src/ReorderingStrategies/RCMReordering.cc), compiled from where they lie
// in the reference tree by oracle/Makefile into oracle/_ref/. Test
// infrastructure only: the tests use it to pin the oracle's and the
// product's colouring / RCM / permutation restatements bit for bit.
#include <GPUSolver/GraphColoring.h>
#include <GPUSolver/Permutation.h>
#include <GPUSolver/ReorderingStrategies/RCMReordering.h>

#include <cstdint>
#include <cstring>
#include <exception>

extern "C" {

int ref_coloring(int64_t n, const int32_t *ptr, const int32_t *col, int32_t *colors,
                 int32_t *vertices, int32_t *offsets, int32_t *ncolors) {
  try {
    gpuSolver::GraphColoring c;
    c.compute(static_cast<std::size_t>(n), ptr, col);
    std::memcpy(colors, c.colors().data(), n * sizeof(int32_t));
    std::memcpy(vertices, c.colorVertices().data(), n * sizeof(int32_t));
    std::memcpy(offsets, c.colorOffsets().data(), c.colorOffsets().size() * sizeof(int32_t));
    *ncolors = static_cast<int32_t>(c.numberOfColors());
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

int ref_rcm(int64_t n, const int32_t *ptr, const int32_t *col, int32_t *o2n, int32_t *n2o) {
  try {
    gpuSolver::RCMReordering r;
    r.compute(static_cast<std::size_t>(n), ptr, col, o2n, n2o);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

int ref_permutation_valid(int64_t n, const int32_t *o2n, const int32_t *n2o) {
  try {
    gpuSolver::Permutation p(static_cast<std::size_t>(n), o2n, n2o);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
}
