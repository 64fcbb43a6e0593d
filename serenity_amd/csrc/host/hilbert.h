// Hilbert-curve key of one grid point, shared by the host sort (grid.cpp) and
// the device grid build (kernels/grid_kernels.hip) so that both give the same
// order.  The points' bounding box is cut into length^3 cells (length a power
// of 2); the key walks the refinement levels of the cell containing the point
// with the child-cube numbering and transition tables of the Hilbert R-tree
// ordering (HilbertRTreeSorting.cpp:54-100).  Points are then taken in
// descending key order, equal keys keeping their input order.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define SXC_HILB_HD __host__ __device__
#else
#define SXC_HILB_HD
#endif

namespace sxc {

SXC_HILB_HD inline int64_t hilbert_key(const double* p, const double* mn, const double* spread, int64_t length) {
#pragma clang fp contract(off)
  // child-cube numbering and the transition table between refinement levels
  const int trans[8][8] = {{0, 7, 6, 1, 2, 5, 4, 3}, {0, 3, 4, 6, 7, 5, 2, 1}, {0, 3, 4, 6, 7, 5, 2, 1},
                           {2, 3, 0, 1, 6, 7, 4, 5}, {2, 3, 0, 1, 6, 7, 4, 5}, {6, 5, 2, 1, 0, 3, 4, 7},
                           {6, 5, 2, 1, 0, 3, 4, 7}, {4, 3, 2, 5, 6, 1, 0, 7}};
  const int val[2][2][2] = {{{5, 6}, {4, 7}}, {{2, 1}, {3, 0}}};
  int64_t pt[3];
  for (int d = 0; d < 3; ++d) pt[d] = (int64_t)(int)((p[d] - mn[d]) * spread[d]);
  int64_t l = length / 2;
  int gx = pt[0] > l, gy = pt[1] > l, gz = pt[2] > l;
  int v = val[gx][gy][gz];
  int64_t idx = v;
  while (l > 1) {
    idx *= 8;
    if (gx) pt[0] -= l;
    if (gy) pt[1] -= l;
    if (gz) pt[2] -= l;
    l /= 2;
    gx = pt[0] > l;
    gy = pt[1] > l;
    gz = pt[2] > l;
    v = trans[v][val[gx][gy][gz]];
    idx += v;
  }
  return idx;
}

}  // namespace sxc
