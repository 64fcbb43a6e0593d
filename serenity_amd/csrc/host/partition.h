// Fuzzy-cell partition weight of one molecular-grid point, shared by the host
// loop (grid.cpp) and the GPU kernel (kernels/grid_kernels.hip), so that both
// evaluate the same arithmetic: Becke cells with Becke's atomic size adjustment
// and k-fold smoothing, Voronoi cells, and the Stratmann-Scuseria-Frisch
// partition with a = 0.64 (GridFactory.cpp:209-262).  FP contraction is off in
// these functions on both targets.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define SXC_PART_HD __host__ __device__
#else
#define SXC_PART_HD
#endif

namespace sxc {

enum { kPartBecke = 0, kPartVoronoi = 1, kPartSSF = 2 };

SXC_PART_HD inline double part_dist(const double* a, const double* b) {
#pragma clang fp contract(off)
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Becke's cell function s(nu) = (1 - p_k(nu)) / 2, p_k = p applied k times,
// p(nu) = (3 - nu^2) nu / 2 (applied once for k <= 1).
SXC_PART_HD inline double becke_cell(double nu, unsigned k) {
#pragma clang fp contract(off)
  const unsigned it = k > 1 ? k : 1;
  for (unsigned i = 0; i < it; ++i) nu = (3.0 - nu * nu) * nu / 2.0;
  return 0.5 * (1 - nu);
}

// Weight of the point gp (raw weight w) of atom k after the partition, or -1
// when the SSF screening drops the point.  xyz: 3 x n atom positions;
// dist: n x n interatomic distances; apair: n x n Becke size adjustments
// (zero for SSF); min_dist_k: distance from atom k to its nearest neighbour.
SXC_PART_HD inline double partition_weight(int type, unsigned smoothing, int n, const double* xyz, const double* dist,
                                           const double* apair, double min_dist_k, int k, const double* gp,
                                           double w) {
#pragma clang fp contract(off)
  double weight = w;
  double cellSum = 0.0;
  if (type == kPartBecke) {
    for (int l = 0; l < n; ++l) {
      if (!(dist[l + (size_t)n * k] < 40.0)) continue;
      const double rl = part_dist(xyz + 3 * l, gp);
      double cell = 1.0;
      for (int j = 0; j < n; ++j) {
        if (!(dist[j + (size_t)n * k] < 40.0) || l == j) continue;
        const double mu = (rl - part_dist(xyz + 3 * j, gp)) / dist[l + (size_t)n * j];
        const double nu = mu + apair[j + (size_t)n * l] * (1.0 - mu * mu);
        cell *= becke_cell(nu, smoothing);
      }
      if (l == k) weight *= cell;
      cellSum += cell;
    }
    weight /= cellSum;
  } else if (type == kPartVoronoi) {
    const double rk = part_dist(xyz + 3 * k, gp);
    for (int l = 0; l < n; ++l) {
      if (!(dist[l + (size_t)n * k] < 40.0)) continue;
      const double mu = (part_dist(xyz + 3 * l, gp) - rk) / dist[l + (size_t)n * k];
      const double nu = mu + apair[k + (size_t)n * l] * (1.0 - mu * mu);
      if (nu < 0.0) {
        weight = 0.0;
        break;
      }
    }
  } else {
    const double rk = part_dist(xyz + 3 * k, gp);
    if (rk >= 0.5 * (1.0 - 0.64) * min_dist_k) {
      for (int j = 0; j < n; ++j) {
        const double nu = (rk - part_dist(xyz + 3 * j, gp)) / dist[k + (size_t)n * j];
        if (nu >= 0.64) return -1.0;
      }
      for (int a = 0; a < n; ++a) {
        const double ra = part_dist(xyz + 3 * a, gp);
        double cell = 1.0;
        for (int j = 0; j < n; ++j) {
          if (a == j) continue;
          double nu = (ra - part_dist(xyz + 3 * j, gp)) / dist[a + (size_t)n * j];
          if (nu <= -0.64) continue;
          if (nu >= 0.64) {
            cell = 0.0;
            break;
          }
          nu /= 0.64;
          const double n3 = nu * nu * nu;
          const double poly = (-5.0 * n3 * n3 * nu + 21.0 * n3 * nu * nu - 35.0 * n3 + 35.0 * nu) / 16.0;
          cell *= 0.5 * (1 - poly);
        }
        if (a == k) weight *= cell;
        cellSum += cell;
      }
      weight /= cellSum;
    }
  }
  return weight;
}

}  // namespace sxc
