// Molecular integration grid: Lebedev-Laikov angular rules, radial
// quadratures, fuzzy-cell (Becke / SSF / Voronoi) partitioning, weight cut and
// Hilbert-curve ordering.  A restatement of Serenity's grid construction
// (src/grid/construction/AtomGridFactory.cpp:78-290, GridFactory.cpp:51-347,
// src/grid/HilbertRTreeSorting.cpp:27-216) written for this library.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

#include "host.h"
#include "hilbert.h"
#include "partition.h"

namespace sxc {

namespace {

struct LebedevOrbit {
  int code;
  double a, b, v;
};
#include "lebedev_generators.inc"

// Per-element parameters, index = nuclear charge (entry 0 is a dummy).
// Ahlrichs radial scaling (AtomGridFactory.cpp:39-52, Treutler & Ahlrichs,
// JCP 102 (1995) 346).
const double kAhlrichsAlpha[] = {1.1, 0.8, 0.9, 1.8, 1.4, 1.3, 1.1, 0.9, 0.9, 0.9, 0.9, 1.4, 1.3,
                                 1.3, 1.2, 1.1, 1.0, 1.0, 1.0, 1.5, 1.4, 1.3, 1.2, 1.2, 1.2, 1.2,
                                 1.2, 1.2, 1.1, 1.1, 1.1, 1.1, 1.0, 0.9, 0.9, 0.9, 0.9};
// Clementi radii used for the angular pruning zones (AtomGridFactory.cpp:53-66).
const double kClementiRadii[] = {1.00, 1.00, 0.59, 3.16, 2.12, 1.64, 1.27, 1.06, 0.91, 0.79, 0.72, 3.59, 2.74,
                                 2.23, 2.10, 1.85, 1.66, 1.49, 1.34, 4.59, 3.67, 3.48, 3.33, 3.23, 3.14, 3.04,
                                 2.95, 2.87, 2.82, 2.74, 2.68, 2.57, 2.36, 2.15, 1.95, 1.78, 1.66};
// Bragg-Slater radii in Angstrom (src/parameters/AtomicParameters.h:454-470;
// hydrogen 0.25 as in src/geometry/AtomTypeFactory.cpp:105-110).
const double kBraggSlaterAngstrom[] = {-1.0, 0.25, 1.20, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.38, 1.80, 1.50,
                                       1.25, 1.10, 1.00, 1.00, 1.00, 0.71, 2.20, 1.80, 1.60, 1.40, 1.35, 1.40, 1.40,
                                       1.40, 1.35, 1.35, 1.35, 1.35, 1.30, 1.25, 1.15, 1.15, 1.15, 0.88};
constexpr int kMaxZ = 36;

// Bohr radius in Angstrom as Serenity defines it (src/parameters/Constants.h:63,79).
constexpr double kBohrToAngstrom = 5.29177210544E-011 * 1.0E+010;
constexpr double kAngstromToBohr = 1.0 / kBohrToAngstrom;

int row_of(int Z) {
  if (Z < 3) return 1;
  if (Z < 11) return 2;
  if (Z < 19) return 3;
  if (Z < 37) return 4;
  if (Z < 55) return 5;
  if (Z < 87) return 6;
  return 7;
}

void check_Z(int Z) {
  if (Z < 1 || Z > kMaxZ) throw std::runtime_error("grid: element Z=" + std::to_string(Z) + " not tabulated");
}

// Expand one octahedral orbit (code 1..6) of a Lebedev rule.
void expand_orbit(const LebedevOrbit& o, std::vector<double>& x, std::vector<double>& y, std::vector<double>& z,
                  std::vector<double>& w) {
  auto put = [&](double px, double py, double pz) {
    x.push_back(px);
    y.push_back(py);
    z.push_back(pz);
    w.push_back(o.v);
  };
  auto signs3 = [&](double p, double q, double r) {
    // all sign combinations of the non-zero components of (p, q, r); the sign
    // of the first component alternates fastest (the published generator's order)
    const double sp[2] = {p, -p}, sq[2] = {q, -q}, sr[2] = {r, -r};
    const int np = (p == 0.0) ? 1 : 2, nq = (q == 0.0) ? 1 : 2, nr = (r == 0.0) ? 1 : 2;
    for (int k = 0; k < nr; ++k)
      for (int j = 0; j < nq; ++j)
        for (int i = 0; i < np; ++i) put(sp[i], sq[j], sr[k]);
  };
  switch (o.code) {
    case 1:  // (1,0,0) and permutations: 6 points
      signs3(1.0, 0.0, 0.0);
      signs3(0.0, 1.0, 0.0);
      signs3(0.0, 0.0, 1.0);
      break;
    case 2: {  // (0,a,a), a = 1/sqrt(2): 12 points
      const double a = std::sqrt(0.5);
      signs3(0.0, a, a);
      signs3(a, 0.0, a);
      signs3(a, a, 0.0);
    } break;
    case 3: {  // (a,a,a), a = 1/sqrt(3): 8 points
      const double a = std::sqrt(1.0 / 3.0);
      signs3(a, a, a);
    } break;
    case 4: {  // (a,a,b), b = sqrt(1-2a^2): 24 points
      const double a = o.a, b = std::sqrt(1.0 - 2.0 * a * a);
      signs3(a, a, b);
      signs3(a, b, a);
      signs3(b, a, a);
    } break;
    case 5: {  // (a,b,0), b = sqrt(1-a^2): 24 points
      const double a = o.a, b = std::sqrt(1.0 - a * a);
      signs3(a, b, 0.0);
      signs3(b, a, 0.0);
      signs3(a, 0.0, b);
      signs3(b, 0.0, a);
      signs3(0.0, a, b);
      signs3(0.0, b, a);
    } break;
    case 6: {  // (a,b,c), c = sqrt(1-a^2-b^2): 48 points
      const double a = o.a, b = o.b, c = std::sqrt(1.0 - a * a - b * b);
      signs3(a, b, c);
      signs3(a, c, b);
      signs3(b, a, c);
      signs3(b, c, a);
      signs3(c, a, b);
      signs3(c, b, a);
    } break;
    default:
      throw std::runtime_error("lebedev: bad orbit code");
  }
}

void radial_grid(RadialType type, double alpha, int n, std::vector<double>& r, std::vector<double>& w) {
  r.assign(n, 0.0);
  w.assign(n, 0.0);
  switch (type) {
    case RadialType::BECKE:
      for (int i = 1; i <= n; ++i) {
        const double xi = std::cos((double)i * M_PI / (n + 1));
        w[n - i] = std::sqrt(std::pow(1.0 + xi, 5) / std::pow(1.0 - xi, 7)) * (2.0 * M_PI) * std::pow(alpha, 3) / (n + 1);
        r[n - i] = alpha * (1.0 + xi) / (1.0 - xi);
      }
      break;
    case RadialType::HANDY:
      for (int i = 1; i <= n; ++i) {
        const double xi = i / (n + 1.0);
        r[i - 1] = alpha * xi * xi / ((1 - xi) * (1 - xi));
        w[i - 1] = 2 * std::pow(alpha, 3) * std::pow(xi, 5) / (n + 1) / std::pow((1 - xi), 7);
      }
      break;
    case RadialType::AHLRICHS:
      // Treutler-Ahlrichs M3 mapping with exponent 0.6 on Chebyshev nodes of the
      // second kind; stored with increasing radius.
      for (int i = 1; i <= n; ++i) {
        const double log2 = std::log(2.0);
        const double tmp = alpha / log2;
        const double xi = std::cos(i * M_PI / (n + 1.0));
        const double ri = tmp * std::pow((xi + 1.0), 0.6) * std::log(2.0 / (1.0 - xi));
        const double ln = std::log((1.0 - xi) / 2.0);
        const double sq = std::sqrt((1.0 + xi) / (1.0 - xi));
        const double wi = (M_PI / (n + 1.0)) * std::pow((1.0 + xi), (3 * 0.6)) * tmp * tmp * tmp *
                          (sq * ln * ln - 0.6 * ln * ln * ln / sq);
        w[n - i] = wi;
        r[n - i] = ri;
      }
      break;
    case RadialType::KNOWLES:
      for (int i = 1; i <= n; ++i) {
        const double xi = i / (n + 1.0);
        const double tmp = 1 - xi * xi * xi;
        const double ln = std::log(tmp);
        r[n - i] = -alpha * ln;
        w[n - i] = 3 * xi * xi * ln * ln * alpha * alpha * alpha / ((n + 1) * tmp);
      }
      break;
    case RadialType::EQUI:
      for (int i = 1; i <= n; ++i) {
        r[i - 1] = i * 5.0 / n;
        w[i - 1] = 1.0 / n;
      }
      break;
  }
}

// Reference grid of one element (not shifted to the atom), AtomGridFactory.cpp:78-203.
void atom_grid(int Z, const GridSettings& s, std::vector<double>& pts, std::vector<double>& wts) {
  check_Z(Z);
  if (s.accuracy < 1 || s.accuracy > 7) throw std::runtime_error("grid accuracy must be in [1,7]");
  static const int radAcc[7] = {13, 13, 13, 14, 15, 16, 17};
  const int row = row_of(Z);
  const int nRadial = (int)(5.0 * (radAcc[s.accuracy - 1] + row - 8));
  const double bsr = kBraggSlaterAngstrom[Z] * kAngstromToBohr;
  double alpha = 0.0;
  switch (s.radial) {
    case RadialType::BECKE:
    case RadialType::HANDY:
      alpha = (Z == 1) ? bsr : 0.5 * bsr;
      break;
    case RadialType::AHLRICHS:
      alpha = kAhlrichsAlpha[Z];
      break;
    case RadialType::KNOWLES:
      alpha = (Z == 1 || Z == 2 || Z == 3 || Z == 4 || Z == 11 || Z == 12 || Z == 19 || Z == 20) ? 5.0 : 7.0;
      break;
    case RadialType::EQUI:
      alpha = 0.0;
      break;
  }
  std::vector<double> rr, rw;
  radial_grid(s.radial, alpha, nRadial, rr, rw);

  // Angular pruning zones (ORCA-style, ldval/ranges tables of AtomGridFactory.cpp:156-180).
  static const int ldval[7][5] = {{4, 4, 4, 4, 4},     {4, 4, 4, 7, 4},      {4, 4, 7, 10, 7},     {4, 7, 10, 13, 10},
                                  {7, 10, 13, 15, 13}, {10, 13, 15, 16, 15}, {13, 15, 16, 17, 16}};
  static const double ranges[3][4] = {{0.25, 0.5, 1.0, 4.5}, {0.1667, 0.5, 0.9, 3.5}, {0.1, 0.4, 0.8, 2.5}};
  const int rrow = (row > 3) ? 2 : row - 1;
  const int redp1 = (row == 1 && s.accuracy > 1) ? 2 : 1;
  double zone[4];
  for (int i = 0; i < 4; ++i) zone[i] = ranges[rrow][i] * kClementiRadii[Z];
  int sphAcc = 0;
  pts.clear();
  wts.clear();
  std::vector<double> sx, sy, sz, sw;
  for (int i = 0; i < nRadial; ++i) {
    if (sphAcc < 4)
      if (rr[i] > zone[sphAcc]) ++sphAcc;
    lebedev_rule(lebedev_points_of_index(ldval[s.accuracy - redp1][sphAcc]), sx, sy, sz, sw);
    for (size_t k = 0; k < sx.size(); ++k) {
      wts.push_back(sw[k] * rw[i] * 4.0 * M_PI);
      pts.push_back(sx[k] * rr[i]);
      pts.push_back(sy[k] * rr[i]);
      pts.push_back(sz[k] * rr[i]);
    }
  }
}

double dist3(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

}  // namespace

int lebedev_points_of_index(int index) {
  if (index < 0 || index >= kLebedevNumRules) throw std::runtime_error("lebedev index out of range");
  return kLebedevRulePoints[index];
}

void lebedev_rule(int npoints, std::vector<double>& x, std::vector<double>& y, std::vector<double>& z,
                  std::vector<double>& w) {
  x.clear();
  y.clear();
  z.clear();
  w.clear();
  int first = 0;
  for (int r = 0; r < kLebedevNumRules; ++r) {
    if (kLebedevRulePoints[r] == npoints) {
      for (int o = 0; o < kLebedevRuleOrbits[r]; ++o) expand_orbit(kLebedevOrbits[first + o], x, y, z, w);
      if ((int)x.size() != npoints) throw std::runtime_error("lebedev: orbit expansion size mismatch");
      return;
    }
    first += kLebedevRuleOrbits[r];
  }
  throw std::runtime_error("lebedev: no rule with " + std::to_string(npoints) + " points");
}

int64_t hilbert_length(int64_t n) {
  int64_t length = 2, nvert = 8;
  while (n * 8 > nvert) {
    length *= 2;
    nvert *= 8;
  }
  return length;
}

void hilbert_sort(std::vector<double>& points, std::vector<double>& weights) {
  const int64_t n = (int64_t)weights.size();
  if (n < 2) return;
  const int64_t length = hilbert_length(n);
  double mn[3], spread[3];
  for (int d = 0; d < 3; ++d) {
    double lo = std::numeric_limits<double>::infinity(), hi = -lo;
    for (int64_t i = 0; i < n; ++i) {
      lo = std::min(lo, points[3 * i + d]);
      hi = std::max(hi, points[3 * i + d]);
    }
    mn[d] = lo;
    spread[d] = double(length) / (hi - lo);
  }
  std::vector<int64_t> key(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) key[i] = hilbert_key(&points[3 * i], mn, spread, length);
  // Descending key order; equal keys keep their input order (the order a
  // single-threaded run of the reference's selection/merge sort produces).
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[a] > key[b]; });
  std::vector<double> p2(points.size()), w2(n);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t s = order[i];
    p2[3 * i + 0] = points[3 * s + 0];
    p2[3 * i + 1] = points[3 * s + 1];
    p2[3 * i + 2] = points[3 * s + 2];
    w2[i] = weights[s];
  }
  points.swap(p2);
  weights.swap(w2);
}

GridProblem grid_problem(const std::vector<int>& Z, const std::vector<double>& xyz, const GridSettings& s,
                         bool with_candidates) {
  const int nAtoms = (int)Z.size();
  if (nAtoms == 0) throw std::runtime_error("grid: empty geometry");
  if ((int)xyz.size() != 3 * nAtoms) throw std::runtime_error("grid: coordinate array size mismatch");
  for (int i = 0; i < nAtoms; ++i) check_Z(Z[i]);
  // atoms closer than 1e-5 bohr count as duplicates (Geometry::hasIdenticalAtoms, Geometry.cpp:540-549)
  for (int i = 0; i < nAtoms; ++i)
    for (int j = 0; j < i; ++j)
      if (dist3(&xyz[3 * i], &xyz[3 * j]) < 1e-5) throw std::runtime_error("grid construction with duplicated atoms");

  std::vector<double> dist((size_t)nAtoms * nAtoms);
  for (int i = 0; i < nAtoms; ++i)
    for (int j = 0; j <= i; ++j) dist[i + (size_t)nAtoms * j] = dist[j + (size_t)nAtoms * i] = dist3(&xyz[3 * i], &xyz[3 * j]);

  // Becke atomic size adjustment (Becke, JCP 88 (1988) 2547, appendix).
  std::vector<double> aPair((size_t)nAtoms * nAtoms, 0.0);
  if (s.type == GridType::BECKE) {
    for (int i = 0; i < nAtoms; ++i)
      for (int j = 0; j < nAtoms; ++j) {
        const double q = std::sqrt(kBraggSlaterAngstrom[Z[j]] * kAngstromToBohr / (kBraggSlaterAngstrom[Z[i]] * kAngstromToBohr));
        const double u = (q - 1) / (q + 1);
        double a = u / (u * u - 1);
        a = std::min(0.5, std::max(-0.5, a));
        aPair[j + (size_t)nAtoms * i] = a;
      }
  }

  // Reference grid per element.
  std::vector<std::vector<double>> refPts(kMaxZ + 1), refW(kMaxZ + 1);
  for (int k = 0; k < nAtoms; ++k)
    if (refW[Z[k]].empty()) atom_grid(Z[k], s, refPts[Z[k]], refW[Z[k]]);

  // candidates: every atom's reference grid translated to the atom, atom order
  GridProblem P;
  P.natoms = nAtoms;
  P.type = (int)s.type;
  P.smoothing = s.smoothing;
  P.xyz = xyz;
  P.dist = dist;
  P.apair = aPair;
  P.min_dist.assign(nAtoms, 999999999.9);
  for (int k = 0; k < nAtoms; ++k)
    for (int l = 0; l < nAtoms; ++l)
      if (l != k) P.min_dist[k] = std::min(P.min_dist[k], dist[l + (size_t)nAtoms * k]);
  std::vector<int64_t> elem_off(kMaxZ + 1, -1);
  for (int k = 0; k < nAtoms; ++k) {
    if (elem_off[Z[k]] < 0) {
      elem_off[Z[k]] = (int64_t)P.ref_w.size();
      P.ref_xyz.insert(P.ref_xyz.end(), refPts[Z[k]].begin(), refPts[Z[k]].end());
      P.ref_w.insert(P.ref_w.end(), refW[Z[k]].begin(), refW[Z[k]].end());
    }
    P.ref_off.push_back(elem_off[Z[k]]);
  }
  P.cand_off.assign(nAtoms + 1, 0);
  for (int k = 0; k < nAtoms; ++k) P.cand_off[k + 1] = P.cand_off[k] + (int64_t)refW[Z[k]].size();
  for (int k = 0; k < nAtoms && with_candidates; ++k) {
    const auto& rp = refPts[Z[k]];
    const auto& rw = refW[Z[k]];
    for (size_t i = 0; i < rw.size(); ++i) {
      P.cand_xyz.push_back(rp[3 * i] + xyz[3 * k]);
      P.cand_xyz.push_back(rp[3 * i + 1] + xyz[3 * k + 1]);
      P.cand_xyz.push_back(rp[3 * i + 2] + xyz[3 * k + 2]);
      P.cand_w.push_back(rw[i]);
      P.cand_atom.push_back(k);
    }
  }
  return P;
}

void partition_weights_host(const GridProblem& P, std::vector<double>& w) {
  const int64_t nc = (int64_t)P.cand_w.size();
  w.resize(nc);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t c = 0; c < nc; ++c) {
    const int k = P.cand_atom[c];
    w[c] = partition_weight(P.type, P.smoothing, P.natoms, P.xyz.data(), P.dist.data(), P.apair.data(),
                            P.min_dist[k], k, &P.cand_xyz[3 * c], P.cand_w[c]);
  }
}

MolecularGrid finish_grid(const GridProblem& P, const std::vector<double>& w, const GridSettings& s) {
  MolecularGrid g;
  g.n_before_cut = (int64_t)P.cand_w.size();
  for (size_t c = 0; c < w.size(); ++c)
    if (w[c] > s.weight_threshold) {
      g.weights.push_back(w[c]);
      g.points.insert(g.points.end(), &P.cand_xyz[3 * c], &P.cand_xyz[3 * c] + 3);
    }
  if (s.sort) hilbert_sort(g.points, g.weights);
  return g;
}

MolecularGrid build_molecular_grid(const std::vector<int>& Z, const std::vector<double>& xyz, const GridSettings& s) {
  const GridProblem P = grid_problem(Z, xyz, s);
  std::vector<double> w;
  partition_weights_host(P, w);
  return finish_grid(P, w, s);
}

}  // namespace sxc
