// Host-side data model of the XC Fock-build library: contracted shells, the
// molecular integration grid and the real solid harmonics.  Everything here is
// plain C++17 on the CPU; device residency lives in xc_api.cpp.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace sxc {

// One contracted shell.  Mirrors what Serenity's Shell keeps for the grid path
// (src/basis/Shell.cpp:30-47 over libint2::Shell): centre, angular momentum,
// spherical flag, exponents and libint2-renormalised contraction coefficients.
struct Shell {
  std::array<double, 3> center{};
  int l = 0;
  bool pure = false;
  std::vector<double> alpha;
  std::vector<double> coef;  // normalised contraction coefficients
  int atom = -1;             // index of the atom the shell sits on
  int nfunc() const { return pure ? 2 * l + 1 : (l + 1) * (l + 2) / 2; }
};

// libint2-style renormalisation of a contraction (Shell::renorm in libint2
// 2.7, used by Serenity's Shell ctor): scale every primitive to unit norm for
// the axis-aligned Cartesian component x^l, then scale the contraction to unit
// self-overlap.
void renormalize_contraction(int l, const std::vector<double>& alpha, std::vector<double>& coef);

// Cartesian component factors sqrt((2l-1)!! / ((2a-1)!!(2b-1)!!(2c-1)!!)),
// components in Serenity's order (a = l..0, b = l-a..0; c = l-a-b)
// (src/basis/Shell.cpp:39-44, src/integrals/Normalization.h:82-87).
std::vector<double> cartesian_norm_factors(int l);

// Real solid harmonics in Racah normalisation, as a (2l+1) x ncart matrix of
// coefficients over the Cartesian monomials (same component order as above),
// rows m = -l..l.  Serenity's generated code (BasisFunctionOnGridController.cpp
// :469-1062, produced by src/data/grid/sharmonics.py) evaluates the same
// polynomials; here they come from the closed form of Helgaker, Jorgensen and
// Olsen, "Molecular Electronic-Structure Theory", eq. (6.4.47)-(6.4.50).
std::vector<double> solid_harmonic_coefficients(int l);

// Parse the Turbomole-format entry of `element` (e.g. "o") from a basis file
// (Serenity: BasisFunctionProvider.cpp:60-130).  Returns the shells with
// primitive data only (center/atom filled by the caller).  Throws on error.
std::vector<Shell> parse_turbomole_basis(const std::string& path, const std::string& element, bool pure);

// Grid construction settings (Serenity Settings.h:160-170 GRID block).
enum class GridType : int { BECKE = 0, VORONOI = 1, SSF = 2 };
enum class RadialType : int { BECKE = 0, HANDY = 1, AHLRICHS = 2, KNOWLES = 3, EQUI = 4 };

struct GridSettings {
  GridType type = GridType::SSF;
  RadialType radial = RadialType::AHLRICHS;
  int accuracy = 4;
  unsigned smoothing = 3;
  double weight_threshold = 1e-14;
  bool sort = true;
};

struct MolecularGrid {
  std::vector<double> points;   // 3 x N, x y z contiguous per point (bohr)
  std::vector<double> weights;  // N
  int64_t n_before_cut = 0;     // points before the weight cutoff
};

// Atom-centred Lebedev x radial grid with fuzzy-cell partitioning, weight cut
// and (optionally) Hilbert-curve ordering
// (GridFactory.cpp:51-300, AtomGridFactory.cpp:78-203, HilbertRTreeSorting.cpp).
MolecularGrid build_molecular_grid(const std::vector<int>& Z, const std::vector<double>& xyz_bohr,
                                   const GridSettings& settings);

// The partition problem of a molecular grid: atoms, interatomic distances,
// Becke size adjustments, and every candidate point (each atom's reference grid
// translated to the atom, in atom order) with its raw weight and atom.
struct GridProblem {
  int natoms = 0;
  int type = 2;
  unsigned smoothing = 3;
  std::vector<double> xyz, dist, apair, min_dist;
  std::vector<double> cand_xyz, cand_w;
  std::vector<int> cand_atom;
  // the same candidates as reference grids: ref_xyz / ref_w hold each element's atom grid once;
  // atom k's candidates are ref points ref_off[k] .. ref_off[k] + (cand_off[k + 1] - cand_off[k])
  // translated to the atom, candidate index cand_off[k] + i
  std::vector<double> ref_xyz, ref_w;
  std::vector<int64_t> ref_off, cand_off;
};
// with_candidates = false leaves cand_xyz / cand_w / cand_atom empty (the device builds them).
GridProblem grid_problem(const std::vector<int>& Z, const std::vector<double>& xyz_bohr, const GridSettings& s,
                         bool with_candidates = true);
// Hilbert-curve box of n points: the cube edge length (a power of 2, at least 2) the keys use.
int64_t hilbert_length(int64_t n);
// Partitioned weight of every candidate (-1: dropped by the SSF screening), on the host.
void partition_weights_host(const GridProblem& P, std::vector<double>& w);
// Weight cut (w > threshold, candidates in order) and optional Hilbert ordering.
MolecularGrid finish_grid(const GridProblem& P, const std::vector<double>& w, const GridSettings& s);

// Lebedev-Laikov rule with `npoints` points (6..5810); unit-sphere points and
// weights summing to 1.  Throws for an unknown order.
void lebedev_rule(int npoints, std::vector<double>& x, std::vector<double>& y, std::vector<double>& z,
                  std::vector<double>& w);
int lebedev_points_of_index(int index);  // Serenity's Lebedev index (0..31) -> number of points

// Hilbert-curve ordering of points (3 x N, xyz per point) and weights, in place
// (HilbertRTreeSorting.cpp:27-216).
void hilbert_sort(std::vector<double>& points, std::vector<double>& weights);

// Analytic overlap matrix S (n x n, column-major) of a shell list.
void overlap_matrix(const std::vector<Shell>& shells, std::vector<double>& S);

// ---- integrals of the SCF driver (integrals.cpp) ----------------------------
// AO pairs are packed as pair(mu, nu) = mu (mu + 1) / 2 + nu for mu >= nu.
// Boys function F_m(T), m <= 17.
double boys_function(int m, double T);
// Kinetic energy T and nuclear attraction V (point nuclei of charge Z at xyz,
// bohr), both n x n column-major.
void one_electron_matrices(const std::vector<Shell>& shells, const std::vector<int>& Z, const std::vector<double>& xyz,
                           std::vector<double>& T, std::vector<double>& V);
// Four-centre Coulomb integrals (mu nu|la si) as an npair x npair matrix
// (npair = n (n + 1) / 2); shell quartets below the Schwarz threshold are zero.
void eri_packed(const std::vector<Shell>& shells, double schwarz_threshold, std::vector<double>& eri);
// Density fitting of the Coulomb matrix: three-centre integrals (P|mu nu) as
// an naux x npair matrix B3 (row P) and the metric (P|Q), naux x naux.
void ri_j_integrals(const std::vector<Shell>& shells, const std::vector<Shell>& aux, double schwarz_threshold,
                    std::vector<double>& B3, std::vector<double>& M);
// The metric (P|Q) alone, naux x naux.
void ri_metric(const std::vector<Shell>& aux, std::vector<double>& M);

}  // namespace sxc
