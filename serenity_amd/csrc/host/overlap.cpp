// Analytic AO overlap matrix (Obara-Saika recursion over Cartesian Gaussians,
// then the shell's angular transform).  Used to build the orthonormal
// synthetic density matrices P = 2 C C^T with C = S^(-1/2) Q of the benchmark
// workloads.
#include <cmath>
#include <stdexcept>

#include "host.h"

namespace sxc {

namespace {

// Angular transform of a shell: AO f = sum_c M[f][c] x^a y^b z^c R(r).
std::vector<double> angular_transform(const Shell& sh) {
  const int ncart = (sh.l + 1) * (sh.l + 2) / 2;
  if (sh.pure) return solid_harmonic_coefficients(sh.l);
  std::vector<double> M((size_t)ncart * ncart, 0.0);
  auto f = cartesian_norm_factors(sh.l);
  for (int c = 0; c < ncart; ++c) M[(size_t)c * ncart + c] = f[c];
  return M;
}

void cart_exponents(int l, std::vector<int>& ax, std::vector<int>& ay, std::vector<int>& az) {
  ax.clear();
  ay.clear();
  az.clear();
  for (int a = l; a >= 0; --a)
    for (int b = l - a; b >= 0; --b) {
      ax.push_back(a);
      ay.push_back(b);
      az.push_back(l - a - b);
    }
}

// 1D overlap table s[i][j], i <= la, j <= lb (Obara-Saika).
void os1d(int la, int lb, double PA, double PB, double oo2p, double s00, double* s, int ld) {
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      double v;
      if (i == 0 && j == 0) {
        v = s00;
      } else if (i > 0) {
        v = PA * s[(i - 1) * ld + j];
        if (i > 1) v += oo2p * (i - 1) * s[(i - 2) * ld + j];
        if (j > 0) v += oo2p * j * s[(i - 1) * ld + j - 1];
      } else {
        v = PB * s[i * ld + j - 1];
        if (j > 1) v += oo2p * (j - 1) * s[i * ld + j - 2];
      }
      s[i * ld + j] = v;
    }
}

}  // namespace

void overlap_matrix(const std::vector<Shell>& shells, std::vector<double>& S) {
  int n = 0;
  std::vector<int> off;
  for (const auto& sh : shells) {
    off.push_back(n);
    n += sh.nfunc();
  }
  S.assign((size_t)n * n, 0.0);
  std::vector<std::vector<double>> M(shells.size());
  for (size_t s = 0; s < shells.size(); ++s) M[s] = angular_transform(shells[s]);
  const int ns = (int)shells.size();
#pragma omp parallel for schedule(dynamic)
  for (int A = 0; A < ns; ++A) {
    std::vector<int> axA, ayA, azA, axB, ayB, azB;
    const Shell& a = shells[A];
    cart_exponents(a.l, axA, ayA, azA);
    const int ncA = (int)axA.size();
    double sx[8 * 8], sy[8 * 8], sz[8 * 8];
    for (int B = 0; B <= A; ++B) {
      const Shell& b = shells[B];
      cart_exponents(b.l, axB, ayB, azB);
      const int ncB = (int)axB.size();
      std::vector<double> Sc((size_t)ncA * ncB, 0.0);
      const double AB2 = (a.center[0] - b.center[0]) * (a.center[0] - b.center[0]) +
                         (a.center[1] - b.center[1]) * (a.center[1] - b.center[1]) +
                         (a.center[2] - b.center[2]) * (a.center[2] - b.center[2]);
      for (size_t p = 0; p < a.alpha.size(); ++p)
        for (size_t q = 0; q < b.alpha.size(); ++q) {
          const double ap = a.alpha[p], bq = b.alpha[q], g = ap + bq;
          const double K = std::exp(-ap * bq / g * AB2);
          if (K * a.coef[p] * b.coef[q] == 0.0) continue;
          const double s00 = std::sqrt(M_PI / g);
          const double oo2p = 0.5 / g;
          double P[3], PA[3], PB[3];
          for (int d = 0; d < 3; ++d) {
            P[d] = (ap * a.center[d] + bq * b.center[d]) / g;
            PA[d] = P[d] - a.center[d];
            PB[d] = P[d] - b.center[d];
          }
          os1d(a.l, b.l, PA[0], PB[0], oo2p, s00, sx, 8);
          os1d(a.l, b.l, PA[1], PB[1], oo2p, s00, sy, 8);
          os1d(a.l, b.l, PA[2], PB[2], oo2p, s00, sz, 8);
          const double pref = K * a.coef[p] * b.coef[q];
          for (int i = 0; i < ncA; ++i)
            for (int j = 0; j < ncB; ++j)
              Sc[(size_t)i * ncB + j] +=
                  pref * sx[axA[i] * 8 + axB[j]] * sy[ayA[i] * 8 + ayB[j]] * sz[azA[i] * 8 + azB[j]];
        }
      // Transform to AO functions: S_AB = M_A Sc M_B^T.
      const int nfA = a.nfunc(), nfB = b.nfunc();
      for (int f = 0; f < nfA; ++f)
        for (int h = 0; h < nfB; ++h) {
          double v = 0.0;
          for (int i = 0; i < ncA; ++i) {
            const double mfi = M[A][(size_t)f * ncA + i];
            if (mfi == 0.0) continue;
            for (int j = 0; j < ncB; ++j) v += mfi * Sc[(size_t)i * ncB + j] * M[B][(size_t)h * ncB + j];
          }
          const size_t r = off[A] + f, c = off[B] + h;
          S[r + (size_t)n * c] = v;
          S[c + (size_t)n * r] = v;
        }
    }
  }
}

}  // namespace sxc
