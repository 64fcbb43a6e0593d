// One- and two-electron integrals for the SCF driver around the XC build:
// kinetic energy, nuclear attraction and electron-repulsion integrals over the
// same contracted shells (and the same Cartesian / solid-harmonic AO
// convention) as the grid path.  McMurchie-Davidson scheme: Hermite expansion
// coefficients E_t^{ij} of a Gaussian product, Hermite Coulomb integrals
// R_{tuv} from the Boys function (McMurchie & Davidson, J. Comput. Phys. 26
// (1978) 218; Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
// Theory", ch. 9).  The electron-repulsion integrals are returned packed over
// AO pairs (mu >= nu), so J[P] is one matrix-vector product per SCF iteration.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "host.h"

namespace sxc {

namespace {

constexpr int kMaxLInt = 4;                        // up to g shells
constexpr int kMaxHerm = 4 * kMaxLInt + 1;         // t + u + v of a quartet plus one

// Boys function F_m(T), m = 0..mmax: series with downward recursion below
// T = 35, asymptotic F_0 = sqrt(pi / T) / 2 with upward recursion above
// (erfc(sqrt(35)) < 1e-16).
void boys(int mmax, double T, double* F) {
  if (T < 35.0) {
    const double e = std::exp(-T);
    // F_mmax by its series  e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
    double term = 1.0 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * mmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[mmax] = e * sum;
    for (int m = mmax; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + e) / (2 * m - 1);
  } else {
    const double e = std::exp(-T);
    F[0] = 0.5 * std::sqrt(M_PI / T);
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
  }
}

// Hermite expansion coefficients E[i][j][t] of one Cartesian direction
// (without the exp(-mu X_AB^2) factor), i <= la, j <= lb, t <= i + j.
struct Herm1D {
  double E[kMaxLInt + 3][kMaxLInt + 3][2 * kMaxLInt + 5];
  void build(int la, int lb, double p, double XPA, double XPB) {
    const double o2p = 0.5 / p;
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j <= lb; ++j)
        for (int t = 0; t <= la + lb + 1; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j <= lb; ++j) {
        if (i == 0 && j == 0) continue;
        for (int t = 0; t <= i + j; ++t) {
          double v;
          if (i > 0) {
            v = XPA * E[i - 1][j][t];
            if (t > 0) v += o2p * E[i - 1][j][t - 1];
            if (t + 1 <= i - 1 + j) v += (t + 1) * E[i - 1][j][t + 1];
          } else {
            v = XPB * E[i][j - 1][t];
            if (t > 0) v += o2p * E[i][j - 1][t - 1];
            if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
          }
          E[i][j][t] = v;
        }
      }
  }
};

// Hermite Coulomb integrals R_{tuv} = R^{(0)}_{tuv}(a, X, Y, Z), t + u + v <= L,
// by the downward recursion over the auxiliary order n:
//   R^{(n)}_{000} = (-2a)^n F_n(a |PC|^2),
//   R^{(n)}_{t+1,u,v} = t R^{(n+1)}_{t-1,u,v} + X R^{(n+1)}_{t,u,v}  (likewise u, v).
struct HermR {
  int L = 0, n1 = 1;
  std::vector<double> cur, nxt;  // index (t * n1 + u) * n1 + v
  double at(int t, int u, int v) const { return nxt[((size_t)t * n1 + u) * n1 + v]; }
  void build(int Lmax, double a, double X, double Y, double Z) {
    L = Lmax;
    n1 = L + 1;
    const size_t sz = (size_t)n1 * n1 * n1;
    if (cur.size() < sz) {
      cur.resize(sz);
      nxt.resize(sz);
    }
    double F[kMaxHerm + 1], pw[kMaxHerm + 1];
    boys(L, a * (X * X + Y * Y + Z * Z), F);
    pw[0] = 1.0;
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (-2.0 * a);
    auto id = [&](int t, int u, int v) { return ((size_t)t * n1 + u) * n1 + v; };
    for (int n = L; n >= 0; --n) {
      // cur = R^{(n)} from nxt = R^{(n+1)}; entries t + u + v <= L - n
      cur[id(0, 0, 0)] = pw[n] * F[n];
      const int top = L - n;
      for (int s = 1; s <= top; ++s)
        for (int t = 0; t <= s; ++t)
          for (int u = 0; u <= s - t; ++u) {
            const int v = s - t - u;
            double val;
            if (t > 0) {
              val = X * nxt[id(t - 1, u, v)];
              if (t > 1) val += (t - 1) * nxt[id(t - 2, u, v)];
            } else if (u > 0) {
              val = Y * nxt[id(t, u - 1, v)];
              if (u > 1) val += (u - 1) * nxt[id(t, u - 2, v)];
            } else {
              val = Z * nxt[id(t, u, v - 1)];
              if (v > 1) val += (v - 1) * nxt[id(t, u, v - 2)];
            }
            cur[id(t, u, v)] = val;
          }
      std::swap(cur, nxt);
    }
  }
};

std::vector<double> angular_transform_of(const Shell& sh) {
  const int ncart = (sh.l + 1) * (sh.l + 2) / 2;
  if (sh.pure) return solid_harmonic_coefficients(sh.l);
  std::vector<double> M((size_t)ncart * ncart, 0.0);
  auto f = cartesian_norm_factors(sh.l);
  for (int c = 0; c < ncart; ++c) M[(size_t)c * ncart + c] = f[c];
  return M;
}

struct CartShell {
  std::vector<int> ax, ay, az;
  std::vector<double> M;  // nfunc x ncart
  int ncart = 0, nfunc = 0, off = 0;
};

std::vector<CartShell> cart_shells(const std::vector<Shell>& shells, int& n) {
  std::vector<CartShell> out(shells.size());
  n = 0;
  for (size_t s = 0; s < shells.size(); ++s) {
    const Shell& sh = shells[s];
    if (sh.l > kMaxLInt) throw std::runtime_error("integrals: angular momentum above g is not supported");
    CartShell& c = out[s];
    for (int a = sh.l; a >= 0; --a)
      for (int b = sh.l - a; b >= 0; --b) {
        c.ax.push_back(a);
        c.ay.push_back(b);
        c.az.push_back(sh.l - a - b);
      }
    c.ncart = (int)c.ax.size();
    c.nfunc = sh.nfunc();
    c.M = angular_transform_of(sh);
    c.off = n;
    n += c.nfunc;
  }
  return out;
}

// Cartesian block -> AO block: out[f][h] = sum_ij MA[f][i] X[i][j] MB[h][j].
void to_ao(const CartShell& A, const CartShell& B, const std::vector<double>& X, std::vector<double>& out) {
  out.assign((size_t)A.nfunc * B.nfunc, 0.0);
  for (int f = 0; f < A.nfunc; ++f)
    for (int i = 0; i < A.ncart; ++i) {
      const double m = A.M[(size_t)f * A.ncart + i];
      if (m == 0.0) continue;
      for (int h = 0; h < B.nfunc; ++h) {
        double v = 0.0;
        for (int j = 0; j < B.ncart; ++j) v += X[(size_t)i * B.ncart + j] * B.M[(size_t)h * B.ncart + j];
        out[(size_t)f * B.nfunc + h] += m * v;
      }
    }
}

// Primitive-pair data of a shell pair.
struct PrimPair {
  double p, K, P[3];  // exponent sum, coef_a coef_b exp(-mu AB^2), centre
  Herm1D ex, ey, ez;
};

std::vector<PrimPair> prim_pairs(const Shell& a, const Shell& b) {
  std::vector<PrimPair> out;
  double AB2 = 0.0;
  for (int d = 0; d < 3; ++d) AB2 += (a.center[d] - b.center[d]) * (a.center[d] - b.center[d]);
  for (size_t i = 0; i < a.alpha.size(); ++i)
    for (size_t j = 0; j < b.alpha.size(); ++j) {
      PrimPair pp;
      const double ai = a.alpha[i], bj = b.alpha[j];
      pp.p = ai + bj;
      pp.K = a.coef[i] * b.coef[j] * std::exp(-ai * bj / pp.p * AB2);
      if (pp.K == 0.0) continue;
      for (int d = 0; d < 3; ++d) pp.P[d] = (ai * a.center[d] + bj * b.center[d]) / pp.p;
      pp.ex.build(a.l, b.l, pp.p, pp.P[0] - a.center[0], pp.P[0] - b.center[0]);
      pp.ey.build(a.l, b.l, pp.p, pp.P[1] - a.center[1], pp.P[1] - b.center[1]);
      pp.ez.build(a.l, b.l, pp.p, pp.P[2] - a.center[2], pp.P[2] - b.center[2]);
      out.push_back(pp);
    }
  return out;
}

// Electron-repulsion integrals of one shell quartet in Cartesian components,
// out[((i * ncB + j) * ncC + k) * ncD + l].
void eri_quartet(const Shell& A, const Shell& B, const Shell& C, const Shell& D, const CartShell& cA,
                 const CartShell& cB, const CartShell& cC, const CartShell& cD, const std::vector<PrimPair>& ab,
                 const std::vector<PrimPair>& cd, HermR& R, std::vector<double>& out) {
  const int nA = cA.ncart, nB = cB.ncart, nC = cC.ncart, nD = cD.ncart;
  out.assign((size_t)nA * nB * nC * nD, 0.0);
  const int L = A.l + B.l + C.l + D.l;
  for (const auto& p : ab)
    for (const auto& q : cd) {
      const double alpha = p.p * q.p / (p.p + q.p);
      R.build(L, alpha, p.P[0] - q.P[0], p.P[1] - q.P[1], p.P[2] - q.P[2]);
      const double pref = 2.0 * std::pow(M_PI, 2.5) / (p.p * q.p * std::sqrt(p.p + q.p)) * p.K * q.K;
      for (int i = 0; i < nA; ++i)
        for (int j = 0; j < nB; ++j) {
          const int tx = cA.ax[i] + cB.ax[j], ty = cA.ay[i] + cB.ay[j], tz = cA.az[i] + cB.az[j];
          for (int k = 0; k < nC; ++k)
            for (int l = 0; l < nD; ++l) {
              const int sx = cC.ax[k] + cD.ax[l], sy = cC.ay[k] + cD.ay[l], sz = cC.az[k] + cD.az[l];
              double v = 0.0;
              for (int t = 0; t <= tx; ++t) {
                const double e1 = p.ex.E[cA.ax[i]][cB.ax[j]][t];
                if (e1 == 0.0) continue;
                for (int u = 0; u <= ty; ++u) {
                  const double e2 = e1 * p.ey.E[cA.ay[i]][cB.ay[j]][u];
                  if (e2 == 0.0) continue;
                  for (int w = 0; w <= tz; ++w) {
                    const double e3 = e2 * p.ez.E[cA.az[i]][cB.az[j]][w];
                    if (e3 == 0.0) continue;
                    double s = 0.0;
                    for (int tt = 0; tt <= sx; ++tt) {
                      const double f1 = q.ex.E[cC.ax[k]][cD.ax[l]][tt];
                      if (f1 == 0.0) continue;
                      for (int uu = 0; uu <= sy; ++uu) {
                        const double f2 = f1 * q.ey.E[cC.ay[k]][cD.ay[l]][uu];
                        if (f2 == 0.0) continue;
                        for (int ww = 0; ww <= sz; ++ww) {
                          const double f3 = f2 * q.ez.E[cC.az[k]][cD.az[l]][ww];
                          if (f3 == 0.0) continue;
                          const double sign = ((tt + uu + ww) & 1) ? -1.0 : 1.0;
                          s += sign * f3 * R.at(t + tt, u + uu, w + ww);
                        }
                      }
                    }
                    v += e3 * s;
                  }
                }
              }
              out[(((size_t)i * nB + j) * nC + k) * nD + l] += pref * v;
            }
        }
    }
}


// Cartesian quartet (ij|kl) -> AO quartet: transform the ket pair, then the bra pair.
void quartet_to_ao(const CartShell& cA, const CartShell& cB, const CartShell& cC, const CartShell& cD,
                   const std::vector<double>& qc, std::vector<double>& half, std::vector<double>& ao) {
  const int nA = cA.ncart, nB = cB.ncart, nC = cC.ncart, nD = cD.ncart;
  const int fA = cA.nfunc, fB = cB.nfunc, fC = cC.nfunc, fD = cD.nfunc;
  half.assign((size_t)nA * nB * fC * fD, 0.0);
  std::vector<double> ket, bra, out;
  for (int ij = 0; ij < nA * nB; ++ij) {
    ket.assign(qc.begin() + (size_t)ij * nC * nD, qc.begin() + (size_t)(ij + 1) * nC * nD);
    to_ao(cC, cD, ket, out);
    std::copy(out.begin(), out.end(), half.begin() + (size_t)ij * fC * fD);
  }
  ao.assign((size_t)fA * fB * fC * fD, 0.0);
  bra.resize((size_t)nA * nB);
  for (int kl = 0; kl < fC * fD; ++kl) {
    for (int ij = 0; ij < nA * nB; ++ij) bra[ij] = half[(size_t)ij * fC * fD + kl];
    to_ao(cA, cB, bra, out);
    for (int fg = 0; fg < fA * fB; ++fg) ao[(size_t)fg * fC * fD + kl] = out[fg];
  }
}

// The s-type unit function (exponent 0, coefficient 1) at the centre of `sh`:
// (P 1|mu nu) is the three-centre integral (P|mu nu), (P 1|Q 1) the two-centre
// metric (P|Q).
Shell unit_shell(const Shell& sh) {
  Shell u;
  u.center = sh.center;
  u.l = 0;
  u.pure = false;
  u.alpha = {0.0};
  u.coef = {1.0};
  return u;
}

}  // namespace

double boys_function(int m, double T) {
  if (m < 0 || m > kMaxHerm) throw std::runtime_error("boys_function: order out of range");
  double F[kMaxHerm + 1];
  boys(m, T, F);
  return F[m];
}

void one_electron_matrices(const std::vector<Shell>& shells, const std::vector<int>& Z, const std::vector<double>& xyz,
                           std::vector<double>& T, std::vector<double>& V) {
  int n = 0;
  const auto cs = cart_shells(shells, n);
  T.assign((size_t)n * n, 0.0);
  V.assign((size_t)n * n, 0.0);
  const int ns = (int)shells.size(), natoms = (int)Z.size();
#pragma omp parallel for schedule(dynamic)
  for (int A = 0; A < ns; ++A) {
    for (int B = 0; B <= A; ++B) {
      const Shell& a = shells[A];
      const Shell& b = shells[B];
      const CartShell& ca = cs[A];
      const CartShell& cb = cs[B];
      std::vector<double> Tc((size_t)ca.ncart * cb.ncart, 0.0), Vc((size_t)ca.ncart * cb.ncart, 0.0);
      // kinetic: overlaps with B's exponent raised by 2 (Obara-Saika form in Hermite language)
      double AB2 = 0.0;
      for (int d = 0; d < 3; ++d) AB2 += (a.center[d] - b.center[d]) * (a.center[d] - b.center[d]);
      for (size_t i = 0; i < a.alpha.size(); ++i)
        for (size_t j = 0; j < b.alpha.size(); ++j) {
          const double ai = a.alpha[i], bj = b.alpha[j], p = ai + bj;
          const double K = a.coef[i] * b.coef[j] * std::exp(-ai * bj / p * AB2);
          if (K == 0.0) continue;
          double P[3];
          Herm1D e[3];
          for (int d = 0; d < 3; ++d) {
            P[d] = (ai * a.center[d] + bj * b.center[d]) / p;
            e[d].build(a.l, b.l + 2, p, P[d] - a.center[d], P[d] - b.center[d]);
          }
          const double s0 = std::sqrt(M_PI / p);
          auto S1 = [&](int d, int ia, int jb) { return jb < 0 ? 0.0 : e[d].E[ia][jb][0] * s0; };
          auto T1 = [&](int d, int ia, int jb) {
            return -2.0 * bj * bj * S1(d, ia, jb + 2) + bj * (2 * jb + 1) * S1(d, ia, jb) -
                   0.5 * jb * (jb - 1) * S1(d, ia, jb - 2);
          };
          HermR R;
          for (int ii = 0; ii < ca.ncart; ++ii)
            for (int jj = 0; jj < cb.ncart; ++jj) {
              const int ax = ca.ax[ii], ay = ca.ay[ii], az = ca.az[ii];
              const int bx = cb.ax[jj], by = cb.ay[jj], bz = cb.az[jj];
              const double sx = S1(0, ax, bx), sy = S1(1, ay, by), sz = S1(2, az, bz);
              Tc[(size_t)ii * cb.ncart + jj] +=
                  K * (T1(0, ax, bx) * sy * sz + sx * T1(1, ay, by) * sz + sx * sy * T1(2, az, bz));
            }
          const int L = a.l + b.l;
          for (int c = 0; c < natoms; ++c) {
            R.build(L, p, P[0] - xyz[3 * c], P[1] - xyz[3 * c + 1], P[2] - xyz[3 * c + 2]);
            const double pref = -Z[c] * 2.0 * M_PI / p * K;
            for (int ii = 0; ii < ca.ncart; ++ii)
              for (int jj = 0; jj < cb.ncart; ++jj) {
                const int ax = ca.ax[ii], ay = ca.ay[ii], az = ca.az[ii];
                const int bx = cb.ax[jj], by = cb.ay[jj], bz = cb.az[jj];
                double v = 0.0;
                for (int t = 0; t <= ax + bx; ++t)
                  for (int u = 0; u <= ay + by; ++u)
                    for (int w = 0; w <= az + bz; ++w)
                      v += e[0].E[ax][bx][t] * e[1].E[ay][by][u] * e[2].E[az][bz][w] * R.at(t, u, w);
                Vc[(size_t)ii * cb.ncart + jj] += pref * v;
              }
          }
        }
      std::vector<double> To, Vo;
      to_ao(ca, cb, Tc, To);
      to_ao(ca, cb, Vc, Vo);
      for (int f = 0; f < ca.nfunc; ++f)
        for (int h = 0; h < cb.nfunc; ++h) {
          const size_t r = ca.off + f, c = cb.off + h;
          T[r + (size_t)n * c] = T[c + (size_t)n * r] = To[(size_t)f * cb.nfunc + h];
          V[r + (size_t)n * c] = V[c + (size_t)n * r] = Vo[(size_t)f * cb.nfunc + h];
        }
    }
  }
}

void eri_packed(const std::vector<Shell>& shells, double schwarz_threshold, std::vector<double>& eri) {
  int n = 0;
  const auto cs = cart_shells(shells, n);
  const int ns = (int)shells.size();
  const size_t npair = (size_t)n * (n + 1) / 2;
  eri.assign(npair * npair, 0.0);
  auto pidx = [](int m, int v) { return m >= v ? (size_t)m * (m + 1) / 2 + v : (size_t)v * (v + 1) / 2 + m; };
  // shell pairs A >= B with their primitive pairs and Schwarz bound sqrt(max |(ab|ab)|)
  std::vector<std::pair<int, int>> pairs;
  for (int A = 0; A < ns; ++A)
    for (int B = 0; B <= A; ++B) pairs.emplace_back(A, B);
  const int np = (int)pairs.size();
  std::vector<std::vector<PrimPair>> pp(np);
  std::vector<double> bound(np, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < np; ++k) {
    const int A = pairs[k].first, B = pairs[k].second;
    pp[k] = prim_pairs(shells[A], shells[B]);
    std::vector<double> q;
    HermR R;
    eri_quartet(shells[A], shells[B], shells[A], shells[B], cs[A], cs[B], cs[A], cs[B], pp[k], pp[k], R, q);
    const int nA = cs[A].ncart, nB = cs[B].ncart;
    double mx = 0.0;
    for (int i = 0; i < nA; ++i)
      for (int j = 0; j < nB; ++j) mx = std::max(mx, std::fabs(q[(((size_t)i * nB + j) * nA + i) * nB + j]));
    bound[k] = std::sqrt(mx);
  }
#pragma omp parallel for schedule(dynamic)
  for (int k1 = 0; k1 < np; ++k1) {
    std::vector<double> qc, half, ao;
    HermR R;
    for (int k2 = 0; k2 <= k1; ++k2) {
      if (bound[k1] * bound[k2] < schwarz_threshold) continue;
      const int A = pairs[k1].first, B = pairs[k1].second, C = pairs[k2].first, D = pairs[k2].second;
      const CartShell &cA = cs[A], &cB = cs[B], &cC = cs[C], &cD = cs[D];
      eri_quartet(shells[A], shells[B], shells[C], shells[D], cA, cB, cC, cD, pp[k1], pp[k2], R, qc);
      quartet_to_ao(cA, cB, cC, cD, qc, half, ao);
      const int fA = cA.nfunc, fB = cB.nfunc, fC = cC.nfunc, fD = cD.nfunc;
      for (int f = 0; f < fA; ++f)
        for (int g = 0; g < fB; ++g) {
          const int mu = cA.off + f, nu = cB.off + g;
          if (A == B && nu > mu) continue;
          const size_t P = pidx(mu, nu);
          for (int h = 0; h < fC; ++h)
            for (int l = 0; l < fD; ++l) {
              const int la = cC.off + h, si = cD.off + l;
              if (C == D && si > la) continue;
              const size_t Q = pidx(la, si);
              const double v = ao[(((size_t)f * fB + g) * fC + h) * fD + l];
              eri[P * npair + Q] = v;
              eri[Q * npair + P] = v;
            }
        }
    }
  }
}

void ri_metric(const std::vector<Shell>& aux, std::vector<double>& M) {
  int naux = 0;
  const auto ca = cart_shells(aux, naux);
  const int na = (int)aux.size();
  M.assign((size_t)naux * naux, 0.0);
  std::vector<Shell> units(na);
  std::vector<CartShell> cu(na);
  std::vector<std::vector<PrimPair>> ap(na);
  for (int P = 0; P < na; ++P) {
    units[P] = unit_shell(aux[P]);
    int dummy = 0;
    cu[P] = cart_shells({units[P]}, dummy)[0];
    ap[P] = prim_pairs(aux[P], units[P]);
  }
#pragma omp parallel for schedule(dynamic)
  for (int P = 0; P < na; ++P) {
    HermR R;
    std::vector<double> qc, half, ao;
    for (int Q = 0; Q <= P; ++Q) {
      eri_quartet(aux[P], units[P], aux[Q], units[Q], ca[P], cu[P], ca[Q], cu[Q], ap[P], ap[Q], R, qc);
      quartet_to_ao(ca[P], cu[P], ca[Q], cu[Q], qc, half, ao);
      for (int f = 0; f < ca[P].nfunc; ++f)
        for (int h = 0; h < ca[Q].nfunc; ++h) {
          const size_t r = ca[P].off + f, c = ca[Q].off + h;
          M[r * naux + c] = M[c * naux + r] = ao[(size_t)f * ca[Q].nfunc + h];
        }
    }
  }
}

void ri_j_integrals(const std::vector<Shell>& shells, const std::vector<Shell>& aux, double schwarz_threshold,
                    std::vector<double>& B3, std::vector<double>& M) {
  int n = 0, naux = 0;
  const auto cs = cart_shells(shells, n);
  const auto ca = cart_shells(aux, naux);
  const int ns = (int)shells.size(), na = (int)aux.size();
  const size_t npair = (size_t)n * (n + 1) / 2;
  B3.assign((size_t)naux * npair, 0.0);
  M.assign((size_t)naux * naux, 0.0);
  std::vector<Shell> units(na);
  std::vector<CartShell> cu(na);
  std::vector<std::vector<PrimPair>> ap(na);
  for (int P = 0; P < na; ++P) {
    units[P] = unit_shell(aux[P]);
    int dummy = 0;
    cu[P] = cart_shells({units[P]}, dummy)[0];
    ap[P] = prim_pairs(aux[P], units[P]);
  }
  // metric (P|Q), and Schwarz bounds of the auxiliary functions
  std::vector<double> abound(na, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int P = 0; P < na; ++P) {
    HermR R;
    std::vector<double> qc, half, ao;
    for (int Q = 0; Q <= P; ++Q) {
      eri_quartet(aux[P], units[P], aux[Q], units[Q], ca[P], cu[P], ca[Q], cu[Q], ap[P], ap[Q], R, qc);
      quartet_to_ao(ca[P], cu[P], ca[Q], cu[Q], qc, half, ao);
      for (int f = 0; f < ca[P].nfunc; ++f)
        for (int h = 0; h < ca[Q].nfunc; ++h) {
          const size_t r = ca[P].off + f, c = ca[Q].off + h;
          M[r * naux + c] = M[c * naux + r] = ao[(size_t)f * ca[Q].nfunc + h];
          if (P == Q && f == h) abound[P] = std::max(abound[P], std::sqrt(std::fabs(ao[(size_t)f * ca[Q].nfunc + h])));
        }
    }
  }
  std::vector<std::pair<int, int>> pairs;
  for (int A = 0; A < ns; ++A)
    for (int B = 0; B <= A; ++B) pairs.emplace_back(A, B);
  const int np = (int)pairs.size();
  std::vector<std::vector<PrimPair>> pp(np);
  std::vector<double> bound(np, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < np; ++k) {
    const int A = pairs[k].first, B = pairs[k].second;
    pp[k] = prim_pairs(shells[A], shells[B]);
    HermR R;
    std::vector<double> q;
    eri_quartet(shells[A], shells[B], shells[A], shells[B], cs[A], cs[B], cs[A], cs[B], pp[k], pp[k], R, q);
    const int nA = cs[A].ncart, nB = cs[B].ncart;
    double mx = 0.0;
    for (int i = 0; i < nA; ++i)
      for (int j = 0; j < nB; ++j) mx = std::max(mx, std::fabs(q[(((size_t)i * nB + j) * nA + i) * nB + j]));
    bound[k] = std::sqrt(mx);
  }
  const double amax = na ? *std::max_element(abound.begin(), abound.end()) : 0.0;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < np; ++k) {
    if (bound[k] * amax < schwarz_threshold) continue;
    const int A = pairs[k].first, B = pairs[k].second;
    HermR R;
    std::vector<double> qc, half, ao;
    for (int P = 0; P < na; ++P) {
      if (bound[k] * abound[P] < schwarz_threshold) continue;
      eri_quartet(aux[P], units[P], shells[A], shells[B], ca[P], cu[P], cs[A], cs[B], ap[P], pp[k], R, qc);
      quartet_to_ao(ca[P], cu[P], cs[A], cs[B], qc, half, ao);
      const int fA = cs[A].nfunc, fB = cs[B].nfunc;
      for (int f = 0; f < ca[P].nfunc; ++f)
        for (int g = 0; g < fA; ++g)
          for (int h = 0; h < fB; ++h) {
            const int mu = cs[A].off + g, nu = cs[B].off + h;
            if (A == B && nu > mu) continue;
            B3[(size_t)(ca[P].off + f) * npair + (size_t)mu * (mu + 1) / 2 + nu] = ao[((size_t)f * fA + g) * fB + h];
          }
    }
  }
}

}  // namespace sxc
