// TEST INFRASTRUCTURE -- the CPU oracle of the XC Fock build.
//
// A plain C++17/OpenMP FP64 restatement of Serenity's FuncPotential path,
// block by block as the reference does it.  Only tests/, the smoke check and
// the CPU-baseline leg of bench.py load it (as libxc_oracle.so, through
// oracle/oracle.py); the library under serenity_amd/ never calls it.
//
// Functions and the reference code each follows:
//   oracle_ao_block      BasisFunctionOnGridController::calculateBasisFunctionData
//                        (src/data/grid/BasisFunctionOnGridController.cpp:150-1105)
//   oracle_density       MatrixOperatorToGridTransformer::transform
//                        (src/data/grid/MatrixOperatorToGridTransformer.cpp:37-198)
//   oracle_functional    XCFun::calcData + calcEnergy (src/dft/functionals/wrappers/XCFun.cpp:38-158,
//                        751-765) for LDA (slaterx + vwn5c), PBE (pbex + pbec), TF (tfk)
//   oracle_fock_add      ScalarOperatorToMatrixAdder::addScalarOperatorToMatrix / addBlock
//                        (src/data/grid/ScalarOperatorToMatrixAdder.cpp:51-113, 179-303)
//   oracle_fock_build    FuncPotential::getMatrix / getEnergy (src/potentials/FuncPotential.cpp:66-111),
//                        with the reference's three AO evaluations per build
//   oracle_xc_gradient   FuncPotential::getGeomGradients (src/potentials/FuncPotential.cpp:113-239)
//   oracle_functional_pol, oracle_fock_build_pol
//                        the unrestricted forms (FuncPotential<UNRESTRICTED>, XCFun.cpp:38-158 with
//                        the spin-density inputs; ScalarOperatorToMatrixAdder for_spin)
//
// Parity pins (tests/test_oracle.py): the reference's Mathematica known-answer
// values for AO values/gradients/Hessians, density/gradient/Hessian and the
// Fock assembly; the spherical normalisation invariant; finite-difference
// checks of every functional derivative and of the AO gradients; the stored
// E_xc of Serenity's He and water PBE test systems and the H2 LDA Fock matrix.
// The solid harmonics here use the real-solid-harmonic recurrences of
// Helgaker, Jorgensen, Olsen (eqs. 6.4.70-6.4.73), evaluated per point, not
// the coefficient tables the product uses.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct OShell {
  double A[3];
  int l, pure, nprim;
  const double* alpha;
  const double* coef;
  int first;  // first AO
  int nfunc;
};

struct Basis {
  std::vector<OShell> sh;
  int nao = 0;
  Basis(int nshell, const int* l, const int* pure, const int* nprim, const double* center, const double* alpha,
        const double* coef) {
    int off = 0, ao = 0;
    for (int s = 0; s < nshell; ++s) {
      OShell o;
      for (int d = 0; d < 3; ++d) o.A[d] = center[3 * s + d];
      o.l = l[s];
      o.pure = pure[s];
      o.nprim = nprim[s];
      o.alpha = alpha + off;
      o.coef = coef + off;
      o.first = ao;
      o.nfunc = o.pure ? 2 * o.l + 1 : (o.l + 1) * (o.l + 2) / 2;
      off += o.nprim;
      ao += o.nfunc;
      sh.push_back(o);
    }
    nao = ao;
  }
};

double dfact(int k) {
  double r = 1.0;
  for (int i = k; i > 1; i -= 2) r *= i;
  return r;
}

// Real solid harmonics S_lm (Racah normalisation) and their first and second
// derivatives by the HJO recurrences (each recurrence differentiated once and
// twice).  out[(l)][m+l] for one point.
struct Harm {
  double v[8][15], x[8][15], y[8][15], z[8][15];
  double xx[8][15], xy[8][15], xz[8][15], yy[8][15], yz[8][15], zz[8][15];
};

void solid_harmonics(int lmax, double X, double Y, double Z, Harm& H) {
  const double r2 = X * X + Y * Y + Z * Z;
  for (int l = 0; l <= lmax; ++l)
    for (int k = 0; k < 2 * l + 1; ++k)
      H.v[l][k] = H.x[l][k] = H.y[l][k] = H.z[l][k] = H.xx[l][k] = H.xy[l][k] = H.xz[l][k] = H.yy[l][k] =
          H.yz[l][k] = H.zz[l][k] = 0.0;
  H.v[0][0] = 1.0;
  for (int l = 0; l < lmax; ++l) {
    // diagonal terms m = +-(l+1)
    const double c = std::sqrt((l == 0 ? 2.0 : 1.0) * (2.0 * l + 1.0) / (2.0 * l + 2.0));
    const double sll = H.v[l][2 * l], slm = H.v[l][0];  // S_{l,l}, S_{l,-l}
    const double f = (l == 0) ? 0.0 : 1.0;
    const int P = 2 * (l + 1), M0 = 0;
    H.v[l + 1][P] = c * (X * sll - f * Y * slm);
    H.x[l + 1][P] = c * (sll + X * H.x[l][2 * l] - f * Y * H.x[l][0]);
    H.y[l + 1][P] = c * (X * H.y[l][2 * l] - f * (slm + Y * H.y[l][0]));
    H.z[l + 1][P] = c * (X * H.z[l][2 * l] - f * Y * H.z[l][0]);
    H.v[l + 1][M0] = c * (Y * sll + f * X * slm);
    H.x[l + 1][M0] = c * (Y * H.x[l][2 * l] + f * (slm + X * H.x[l][0]));
    H.y[l + 1][M0] = c * (sll + Y * H.y[l][2 * l] + f * X * H.y[l][0]);
    H.z[l + 1][M0] = c * (Y * H.z[l][2 * l] + f * X * H.z[l][0]);
    {
      // second derivatives of S_{l+1,l+1} = c (X s - f Y t) and S_{l+1,-l-1} = c (Y s + f X t),
      // s = S_{l,l}, t = S_{l,-l}
      const int a = 2 * l, b = 0;
      H.xx[l + 1][P] = c * (2.0 * H.x[l][a] + X * H.xx[l][a] - f * Y * H.xx[l][b]);
      H.xy[l + 1][P] = c * (H.y[l][a] + X * H.xy[l][a] - f * (H.x[l][b] + Y * H.xy[l][b]));
      H.xz[l + 1][P] = c * (H.z[l][a] + X * H.xz[l][a] - f * Y * H.xz[l][b]);
      H.yy[l + 1][P] = c * (X * H.yy[l][a] - f * (2.0 * H.y[l][b] + Y * H.yy[l][b]));
      H.yz[l + 1][P] = c * (X * H.yz[l][a] - f * (H.z[l][b] + Y * H.yz[l][b]));
      H.zz[l + 1][P] = c * (X * H.zz[l][a] - f * Y * H.zz[l][b]);
      H.xx[l + 1][M0] = c * (Y * H.xx[l][a] + f * (2.0 * H.x[l][b] + X * H.xx[l][b]));
      H.xy[l + 1][M0] = c * (H.x[l][a] + Y * H.xy[l][a] + f * (H.y[l][b] + X * H.xy[l][b]));
      H.xz[l + 1][M0] = c * (Y * H.xz[l][a] + f * (H.z[l][b] + X * H.xz[l][b]));
      H.yy[l + 1][M0] = c * (2.0 * H.y[l][a] + Y * H.yy[l][a] + f * X * H.yy[l][b]);
      H.yz[l + 1][M0] = c * (H.z[l][a] + Y * H.yz[l][a] + f * X * H.yz[l][b]);
      H.zz[l + 1][M0] = c * (Y * H.zz[l][a] + f * X * H.zz[l][b]);
    }
    // |m| <= l by the z recurrence
    for (int m = -l; m <= l; ++m) {
      const double a = (2.0 * l + 1.0);
      const double b = (std::abs(m) <= l - 1) ? std::sqrt((double)(l + m) * (l - m)) : 0.0;
      const double d = std::sqrt((double)(l + m + 1) * (l - m + 1));
      const int k1 = m + l + 1;  // index in l+1
      const int k0 = m + l;      // index in l
      double sm1 = 0.0, sm1x = 0.0, sm1y = 0.0, sm1z = 0.0;
      double w2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // xx xy xz yy yz zz of S_{l-1,m}
      if (b != 0.0) {
        const int km = m + l - 1;
        sm1 = H.v[l - 1][km];
        sm1x = H.x[l - 1][km];
        sm1y = H.y[l - 1][km];
        sm1z = H.z[l - 1][km];
        w2[0] = H.xx[l - 1][km];
        w2[1] = H.xy[l - 1][km];
        w2[2] = H.xz[l - 1][km];
        w2[3] = H.yy[l - 1][km];
        w2[4] = H.yz[l - 1][km];
        w2[5] = H.zz[l - 1][km];
      }
      H.v[l + 1][k1] = (a * Z * H.v[l][k0] - b * r2 * sm1) / d;
      H.x[l + 1][k1] = (a * Z * H.x[l][k0] - b * (2.0 * X * sm1 + r2 * sm1x)) / d;
      H.y[l + 1][k1] = (a * Z * H.y[l][k0] - b * (2.0 * Y * sm1 + r2 * sm1y)) / d;
      H.z[l + 1][k1] = (a * (H.v[l][k0] + Z * H.z[l][k0]) - b * (2.0 * Z * sm1 + r2 * sm1z)) / d;
      // S = (a Z u - b r^2 w) / d:  S_ij = (a (dz_i u_j + dz_j u_i + Z u_ij)
      //                                   - b (2 d_ij w + 2 x_i w_j + 2 x_j w_i + r^2 w_ij)) / d
      const double u = H.v[l][k0], ux = H.x[l][k0], uy = H.y[l][k0], uz = H.z[l][k0];
      H.xx[l + 1][k1] = (a * Z * H.xx[l][k0] - b * (2.0 * sm1 + 4.0 * X * sm1x + r2 * w2[0])) / d;
      H.xy[l + 1][k1] = (a * Z * H.xy[l][k0] - b * (2.0 * X * sm1y + 2.0 * Y * sm1x + r2 * w2[1])) / d;
      H.xz[l + 1][k1] = (a * (ux + Z * H.xz[l][k0]) - b * (2.0 * X * sm1z + 2.0 * Z * sm1x + r2 * w2[2])) / d;
      H.yy[l + 1][k1] = (a * Z * H.yy[l][k0] - b * (2.0 * sm1 + 4.0 * Y * sm1y + r2 * w2[3])) / d;
      H.yz[l + 1][k1] = (a * (uy + Z * H.yz[l][k0]) - b * (2.0 * Y * sm1z + 2.0 * Z * sm1y + r2 * w2[4])) / d;
      H.zz[l + 1][k1] = (a * (2.0 * uz + Z * H.zz[l][k0]) - b * (2.0 * sm1 + 4.0 * Z * sm1z + r2 * w2[5])) / d;
      (void)u;
    }
  }
}

struct BlockAO {
  int B = 0, n = 0, deriv = 0;
  std::vector<double> val, gx, gy, gz;                // B x n column-major
  std::vector<double> hxx, hxy, hxz, hyy, hyz, hzz;   // B x n
  std::vector<int> neg;                               // n flags, 1 = negligible
};

// AO values of one block, dense over all AOs, with the reference's screening:
// shell negligible for the whole block when (|centre - A| - spread) >= 1 and
// |sum_k c_k exp(-a_k d^2)| < thr; per point, primitives with a r^2 >= -ln thr
// are skipped and the shell is zero when |R| < thr.
void ao_block(const Basis& bas, const double* pts, int B, double thr, int deriv, BlockAO& out) {
  const int n = bas.nao;
  out.B = B;
  out.n = n;
  out.deriv = deriv;
  out.val.assign((size_t)B * n, 0.0);
  if (deriv >= 1) {
    out.gx.assign((size_t)B * n, 0.0);
    out.gy.assign((size_t)B * n, 0.0);
    out.gz.assign((size_t)B * n, 0.0);
  }
  if (deriv >= 2) {
    for (auto* h : {&out.hxx, &out.hxy, &out.hxz, &out.hyy, &out.hyz, &out.hzz}) h->assign((size_t)B * n, 0.0);
  }
  out.neg.assign(n, 0);
  double c[3] = {0, 0, 0};
  for (int p = 0; p < B; ++p)
    for (int d = 0; d < 3; ++d) c[d] += pts[3 * p + d];
  for (int d = 0; d < 3; ++d) c[d] /= B;
  double spread = 0.0;
  for (int p = 0; p < B; ++p) {
    const double dx = pts[3 * p] - c[0], dy = pts[3 * p + 1] - c[1], dz = pts[3 * p + 2] - c[2];
    spread = std::max(spread, std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  std::vector<char> skip(bas.sh.size(), 0);
  for (size_t s = 0; s < bas.sh.size(); ++s) {
    const OShell& o = bas.sh[s];
    const double dx = c[0] - o.A[0], dy = c[1] - o.A[1], dz = c[2] - o.A[2];
    double d = std::sqrt(dx * dx + dy * dy + dz * dz) - spread;
    if (d < 1.0) continue;
    d *= d;
    double radial = 0.0;
    for (int k = 0; k < o.nprim; ++k) radial += o.coef[k] * std::exp(-o.alpha[k] * d);
    if (std::fabs(radial) < thr) {
      skip[s] = 1;
      for (int f = 0; f < o.nfunc; ++f) out.neg[o.first + f] = 1;
    }
  }
  const double expthr = -std::log(thr);
  Harm H;
  for (int p = 0; p < B; ++p) {
    for (size_t s = 0; s < bas.sh.size(); ++s) {
      if (skip[s]) continue;
      const OShell& o = bas.sh[s];
      const double X = pts[3 * p] - o.A[0], Y = pts[3 * p + 1] - o.A[1], Z = pts[3 * p + 2] - o.A[2];
      const double r2 = X * X + Y * Y + Z * Z;
      double R = 0.0, dR = 0.0, ddR = 0.0;
      for (int k = 0; k < o.nprim; ++k) {
        const double t = o.alpha[k] * r2;
        if (t < expthr) {
          const double e = std::exp(-t);
          R += o.coef[k] * e;
          dR -= 2.0 * o.alpha[k] * o.coef[k] * e;
          ddR += 4.0 * o.alpha[k] * o.alpha[k] * o.coef[k] * e;
        }
      }
      if (std::fabs(R) < thr) continue;  // stays zero
      const int l = o.l;
      if (!o.pure) {
        int f = 0;
        const double top = dfact(2 * l - 1);
        for (int a = l; a >= 0; --a)
          for (int b = l - a; b >= 0; --b, ++f) {
            const int cc = l - a - b;
            const double N = std::sqrt(top / (dfact(2 * a - 1) * dfact(2 * b - 1) * dfact(2 * cc - 1)));
            auto pw = [](double x, int e) { return e < 0 ? 0.0 : std::pow(x, e); };
            const double mono = pw(X, a) * pw(Y, b) * pw(Z, cc);
            const size_t o_ = (size_t)(o.first + f) * B + p;
            out.val[o_] = N * R * mono;
            if (deriv >= 1) {
              const double mx = a * pw(X, a - 1) * pw(Y, b) * pw(Z, cc);
              const double my = b * pw(X, a) * pw(Y, b - 1) * pw(Z, cc);
              const double mz = cc * pw(X, a) * pw(Y, b) * pw(Z, cc - 1);
              out.gx[o_] = N * (R * mx + dR * X * mono);
              out.gy[o_] = N * (R * my + dR * Y * mono);
              out.gz[o_] = N * (R * mz + dR * Z * mono);
              if (deriv >= 2) {
                // second derivatives of R(r) m(x,y,z): R'' x_i x_j m + R' (d_ij m + x_i m_j + x_j m_i) + R m_ij
                const double mxx = a * (a - 1) * pw(X, a - 2) * pw(Y, b) * pw(Z, cc);
                const double myy = b * (b - 1) * pw(X, a) * pw(Y, b - 2) * pw(Z, cc);
                const double mzz = cc * (cc - 1) * pw(X, a) * pw(Y, b) * pw(Z, cc - 2);
                const double mxy = a * b * pw(X, a - 1) * pw(Y, b - 1) * pw(Z, cc);
                const double mxz = a * cc * pw(X, a - 1) * pw(Y, b) * pw(Z, cc - 1);
                const double myz = b * cc * pw(X, a) * pw(Y, b - 1) * pw(Z, cc - 1);
                out.hxx[o_] = N * (ddR * X * X * mono + dR * (mono + 2 * X * mx) + R * mxx);
                out.hyy[o_] = N * (ddR * Y * Y * mono + dR * (mono + 2 * Y * my) + R * myy);
                out.hzz[o_] = N * (ddR * Z * Z * mono + dR * (mono + 2 * Z * mz) + R * mzz);
                out.hxy[o_] = N * (ddR * X * Y * mono + dR * (X * my + Y * mx) + R * mxy);
                out.hxz[o_] = N * (ddR * X * Z * mono + dR * (X * mz + Z * mx) + R * mxz);
                out.hyz[o_] = N * (ddR * Y * Z * mono + dR * (Y * mz + Z * my) + R * myz);
              }
            }
          }
      } else {
        solid_harmonics(l, X, Y, Z, H);
        for (int m = 0; m < 2 * l + 1; ++m) {
          const size_t o_ = (size_t)(o.first + m) * B + p;
          out.val[o_] = R * H.v[l][m];
          if (deriv >= 1) {
            out.gx[o_] = R * H.x[l][m] + dR * X * H.v[l][m];
            out.gy[o_] = R * H.y[l][m] + dR * Y * H.v[l][m];
            out.gz[o_] = R * H.z[l][m] + dR * Z * H.v[l][m];
          }
          if (deriv >= 2) {
            // d_i d_j (R S) = R S_ij + R' (x_i S_j + x_j S_i) + d_ij R' S + R'' x_i x_j S
            const double S0 = H.v[l][m], Sx = H.x[l][m], Sy = H.y[l][m], Sz = H.z[l][m];
            out.hxx[o_] = R * H.xx[l][m] + dR * (2.0 * X * Sx + S0) + ddR * X * X * S0;
            out.hyy[o_] = R * H.yy[l][m] + dR * (2.0 * Y * Sy + S0) + ddR * Y * Y * S0;
            out.hzz[o_] = R * H.zz[l][m] + dR * (2.0 * Z * Sz + S0) + ddR * Z * Z * S0;
            out.hxy[o_] = R * H.xy[l][m] + dR * (X * Sy + Y * Sx) + ddR * X * Y * S0;
            out.hxz[o_] = R * H.xz[l][m] + dR * (X * Sz + Z * Sx) + ddR * X * Z * S0;
            out.hyz[o_] = R * H.yz[l][m] + dR * (Y * Sz + Z * Sy) + ddR * Y * Z * S0;
          }
        }
      }
    }
  }
}

// ---- functionals: closed forms with hand-derived first derivatives -------
// XCFun's PBE (Serenity's default evaluator): mu = 0.066725 pi^2 / 3; LibXC's:
// mu = 0.2195149727645171 (= beta pi^2 / 3 with the accurate beta).  Both: PW92
// A = 0.0310907, accurate beta in the correlation.  Decided by the full-SCF
// energies of qcserenity/tests/test_tasks.py:50-76 (water/STO-3G PBE, XCFun
// and LibXC runs).
struct FParams {
  double vwn_A = 0.0310907, vwn_b = 3.72744, vwn_c = 12.9352, vwn_x0 = -0.10498;
  double pw_A = 0.0310907, pw_a1 = 0.21370, pw_b1 = 7.5957, pw_b2 = 3.5876, pw_b3 = 1.6382, pw_b4 = 0.49294;
  double kappa = 0.804, mu = 0.066725 * M_PI * M_PI / 3.0, beta = 0.06672455060314922;
  double mu_libxc = 0.2195149727645171;
  // BP86 (Becke 1988 exchange, Perdew 1986 correlation on Perdew-Zunger 1981): the published
  // constants; the prefactor of the P86 Phi is f~ (9 pi)^(1/6), which the reference's BP86 Fock
  // matrix of H2/def2-TZVP (FuncPotential_test.cpp:148-232) reproduces with the original PZ81
  // C and D (1.745 f~ or the continuity-corrected C, D do not)
  double b88_beta = 0.0042;
  double pzU[7] = {-0.1423, 1.0529, 0.3334, 0.0311, -0.048, 0.0020, -0.0116};     // gamma b1 b2 A B C D
  double pzP[7] = {-0.0843, 1.3981, 0.2611, 0.01555, -0.0269, 0.0007, -0.0048};
  double p86_alpha = 0.023266, p86_beta = 7.389e-6, p86_gamma = 8.723, p86_delta = 0.472, p86_c0 = 0.001667,
         p86_c1 = 0.002568, p86_ft = 0.11;
};

// e, de/drho, de/dsigma
struct EV {
  double e, dr, ds;
};

EV slater(double r) {
  const double cx = -0.75 * std::cbrt(3.0 / M_PI);
  const double r13 = std::cbrt(r);
  return {cx * r * r13, (4.0 / 3.0) * cx * r13, 0.0};
}

EV vwn5(double r, const FParams& P) {
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * r));
  const double x = std::sqrt(rs);
  const double b = P.vwn_b, c = P.vwn_c, x0 = P.vwn_x0, A = P.vwn_A;
  const double Q = std::sqrt(4.0 * c - b * b);
  const double X = x * x + b * x + c, X0 = x0 * x0 + b * x0 + c;
  const double at = std::atan(Q / (2.0 * x + b));
  const double eps = A * (std::log(x * x / X) + 2.0 * b / Q * at -
                          b * x0 / X0 * (std::log((x - x0) * (x - x0) / X) + 2.0 * (b + 2.0 * x0) / Q * at));
  const double deps_dx = A * (2.0 / x - (2.0 * x + b) / X - b / X -
                              b * x0 / X0 * (2.0 / (x - x0) - (2.0 * x + b) / X - (b + 2.0 * x0) / X));
  return {r * eps, eps - x / 6.0 * deps_dx, 0.0};
}

void pw92(double rs, const FParams& P, double& eps, double& deps_drs) {
  const double A = P.pw_A;
  const double s = std::sqrt(rs);
  const double Q1 = 2.0 * A * (P.pw_b1 * s + P.pw_b2 * rs + P.pw_b3 * rs * s + P.pw_b4 * rs * rs);
  const double Q1p = 2.0 * A * (0.5 * P.pw_b1 / s + P.pw_b2 + 1.5 * P.pw_b3 * s + 2.0 * P.pw_b4 * rs);
  const double lg = std::log(1.0 + 1.0 / Q1);
  eps = -2.0 * A * (1.0 + P.pw_a1 * rs) * lg;
  deps_drs = -2.0 * A * P.pw_a1 * lg + 2.0 * A * (1.0 + P.pw_a1 * rs) * Q1p / (Q1 * Q1 + Q1);
}

EV pbex(double r, double sig, const FParams& P) {
  const double k = 4.0 * std::pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  const double r83 = std::pow(r, 8.0 / 3.0);
  const double s2 = sig / (k * r83);
  const double den = 1.0 + P.mu * s2 / P.kappa;
  const double F = 1.0 + P.kappa - P.kappa / den;
  const double dF = P.mu / (den * den);
  const EV lda = slater(r);
  return {lda.e * F, lda.dr * F + lda.e * dF * (-8.0 / 3.0) * s2 / r, lda.e * dF / (k * r83)};
}

EV pbec(double r, double sig, const FParams& P) {
  const double gamma = (1.0 - std::log(2.0)) / (M_PI * M_PI);
  const double bg = P.beta / gamma;
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * r));
  double eps, deps_drs;
  pw92(rs, P, eps, deps_drs);
  const double deps = deps_drs * (-rs / (3.0 * r));
  const double kt = M_PI / (16.0 * std::cbrt(3.0 * M_PI * M_PI));
  const double r73 = std::pow(r, 7.0 / 3.0);
  const double y = kt * sig / r73;
  const double ex = std::exp(-eps / gamma);
  const double A = bg / (ex - 1.0);
  const double dA_deps = bg * (ex / gamma) / ((ex - 1.0) * (ex - 1.0));
  const double N = 1.0 + A * y, D = 1.0 + A * y + A * A * y * y;
  const double Phi = bg * y * N / D;
  const double dPhi_dy = bg * (N / D + y * (A * D - N * (A + 2.0 * A * A * y)) / (D * D));
  const double dPhi_dA = bg * y * (y * D - N * (y + 2.0 * A * y * y)) / (D * D);
  const double H = gamma * std::log(1.0 + Phi);
  const double dH_dPhi = gamma / (1.0 + Phi);
  const double dH_dr = dH_dPhi * (dPhi_dy * (-7.0 / 3.0) * y / r + dPhi_dA * dA_deps * deps);
  const double dH_ds = dH_dPhi * dPhi_dy * kt / r73;
  return {r * (eps + H), eps + H + r * (deps + dH_dr), r * dH_ds};
}

EV tf(double r) {
  const double cf = 0.3 * std::pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  const double r23 = std::pow(r, 2.0 / 3.0);
  return {cf * r * r23, (5.0 / 3.0) * cf * r23, 0.0};
}

// Becke 1988 exchange, closed shell (rho_s = rho / 2, |grad rho_s| = |grad rho| / 2): Slater plus
// two spins of -beta rho_s^(4/3) x^2 / Dn, Dn = 1 + 6 beta x asinh x, x = |grad rho_s| / rho_s^(4/3),
// i.e. -(beta / 2) sigma rho_s^(-4/3) / Dn; with q = x Dn'(x) / Dn its derivatives are
// d/dsigma = -(beta / 2) rho_s^(-4/3) (1 - q / 2) / Dn, d/drho = (beta / 3) sigma rho_s^(-7/3) (1 - q) / Dn
EV b88(double r, double sig, const FParams& P) {
  const EV lda = slater(r);
  const double b = P.b88_beta;
  const double ns = 0.5 * r, ns43 = ns * std::cbrt(ns);
  const double x = 0.5 * std::sqrt(sig) / ns43;
  const double Dn = 1.0 + 6.0 * b * x * std::asinh(x);
  const double q = x * 6.0 * b * (std::asinh(x) + x / std::sqrt(1.0 + x * x)) / Dn;
  const double term = -0.5 * b * sig / ns43 / Dn;
  return {lda.e + term, lda.dr + (b * sig / 3.0) / (ns43 * ns) / Dn * (1.0 - q),
          -0.5 * b / ns43 / Dn * (1.0 - 0.5 * q)};
}

// Perdew-Zunger 1981 correlation per particle and its r_s derivative (c = gamma b1 b2 A B C D)
void pz81(double rs, const double* c, double& eps, double& deps) {
  if (rs >= 1.0) {
    const double s = std::sqrt(rs), den = 1.0 + c[1] * s + c[2] * rs;
    eps = c[0] / den;
    deps = -c[0] * (0.5 * c[1] / s + c[2]) / (den * den);
  } else {
    const double l = std::log(rs);
    eps = c[3] * l + c[4] + c[5] * rs * l + c[6] * rs;
    deps = c[3] / rs + c[5] * (l + 1.0) + c[6];
  }
}

EV pz81c(double r, const FParams& P) {
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * r));
  double eps, de;
  pz81(rs, P.pzU, eps, de);
  return {r * eps, eps - rs / 3.0 * de, 0.0};
}

// Perdew 1986 gradient correction, closed shell: e = exp(-Phi) C(n) sigma / n^(4/3),
// Phi = f~ (9 pi)^(1/6) C(inf) / C(n) sqrt(sigma) / n^(7/6), C(inf) = c0 + c1
EV p86(double r, double sig, const FParams& P) {
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * r));
  const double num = P.p86_c1 + P.p86_alpha * rs + P.p86_beta * rs * rs;
  const double den = 1.0 + P.p86_gamma * rs + P.p86_delta * rs * rs + 1.0e4 * P.p86_beta * rs * rs * rs;
  const double Cn = P.p86_c0 + num / den;
  const double dCn_drs = ((P.p86_alpha + 2.0 * P.p86_beta * rs) * den -
                          num * (P.p86_gamma + 2.0 * P.p86_delta * rs + 3.0e4 * P.p86_beta * rs * rs)) /
                         (den * den);
  const double dCn = dCn_drs * (-rs / (3.0 * r));
  const double k = P.p86_ft * std::pow(9.0 * M_PI, 1.0 / 6.0) * (P.p86_c0 + P.p86_c1);
  const double r43 = r * std::cbrt(r);
  const double Phi = k * std::sqrt(sig) / (Cn * std::pow(r, 7.0 / 6.0));
  const double ex = std::exp(-Phi);
  const double e = ex * Cn * sig / r43;
  const double dPhi_dr = Phi * (-dCn / Cn - 7.0 / (6.0 * r));
  return {e, e * (-dPhi_dr + dCn / Cn - 4.0 / (3.0 * r)), ex * Cn / r43 * (1.0 - 0.5 * Phi)};
}

bool is_gga(int f) { return f == 1 || f == 5 || f == 6 || f == 7 || f == 8; }

EV eval(int f, double r, double sig, const FParams& P) {
  EV a{0, 0, 0}, b{0, 0, 0};
  switch (f) {
    case 0: a = slater(r); b = vwn5(r, P); break;
    case 1: a = pbex(r, sig, P); b = pbec(r, sig, P); break;
    case 2: a = tf(r); break;
    case 3: a = slater(r); break;
    case 4: a = vwn5(r, P); break;
    case 5: a = pbex(r, sig, P); break;
    case 6: a = pbec(r, sig, P); break;
    case 7: {
      FParams L = P;
      L.mu = P.mu_libxc;
      a = pbex(r, sig, L);
      b = pbec(r, sig, P);
    } break;
    case 8: {
      a = b88(r, sig, P);
      const EV c = pz81c(r, P), d = p86(r, sig, P);
      b = {c.e + d.e, c.dr + d.dr, c.ds + d.ds};
    } break;
    default: break;
  }
  return {a.e + b.e, a.dr + b.dr, a.ds + b.ds};
}

// ---- spin-polarised functionals (unrestricted) ------------------------------
// A small forward-mode derivative type over (rho_a, rho_b, sigma_aa, sigma_ab,
// sigma_bb), written for this restatement.  Formulas: Slater exchange in its
// spin form -(3/2)(3/(4 pi))^(1/3) sum_s rho_s^(4/3); PBE exchange by spin
// scaling E_x[2 rho_s, 4 sigma_ss] / 2; VWN5 with its spin interpolation
// (paramagnetic, ferromagnetic and spin-stiffness fits); PBE correlation with
// the spin-interpolated PW92 energy, phi(zeta) and t^2 = sigma / (4 phi^2 k_s^2 rho^2).
struct A5 {
  double v, d[5];
};
A5 cst(double v) { return {v, {0, 0, 0, 0, 0}}; }
A5 un(const A5& a, double f, double df) {
  A5 r{f, {}};
  for (int i = 0; i < 5; ++i) r.d[i] = df * a.d[i];
  return r;
}
A5 operator+(const A5& a, const A5& b) {
  A5 r{a.v + b.v, {}};
  for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
A5 operator-(const A5& a, const A5& b) {
  A5 r{a.v - b.v, {}};
  for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
A5 operator*(const A5& a, const A5& b) {
  A5 r{a.v * b.v, {}};
  for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
A5 operator/(const A5& a, const A5& b) {
  A5 r{a.v / b.v, {}};
  for (int i = 0; i < 5; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
A5 operator+(const A5& a, double b) { return a + cst(b); }
A5 operator+(double b, const A5& a) { return cst(b) + a; }
A5 operator-(const A5& a, double b) { return a - cst(b); }
A5 operator-(double b, const A5& a) { return cst(b) - a; }
A5 operator*(double b, const A5& a) { return un(a, a.v * b, b); }
A5 operator/(const A5& a, double b) { return un(a, a.v / b, 1.0 / b); }
A5 operator/(double b, const A5& a) { return cst(b) / a; }
A5 aln(const A5& a) { return un(a, std::log(a.v), 1.0 / a.v); }
A5 aexp(const A5& a) { return un(a, std::exp(a.v), std::exp(a.v)); }
A5 asqrt(const A5& a) { return un(a, std::sqrt(a.v), 0.5 / std::sqrt(a.v)); }
A5 aatan(const A5& a) { return un(a, std::atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
A5 apow(const A5& a, double p) { return un(a, std::pow(a.v, p), a.v > 0 ? p * std::pow(a.v, p - 1.0) : 0.0); }
A5 aasinh(const A5& a) { return un(a, std::asinh(a.v), 1.0 / std::sqrt(1.0 + a.v * a.v)); }

A5 vwn_g(const A5& x, double A, double b, double c, double x0) {
  // VWN fit in x = sqrt(r_s)
  const double Q = std::sqrt(4.0 * c - b * b), X0 = x0 * x0 + b * x0 + c;
  const A5 X = x * x + b * x + c;
  const A5 at = aatan(Q / (2.0 * x + b));
  return A * (aln(x * x / X) + (2.0 * b / Q) * at -
              (b * x0 / X0) * (aln((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / Q) * at));
}
A5 pw_g(const A5& rs, double A, double a1, double b1, double b2, double b3, double b4) {
  const A5 s = asqrt(rs);
  return -2.0 * A * (1.0 + a1 * rs) * aln(1.0 + 1.0 / (2.0 * A * (b1 * s + b2 * rs + b3 * rs * s + b4 * rs * rs)));
}
const double kF2 = 4.0 / (9.0 * (std::cbrt(2.0) - 1.0));  // f''(0)
A5 fzeta(const A5& z) {
  return (apow(1.0 + z, 4.0 / 3.0) + apow(1.0 - z, 4.0 / 3.0) - 2.0) / (2.0 * std::cbrt(2.0) - 2.0);
}
// kind: 0 Slater, 1 TF, 2 PBE exchange with mu_xcfun, 3 PBE exchange with mu_libxc, 4 Becke 88
A5 x_pol(int kind, const A5& ra, const A5& rb, const A5& saa, const A5& sbb, bool aon, bool bon, const FParams& P) {
  A5 e = cst(0.0);
  for (int s = 0; s < 2; ++s) {
    if (!(s == 0 ? aon : bon)) continue;
    const A5& r = s == 0 ? ra : rb;
    const A5& sg = s == 0 ? saa : sbb;
    if (kind == 0) {
      e = e + (-1.5 * std::cbrt(3.0 / (4.0 * M_PI))) * apow(r, 4.0 / 3.0);
    } else if (kind == 1) {
      e = e + 0.5 * (0.3 * std::pow(3.0 * M_PI * M_PI, 2.0 / 3.0)) * apow(2.0 * r, 5.0 / 3.0);
    } else if (kind == 4) {
      // -C rho_s^(4/3) - beta rho_s^(4/3) x^2 / (1 + 6 beta x asinh x), x^2 = sigma_ss / rho_s^(8/3)
      const A5 r43 = apow(r, 4.0 / 3.0);
      const A5 u = sg / (r43 * r43);
      A5 xas = u - u * u / 6.0;  // x asinh x for tiny x
      if (u.v >= 1e-12) {
        const A5 x = asqrt(u);
        xas = x * aasinh(x);
      }
      e = e + (-1.5 * std::cbrt(3.0 / (4.0 * M_PI))) * r43 - P.b88_beta * (sg / r43) / (1.0 + 6.0 * P.b88_beta * xas);
    } else {
      const double mu = kind == 2 ? P.mu : P.mu_libxc;
      const A5 r2 = 2.0 * r, s4 = 4.0 * sg;
      const A5 s2 = s4 / (4.0 * std::pow(3.0 * M_PI * M_PI, 2.0 / 3.0) * apow(r2, 8.0 / 3.0));
      const A5 F = (1.0 + P.kappa) - P.kappa / (1.0 + mu * s2 / P.kappa);
      e = e + 0.5 * (-0.75 * std::cbrt(3.0 / M_PI)) * apow(r2, 4.0 / 3.0) * F;
    }
  }
  return e;
}
A5 vwn_pol(const A5& ra, const A5& rb, const FParams& P) {
  const A5 n = ra + rb, z = (ra - rb) / n;
  const A5 x = asqrt(apow(3.0 / (4.0 * M_PI * n), 1.0 / 3.0));
  const A5 eP = vwn_g(x, P.vwn_A, P.vwn_b, P.vwn_c, P.vwn_x0);
  const A5 eF = vwn_g(x, 0.01554535, 7.06042, 18.0578, -0.32500);
  const A5 aS = vwn_g(x, -1.0 / (6.0 * M_PI * M_PI), 1.13107, 13.0045, -0.0047584);
  const A5 f = fzeta(z), z4 = z * z * z * z;
  return n * (eP + aS * f / kF2 * (1.0 - z4) + (eF - eP) * f * z4);
}
A5 pbec_pol(const A5& ra, const A5& rb, const A5& sig, const FParams& P) {
  const A5 n = ra + rb, z = (ra - rb) / n;
  const A5 rs = apow(3.0 / (4.0 * M_PI * n), 1.0 / 3.0);
  const A5 f = fzeta(z), z4 = z * z * z * z;
  const A5 e0 = pw_g(rs, P.pw_A, P.pw_a1, P.pw_b1, P.pw_b2, P.pw_b3, P.pw_b4);
  const A5 e1 = pw_g(rs, 0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  const A5 ac = -1.0 * pw_g(rs, 0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);
  const A5 ec = e0 + ac * f / kF2 * (1.0 - z4) + (e1 - e0) * f * z4;
  A5 opz = 1.0 + z, omz = 1.0 - z;
  if (opz.v < 1e-12) opz.v = 1e-12;
  if (omz.v < 1e-12) omz.v = 1e-12;
  const A5 phi = 0.5 * (apow(opz, 2.0 / 3.0) + apow(omz, 2.0 / 3.0));
  const A5 kF = apow(3.0 * M_PI * M_PI * n, 1.0 / 3.0);
  const A5 ks2 = 4.0 * kF / M_PI;
  const A5 t2 = sig / (4.0 * phi * phi * ks2 * n * n);
  const double gam = (1.0 - std::log(2.0)) / (M_PI * M_PI);
  const A5 phi3 = phi * phi * phi;
  const A5 A = (P.beta / gam) / (aexp(-1.0 * ec / (gam * phi3)) - 1.0);
  const A5 H = gam * phi3 * aln(1.0 + (P.beta / gam) * t2 * (1.0 + A * t2) / (1.0 + A * t2 + A * A * t2 * t2));
  return n * (ec + H);
}
A5 pz_g(const A5& rs, const double* c) {
  if (rs.v >= 1.0) return c[0] / (1.0 + c[1] * asqrt(rs) + c[2] * rs);
  const A5 l = aln(rs);
  return c[3] * l + c[4] + c[5] * rs * l + c[6] * rs;
}
// PZ81 with the spin interpolation f(zeta), and the P86 correction of the total density with
// d(zeta) = 2^(1/3) sqrt(((1+zeta)/2)^(5/3) + ((1-zeta)/2)^(5/3))
A5 bp86c_pol(const A5& ra, const A5& rb, const A5& sig, const FParams& P) {
  const A5 n = ra + rb, z = (ra - rb) / n;
  const A5 rs = apow(3.0 / (4.0 * M_PI * n), 1.0 / 3.0);
  const A5 eU = pz_g(rs, P.pzU), eP = pz_g(rs, P.pzP);
  const A5 dz = std::cbrt(2.0) * asqrt(apow(0.5 * (1.0 + z), 5.0 / 3.0) + apow(0.5 * (1.0 - z), 5.0 / 3.0));
  const A5 Cn = P.p86_c0 + (P.p86_c1 + P.p86_alpha * rs + P.p86_beta * rs * rs) /
                               (1.0 + P.p86_gamma * rs + P.p86_delta * rs * rs + 1.0e4 * P.p86_beta * rs * rs * rs);
  const double k = P.p86_ft * std::pow(9.0 * M_PI, 1.0 / 6.0) * (P.p86_c0 + P.p86_c1);
  const A5 g = sig.v > 0.0 ? asqrt(sig) : 0.0 * sig;
  const A5 Phi = k * g / (Cn * apow(n, 7.0 / 6.0));
  return n * (eU + fzeta(z) * (eP - eU)) + aexp(-1.0 * Phi) * Cn * sig / (dz * apow(n, 4.0 / 3.0));
}
A5 eval_pol(int f, const A5& ra, const A5& rb, const A5& saa, const A5& sab, const A5& sbb, bool aon, bool bon,
            const FParams& P) {
  const A5 sig = saa + 2.0 * sab + sbb;
  const A5 rac = aon ? ra : cst(0.0), rbc = bon ? rb : cst(0.0);
  switch (f) {
    case 0: return x_pol(0, ra, rb, saa, sbb, aon, bon, P) + vwn_pol(rac, rbc, P);
    case 1: return x_pol(2, ra, rb, saa, sbb, aon, bon, P) + pbec_pol(rac, rbc, sig, P);
    case 2: return x_pol(1, ra, rb, saa, sbb, aon, bon, P);
    case 3: return x_pol(0, ra, rb, saa, sbb, aon, bon, P);
    case 4: return vwn_pol(rac, rbc, P);
    case 5: return x_pol(2, ra, rb, saa, sbb, aon, bon, P);
    case 6: return pbec_pol(rac, rbc, sig, P);
    case 7: return x_pol(3, ra, rb, saa, sbb, aon, bon, P) + pbec_pol(rac, rbc, sig, P);
    case 8: return x_pol(4, ra, rb, saa, sbb, aon, bon, P) + bp86c_pol(rac, rbc, sig, P);
    default: return cst(0.0);
  }
}

FParams params_from(const double* p) {
  FParams P;
  if (p) {
    P.vwn_A = p[0];
    P.pw_A = p[1];
    P.kappa = p[2];
    P.mu = p[3];
    P.beta = p[4];
  }
  return P;
}

// The two dense block contractions, register-blocked.  Every output element is
// summed in the same order as the plain triple loop (k ascending for BP, p
// ascending for T), so the blocking changes speed only, not results.
//
// BP[j][p] += sum_k phi[sig[k]][p] Ps[k][j]   (phi AO-major with B points per AO, BP j-major)
void phi_times_p(const double* phi, const std::vector<int>& sig, int B, const double* Ps, int m, double* BP) {
  constexpr int JB = 4, PB = 8;
  int j0 = 0;
  for (; j0 + JB <= m; j0 += JB) {
    int p0 = 0;
    for (; p0 + PB <= B; p0 += PB) {
      double acc[JB][PB];
      for (int u = 0; u < JB; ++u)
        for (int q = 0; q < PB; ++q) acc[u][q] = BP[(size_t)(j0 + u) * B + p0 + q];
      for (int k = 0; k < m; ++k) {
        const double* ph = phi + (size_t)sig[k] * B + p0;
        for (int u = 0; u < JB; ++u) {
          const double pk = Ps[k + (size_t)m * (j0 + u)];
          for (int q = 0; q < PB; ++q) acc[u][q] += ph[q] * pk;
        }
      }
      for (int u = 0; u < JB; ++u)
        for (int q = 0; q < PB; ++q) BP[(size_t)(j0 + u) * B + p0 + q] = acc[u][q];
    }
    for (int u = 0; u < JB; ++u)
      for (int k = 0; k < m; ++k) {
        const double pk = Ps[k + (size_t)m * (j0 + u)];
        const double* ph = phi + (size_t)sig[k] * B;
        double* bp = BP + (size_t)(j0 + u) * B;
        for (int p = p0; p < B; ++p) bp[p] += ph[p] * pk;
      }
  }
  for (int j = j0; j < m; ++j)
    for (int k = 0; k < m; ++k) {
      const double pk = Ps[k + (size_t)m * j];
      const double* ph = phi + (size_t)sig[k] * B;
      double* bp = BP + (size_t)j * B;
      for (int p = 0; p < B; ++p) bp[p] += ph[p] * pk;
    }
}

// T[i][j] = sum_p phi[sig[i]][p] G[j][p]   (G j-major with B points per column; T i-major m x m)
void phi_t_times_g(const double* phi, const std::vector<int>& sig, int B, const double* G, int m, double* T) {
  constexpr int IB = 4, JB = 8;
  std::vector<double> Gt((size_t)B * m);  // point-major copy of G
  for (int j = 0; j < m; ++j)
    for (int p = 0; p < B; ++p) Gt[(size_t)p * m + j] = G[(size_t)j * B + p];
  for (int i0 = 0; i0 < m; i0 += IB) {
    const int ni = std::min(IB, m - i0);
    for (int j0 = 0; j0 < m; j0 += JB) {
      const int nj = std::min(JB, m - j0);
      double acc[IB][JB] = {};
      if (ni == IB && nj == JB) {
        for (int p = 0; p < B; ++p) {
          const double* g = &Gt[(size_t)p * m + j0];
          for (int u = 0; u < IB; ++u) {
            const double f = phi[(size_t)sig[i0 + u] * B + p];
            for (int v = 0; v < JB; ++v) acc[u][v] += f * g[v];
          }
        }
      } else {
        for (int p = 0; p < B; ++p)
          for (int u = 0; u < ni; ++u) {
            const double f = phi[(size_t)sig[i0 + u] * B + p];
            for (int v = 0; v < nj; ++v) acc[u][v] += f * Gt[(size_t)p * m + j0 + v];
          }
      }
      for (int u = 0; u < ni; ++u)
        for (int v = 0; v < nj; ++v) T[(size_t)(i0 + u) * m + j0 + v] = acc[u][v];
    }
  }
}

}  // namespace

extern "C" {

// AO data of points[0..B) as one block: values (B x n col-major), gradients,
// Hessian (Cartesian shells only; deriv = 2), negligible flags.
int oracle_ao_block(int nshell, const int* l, const int* pure, const int* nprim, const double* center,
                    const double* alpha, const double* coef, const double* pts, int B, double thr, int deriv,
                    double* val, double* gx, double* gy, double* gz, double* hess /*6 x B x n*/, int* neg) {
  Basis bas(nshell, l, pure, nprim, center, alpha, coef);
  BlockAO ao;
  ao_block(bas, pts, B, thr, deriv, ao);
  const size_t sz = (size_t)B * bas.nao;
  std::memcpy(val, ao.val.data(), sz * sizeof(double));
  if (deriv >= 1 && gx) {
    std::memcpy(gx, ao.gx.data(), sz * sizeof(double));
    std::memcpy(gy, ao.gy.data(), sz * sizeof(double));
    std::memcpy(gz, ao.gz.data(), sz * sizeof(double));
  }
  if (deriv >= 2 && hess) {
    const std::vector<double>* h[6] = {&ao.hxx, &ao.hxy, &ao.hxz, &ao.hyy, &ao.hyz, &ao.hzz};
    for (int c = 0; c < 6; ++c) std::memcpy(hess + c * sz, h[c]->data(), sz * sizeof(double));
  }
  if (neg) std::memcpy(neg, ao.neg.data(), bas.nao * sizeof(int));
  return 0;
}

// rho (N), grad (3 x N SoA, nullable), hess (6 x N SoA, nullable; needs deriv 2
// and Cartesian shells), nonneg (nblocks, nullable).
int oracle_density(int nshell, const int* l, const int* pure, const int* nprim, const double* center,
                   const double* alpha, const double* coef, const double* pts, int64_t N, int bs, double thr,
                   const double* P, double* rho, double* grad, double* hess, int* nonneg) {
  Basis bas(nshell, l, pure, nprim, center, alpha, coef);
  const int n = bas.nao;
  const int64_t nb = (N + bs - 1) / bs;
  const int deriv = hess ? 2 : (grad ? 1 : 0);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t first = b * bs;
    const int B = (int)std::min<int64_t>(bs, N - first);
    BlockAO ao;
    ao_block(bas, pts + 3 * first, B, thr, deriv, ao);
    std::vector<int> sig;
    for (int i = 0; i < n; ++i)
      if (!ao.neg[i]) sig.push_back(i);
    for (int p = 0; p < B; ++p) {
      rho[first + p] = 0.0;
      if (grad)
        for (int d = 0; d < 3; ++d) grad[d * N + first + p] = 0.0;
      if (hess)
        for (int d = 0; d < 6; ++d) hess[d * N + first + p] = 0.0;
    }
    if (nonneg) nonneg[b] = sig.empty() ? 0 : 1;
    if (sig.empty()) continue;
    const int m = (int)sig.size();
    // Psub = P[sig, sig];  BP = phi_sub Psub  (B x m)
    std::vector<double> Ps((size_t)m * m), BP((size_t)B * m, 0.0);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) Ps[i + (size_t)m * j] = P[sig[i] + (size_t)n * sig[j]];
    phi_times_p(ao.val.data(), sig, B, Ps.data(), m, BP.data());
    std::vector<double> GX, GY, GZ;  // grad phi . Psub for the Hessian
    if (hess) {
      GX.assign((size_t)B * m, 0.0);
      GY.assign((size_t)B * m, 0.0);
      GZ.assign((size_t)B * m, 0.0);
      for (int j = 0; j < m; ++j)
        for (int k = 0; k < m; ++k) {
          const double pkj = Ps[k + (size_t)m * j];
          for (int p = 0; p < B; ++p) {
            GX[(size_t)j * B + p] += ao.gx[(size_t)sig[k] * B + p] * pkj;
            GY[(size_t)j * B + p] += ao.gy[(size_t)sig[k] * B + p] * pkj;
            GZ[(size_t)j * B + p] += ao.gz[(size_t)sig[k] * B + p] * pkj;
          }
        }
    }
    for (int p = 0; p < B; ++p) {
      double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
      double hxx = 0, hxy = 0, hxz = 0, hyy = 0, hyz = 0, hzz = 0;
      for (int j = 0; j < m; ++j) {
        const size_t o = (size_t)sig[j] * B + p;
        const double t = BP[(size_t)j * B + p];
        r += t * ao.val[o];
        if (grad || hess) {
          x += t * ao.gx[o];
          y += t * ao.gy[o];
          z += t * ao.gz[o];
        }
        if (hess) {
          const size_t q = (size_t)j * B + p;
          hxx += t * ao.hxx[o] + GX[q] * ao.gx[o];
          hxy += t * ao.hxy[o] + GX[q] * ao.gy[o];
          hxz += t * ao.hxz[o] + GX[q] * ao.gz[o];
          hyy += t * ao.hyy[o] + GY[q] * ao.gy[o];
          hyz += t * ao.hyz[o] + GY[q] * ao.gz[o];
          hzz += t * ao.hzz[o] + GZ[q] * ao.gz[o];
        }
      }
      rho[first + p] = r;
      if (grad) {
        grad[first + p] = 2.0 * x;
        grad[N + first + p] = 2.0 * y;
        grad[2 * N + first + p] = 2.0 * z;
      }
      if (hess) {
        const double hv[6] = {hxx, hxy, hxz, hyy, hyz, hzz};
        for (int d = 0; d < 6; ++d) hess[d * N + first + p] = 2.0 * hv[d];
      }
    }
  }
  return 0;
}

// Functional on N points in blocks of bs: eps, v, g (3 x N, nullable), energy.
// params: NULL or {vwn_A, pw_A, kappa, mu, beta} (mu of the XCFun form; id 7 keeps LibXC's mu).
int oracle_functional(int func, int64_t N, int bs, const double* rho, const double* grad, const double* w,
                      const double* params, double* eps, double* v, double* g, double* energy) {
  const FParams P = params_from(params);
  const bool gga = is_gga(func);
  const int64_t nb = (N + bs - 1) / bs;
  std::vector<double> eb(nb, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t first = b * bs;
    const int B = (int)std::min<int64_t>(bs, N - first);
    double s = 0.0;
    for (int p = 0; p < B; ++p) s += std::fabs(rho[first + p]);
    const bool skip = s < B * 1e-12;
    double e_acc = 0.0;
    for (int p = 0; p < B; ++p) {
      const int64_t q = first + p;
      double e = 0.0, vv = 0.0, gg[3] = {0, 0, 0};
      const double r = rho[q];
      if (!skip && r >= 1e-14) {
        double gx = 0, gy = 0, gz = 0;
        if (gga) {
          gx = grad[q];
          gy = grad[N + q];
          gz = grad[2 * N + q];
        }
        const EV ev = eval(func, r, gx * gx + gy * gy + gz * gz, P);
        e = ev.e;
        vv = ev.dr;
        if (gga) {
          gg[0] = 2.0 * ev.ds * gx;
          gg[1] = 2.0 * ev.ds * gy;
          gg[2] = 2.0 * ev.ds * gz;
        }
      }
      if (eps) eps[q] = e;
      if (v) v[q] = vv;
      if (g) {
        g[q] = gg[0];
        g[N + q] = gg[1];
        g[2 * N + q] = gg[2];
      }
      e_acc += w[q] * e;
    }
    eb[b] = e_acc;
  }
  double E = 0.0;
  for (int64_t b = 0; b < nb; ++b) E += eb[b];
  if (energy) *energy = E;
  return 0;
}

// F += sum_b phi_b^T G_b + G_b^T phi_b (addBlock semantics); v: N, g: 3 x N or NULL.
// Blocks [block_begin, block_end) only (block_end < 0: all).
int oracle_fock_add(int nshell, const int* l, const int* pure, const int* nprim, const double* center,
                    const double* alpha, const double* coef, const double* pts, const double* w, int64_t N, int bs,
                    double thr, double block_ave_thr, const double* v, const double* g, double* F,
                    int64_t block_begin, int64_t block_end) {
  Basis bas(nshell, l, pure, nprim, center, alpha, coef);
  const int n = bas.nao;
  const int64_t nb = (N + bs - 1) / bs;
  if (block_end < 0 || block_end > nb) block_end = nb;
  const int nthreads = omp_get_max_threads();
  std::vector<std::vector<double>> Ft(nthreads, std::vector<double>((size_t)n * n, 0.0));
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = block_begin; b < block_end; ++b) {
    auto& Fm = Ft[omp_get_thread_num()];
    const int64_t first = b * bs;
    const int B = (int)std::min<int64_t>(bs, N - first);
    std::vector<double> s(B), gw[3];
    double avg = 0.0;
    for (int p = 0; p < B; ++p) {
      s[p] = v[first + p] * w[first + p];
      avg += std::fabs(s[p]);
    }
    if (g) {
      for (int d = 0; d < 3; ++d) {
        gw[d].resize(B);
        for (int p = 0; p < B; ++p) {
          gw[d][p] = w[first + p] * g[d * N + first + p];
          avg += std::fabs(gw[d][p]);
        }
      }
    }
    if (avg / B < block_ave_thr) continue;
    // the reference fetches the block twice (controllers A and B), recomputing the AOs
    BlockAO ao, aoB;
    ao_block(bas, pts + 3 * first, B, thr, g ? 1 : 0, ao);
    ao_block(bas, pts + 3 * first, B, thr, g ? 1 : 0, aoB);
    std::vector<int> sig;
    for (int i = 0; i < n; ++i)
      if (!ao.neg[i]) sig.push_back(i);
    const int m = (int)sig.size();
    if (m == 0) continue;
    // G (B x m)
    std::vector<double> G((size_t)B * m);
    for (int j = 0; j < m; ++j)
      for (int p = 0; p < B; ++p) {
        const size_t o = (size_t)sig[j] * B + p;
        double gv = 0.5 * s[p] * ao.val[o];
        if (g) gv += gw[0][p] * ao.gx[o] + gw[1][p] * ao.gy[o] + gw[2][p] * ao.gz[o];
        G[(size_t)j * B + p] = gv;
      }
    // T = phi^T G (m x m); F[sig,sig] += T + T^T
    std::vector<double> T((size_t)m * m);
    phi_t_times_g(aoB.val.data(), sig, B, G.data(), m, T.data());
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        const double t = T[(size_t)i * m + j];
        Fm[sig[i] + (size_t)n * sig[j]] += t;
        Fm[sig[j] + (size_t)n * sig[i]] += t;
      }
  }
  for (int t = 0; t < nthreads; ++t)
    for (size_t e = 0; e < (size_t)n * n; ++e) F[e] += Ft[t][e];
  return 0;
}

// FuncPotential::getMatrix/getEnergy: F = V_xc (overwritten), energy.
// Optional outputs rho/grad (N, 3N) for inspection.
int oracle_fock_build(int nshell, const int* l, const int* pure, const int* nprim, const double* center,
                      const double* alpha, const double* coef, const double* pts, const double* w, int64_t N, int bs,
                      double thr, double block_ave_thr, int func, const double* params, const double* P, double* F,
                      double* energy, double* rho_out, double* grad_out) {
  const bool gga = is_gga(func);
  std::vector<double> rho(N), grad(gga ? 3 * N : 0), v(N), g(gga ? 3 * N : 0);
  oracle_density(nshell, l, pure, nprim, center, alpha, coef, pts, N, bs, thr, P, rho.data(),
                 gga ? grad.data() : nullptr, nullptr, nullptr);
  oracle_functional(func, N, 128, rho.data(), gga ? grad.data() : nullptr, w, params, nullptr, v.data(),
                    gga ? g.data() : nullptr, energy);
  Basis bas(nshell, l, pure, nprim, center, alpha, coef);
  std::fill(F, F + (size_t)bas.nao * bas.nao, 0.0);
  oracle_fock_add(nshell, l, pure, nprim, center, alpha, coef, pts, w, N, bs, thr, block_ave_thr, v.data(),
                  gga ? g.data() : nullptr, F, 0, -1);
  if (rho_out) std::copy(rho.begin(), rho.end(), rho_out);
  if (grad_out && gga) std::copy(grad.begin(), grad.end(), grad_out);
  return 0;
}

// FuncPotential::getGeomGradients (src/potentials/FuncPotential.cpp:113-239): the XC part of the
// nuclear gradient at a fixed grid (no weight derivatives), natoms x 3 column-major as Eigen
// stores it.  ao_atom[nu] is the atom of AO nu (BasisController::getAtomIndicesOfBasis).  For
// every block and every pair of non-negligible AOs (mu, nu):
//   grad[atom(nu)][c] -= 2 P_mu,nu sum_p w v phi_mu d_c phi_nu
//   (GGA)             -= 2 P_mu,nu sum_p w sum_k g_k (phi_mu d_k d_c phi_nu + d_k phi_mu d_c phi_nu)
int oracle_xc_gradient(int nshell, const int* l, const int* pure, const int* nprim, const double* center,
                       const double* alpha, const double* coef, const double* pts, const double* w, int64_t N,
                       int bs, double thr, int func, const double* params, const double* P, int natoms,
                       const int* ao_atom, double* grad, double* energy) {
  const bool gga = is_gga(func);
  std::vector<double> rho(N), gr(gga ? 3 * N : 0), v(N), g(gga ? 3 * N : 0);
  oracle_density(nshell, l, pure, nprim, center, alpha, coef, pts, N, bs, thr, P, rho.data(),
                 gga ? gr.data() : nullptr, nullptr, nullptr);
  oracle_functional(func, N, 128, rho.data(), gga ? gr.data() : nullptr, w, params, nullptr, v.data(),
                    gga ? g.data() : nullptr, energy);
  Basis bas(nshell, l, pure, nprim, center, alpha, coef);
  const int n = bas.nao;
  const int64_t nb = (N + bs - 1) / bs;
  const int nthreads = omp_get_max_threads();
  std::vector<std::vector<double>> Gt(nthreads, std::vector<double>((size_t)natoms * 3, 0.0));
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < nb; ++b) {
    auto& gp = Gt[omp_get_thread_num()];
    const int64_t first = b * bs;
    const int B = (int)std::min<int64_t>(bs, N - first);
    BlockAO ao;
    ao_block(bas, pts + 3 * first, B, thr, gga ? 2 : 1, ao);
    std::vector<double> potMu(B), pre(B);
    for (int mu = 0; mu < n; ++mu) {
      if (ao.neg[mu]) continue;
      const double* fm = &ao.val[(size_t)mu * B];
      for (int p = 0; p < B; ++p) potMu[p] = w[first + p] * v[first + p] * fm[p];
      for (int nu = 0; nu < n; ++nu) {
        if (ao.neg[nu]) continue;
        const int A = ao_atom[nu];
        const double prefac = 2.0 * P[mu + (size_t)n * nu];
        const size_t o = (size_t)nu * B;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int p = 0; p < B; ++p) {
          sx += potMu[p] * ao.gx[o + p];
          sy += potMu[p] * ao.gy[o + p];
          sz += potMu[p] * ao.gz[o + p];
        }
        gp[A] -= prefac * sx;
        gp[A + natoms] -= prefac * sy;
        gp[A + 2 * natoms] -= prefac * sz;
        if (gga) {
          const size_t om = (size_t)mu * B;
          double tx = 0.0, ty = 0.0, tz = 0.0;
          for (int p = 0; p < B; ++p) {
            const double c = prefac * w[first + p];
            const double px = c * g[first + p], py = c * g[N + first + p], pz = c * g[2 * N + first + p];
            const double f = fm[p];
            const double mx = ao.gx[om + p], my = ao.gy[om + p], mz = ao.gz[om + p];
            const double nx = ao.gx[o + p], ny = ao.gy[o + p], nz = ao.gz[o + p];
            tx += px * (f * ao.hxx[o + p] + mx * nx) + py * (f * ao.hxy[o + p] + my * nx) +
                  pz * (f * ao.hxz[o + p] + mz * nx);
            ty += px * (f * ao.hxy[o + p] + mx * ny) + py * (f * ao.hyy[o + p] + my * ny) +
                  pz * (f * ao.hyz[o + p] + mz * ny);
            tz += px * (f * ao.hxz[o + p] + mx * nz) + py * (f * ao.hyz[o + p] + my * nz) +
                  pz * (f * ao.hzz[o + p] + mz * nz);
          }
          gp[A] -= tx;
          gp[A + natoms] -= ty;
          gp[A + 2 * natoms] -= tz;
        }
      }
    }
  }
  std::fill(grad, grad + (size_t)natoms * 3, 0.0);
  for (int t = 0; t < nthreads; ++t)
    for (size_t e = 0; e < (size_t)natoms * 3; ++e) grad[e] += Gt[t][e];
  return 0;
}

// Unrestricted XCFun/LibXC calcData (XC_A_B_AX_AY_AZ_BX_BY_BZ): the block is skipped when
// both spin densities are negligible (XCFun.cpp:136-140 with for_spin); a point counts when
// rho_a + rho_b >= 1e-14, a spin channel below 1e-14 is dropped.  g_s = d eps / d grad rho_s.
int oracle_functional_pol(int func, int64_t N, int bs, const double* ra, const double* rb, const double* gra,
                          const double* grb, const double* w, const double* params, double* eps, double* va,
                          double* vb, double* ga, double* gb, double* energy) {
  const FParams P = params_from(params);
  const bool gga = is_gga(func);
  const int64_t nb = (N + bs - 1) / bs;
  std::vector<double> eb(nb, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t first = b * bs;
    const int B = (int)std::min<int64_t>(bs, N - first);
    double sa = 0.0, sb = 0.0;
    for (int p = 0; p < B; ++p) {
      sa += std::fabs(ra[first + p]);
      sb += std::fabs(rb[first + p]);
    }
    const bool skip = sa < B * 1e-12 && sb < B * 1e-12;
    double e_acc = 0.0;
    for (int p = 0; p < B; ++p) {
      const int64_t q = first + p;
      double e = 0.0, v2[2] = {0, 0}, g6[6] = {0, 0, 0, 0, 0, 0};
      const double a = ra[q], bb = rb[q];
      const bool aon = a >= 1e-14, bon = bb >= 1e-14;
      if (!skip && a + bb >= 1e-14 && (aon || bon)) {
        double x[3] = {0, 0, 0}, y[3] = {0, 0, 0};
        if (gga)
          for (int c = 0; c < 3; ++c) {
            x[c] = gra[c * N + q];
            y[c] = grb[c * N + q];
          }
        A5 in[5];
        const double vals[5] = {aon ? a : 0.0, bon ? bb : 0.0, x[0] * x[0] + x[1] * x[1] + x[2] * x[2],
                                x[0] * y[0] + x[1] * y[1] + x[2] * y[2], y[0] * y[0] + y[1] * y[1] + y[2] * y[2]};
        for (int k = 0; k < 5; ++k) {
          in[k] = cst(vals[k]);
          in[k].d[k] = 1.0;
        }
        const A5 r = eval_pol(func, in[0], in[1], in[2], in[3], in[4], aon, bon, P);
        e = r.v;
        v2[0] = r.d[0];
        v2[1] = r.d[1];
        if (gga)
          for (int c = 0; c < 3; ++c) {
            g6[c] = 2.0 * r.d[2] * x[c] + r.d[3] * y[c];
            g6[3 + c] = 2.0 * r.d[4] * y[c] + r.d[3] * x[c];
          }
      }
      if (eps) eps[q] = e;
      if (va) va[q] = v2[0];
      if (vb) vb[q] = v2[1];
      for (int c = 0; c < 3; ++c) {
        if (ga) ga[c * N + q] = g6[c];
        if (gb) gb[c * N + q] = g6[3 + c];
      }
      e_acc += w[q] * e;
    }
    eb[b] = e_acc;
  }
  double E = 0.0;
  for (int64_t b = 0; b < nb; ++b) E += eb[b];
  if (energy) *energy = E;
  return 0;
}

// FuncPotential<UNRESTRICTED>::getMatrix / getEnergy: F_a, F_b overwritten.
int oracle_fock_build_pol(int nshell, const int* l, const int* pure, const int* nprim, const double* center,
                          const double* alpha, const double* coef, const double* pts, const double* w, int64_t N,
                          int bs, double thr, double block_ave_thr, int func, const double* params, const double* Pa,
                          const double* Pb, double* Fa, double* Fb, double* energy) {
  const bool gga = is_gga(func);
  std::vector<double> ra(N), rb(N), gra(gga ? 3 * N : 0), grb(gga ? 3 * N : 0), va(N), vb(N), ga(gga ? 3 * N : 0),
      gb(gga ? 3 * N : 0);
  oracle_density(nshell, l, pure, nprim, center, alpha, coef, pts, N, bs, thr, Pa, ra.data(),
                 gga ? gra.data() : nullptr, nullptr, nullptr);
  oracle_density(nshell, l, pure, nprim, center, alpha, coef, pts, N, bs, thr, Pb, rb.data(),
                 gga ? grb.data() : nullptr, nullptr, nullptr);
  oracle_functional_pol(func, N, 128, ra.data(), rb.data(), gga ? gra.data() : nullptr, gga ? grb.data() : nullptr,
                        w, params, nullptr, va.data(), vb.data(), gga ? ga.data() : nullptr,
                        gga ? gb.data() : nullptr, energy);
  Basis bas(nshell, l, pure, nprim, center, alpha, coef);
  std::fill(Fa, Fa + (size_t)bas.nao * bas.nao, 0.0);
  std::fill(Fb, Fb + (size_t)bas.nao * bas.nao, 0.0);
  oracle_fock_add(nshell, l, pure, nprim, center, alpha, coef, pts, w, N, bs, thr, block_ave_thr, va.data(),
                  gga ? ga.data() : nullptr, Fa, 0, -1);
  oracle_fock_add(nshell, l, pure, nprim, center, alpha, coef, pts, w, N, bs, thr, block_ave_thr, vb.data(),
                  gga ? gb.data() : nullptr, Fb, 0, -1);
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
