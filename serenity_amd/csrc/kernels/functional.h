// Exchange-correlation functionals of the XC Fock build, FP64, evaluated per
// grid point with forward-mode derivatives (value, d/d rho, d/d sigma) so the
// potential terms come from the same code as the energy.
//
// The formulas restate the basic functionals Serenity takes from XCFun for
// the LDA and PBE composites (functional_definitions.dat:44,47;
// basic_functionals.dat:35,78,99,168,230): Slater exchange, VWN5 correlation
// (Vosko-Wilk-Nusair 1980, paramagnetic fit), PBE exchange and correlation
// (Perdew-Burke-Ernzerhof 1996 with the Perdew-Wang 1992 LDA correlation) and
// Thomas-Fermi kinetic energy.  Inputs are restricted (closed-shell) total
// density rho and sigma = |grad rho|^2; XCFun's XC_N_NX_NY_NZ outputs are then
//   eps (energy per volume), v = d eps/d rho, g = 2 d eps/d sigma * grad rho.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define SXC_HD __host__ __device__ inline
#else
#define SXC_HD inline
#endif

namespace sxc {

// Value and first partials with respect to (rho, sigma).
struct D2 {
  double v, dr, ds;
};

SXC_HD D2 mk(double v) { return {v, 0.0, 0.0}; }
SXC_HD D2 operator+(D2 a, D2 b) { return {a.v + b.v, a.dr + b.dr, a.ds + b.ds}; }
SXC_HD D2 operator-(D2 a, D2 b) { return {a.v - b.v, a.dr - b.dr, a.ds - b.ds}; }
SXC_HD D2 operator-(D2 a) { return {-a.v, -a.dr, -a.ds}; }
SXC_HD D2 operator*(D2 a, D2 b) { return {a.v * b.v, a.dr * b.v + a.v * b.dr, a.ds * b.v + a.v * b.ds}; }
SXC_HD D2 operator/(D2 a, D2 b) {
  const double ib = 1.0 / b.v, q = a.v * ib;
  return {q, (a.dr - q * b.dr) * ib, (a.ds - q * b.ds) * ib};
}
SXC_HD D2 operator+(D2 a, double b) { return {a.v + b, a.dr, a.ds}; }
SXC_HD D2 operator+(double b, D2 a) { return {a.v + b, a.dr, a.ds}; }
SXC_HD D2 operator-(D2 a, double b) { return {a.v - b, a.dr, a.ds}; }
SXC_HD D2 operator-(double b, D2 a) { return {b - a.v, -a.dr, -a.ds}; }
SXC_HD D2 operator*(D2 a, double b) { return {a.v * b, a.dr * b, a.ds * b}; }
SXC_HD D2 operator*(double b, D2 a) { return {a.v * b, a.dr * b, a.ds * b}; }
SXC_HD D2 operator/(D2 a, double b) { return {a.v / b, a.dr / b, a.ds / b}; }
SXC_HD D2 operator/(double b, D2 a) {
  const double q = b / a.v, d = -q / a.v;
  return {q, d * a.dr, d * a.ds};
}
SXC_HD D2 chain(D2 a, double f, double df) { return {f, df * a.dr, df * a.ds}; }
SXC_HD D2 dlog(D2 a) { return chain(a, log(a.v), 1.0 / a.v); }
SXC_HD D2 dexpm1(D2 a) { return chain(a, expm1(a.v), exp(a.v)); }
SXC_HD D2 dsqrt(D2 a) {
  const double s = sqrt(a.v);
  return chain(a, s, 0.5 / s);
}
SXC_HD D2 datan(D2 a) { return chain(a, atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
SXC_HD D2 dpow(D2 a, double p) {
  const double f = pow(a.v, p);
  return chain(a, f, p * f / a.v);
}

// Constants of the functionals (the published VWN5, PW92 and PBE values).
// The two PBE evaluators Serenity offers differ in the exchange mu: XCFun
// (its default for LDA/PBE) takes mu = 0.066725 pi^2 / 3 from the PBE paper's
// rounded beta, LibXC takes mu = beta pi^2 / 3 with the accurate beta.  Both
// use A = 0.0310907 in PW92 and the accurate beta in the correlation.  The
// full-SCF energies of Serenity's water/STO-3G PBE tests decide these values
// (qcserenity/tests/test_tasks.py:50-76: 2.8e-8 for XCFun, 2.2e-10 for LibXC).
struct XcParams {
  double vwn_A = 0.0310907, vwn_b = 3.72744, vwn_c = 12.9352, vwn_x0 = -0.10498;
  double pw_A = 0.0310907, pw_a1 = 0.21370, pw_b1 = 7.5957, pw_b2 = 3.5876, pw_b3 = 1.6382, pw_b4 = 0.49294;
  double pbe_kappa = 0.804, pbe_mu = 0.066725 * M_PI * M_PI / 3.0, pbe_beta = 0.06672455060314922;
  double pbe_mu_libxc = 0.2195149727645171;
};

SXC_HD D2 slater_x(D2 rho) {
  // -(3/4) (3/pi)^(1/3) rho^(4/3)
  const double c = -0.75 * cbrt(3.0 / M_PI);
  return c * dpow(rho, 4.0 / 3.0);
}

SXC_HD D2 vwn5_eps(D2 rs, const XcParams& p) {
  // per-particle VWN5 paramagnetic correlation as a function of x = sqrt(r_s)
  const double b = p.vwn_b, c = p.vwn_c, x0 = p.vwn_x0;
  const double Q = sqrt(4.0 * c - b * b);
  const double X0 = x0 * x0 + b * x0 + c;
  D2 x = dsqrt(rs);
  D2 X = x * x + b * x + c;
  D2 at = datan(Q / (2.0 * x + b));
  D2 t1 = dlog(x * x / X) + (2.0 * b / Q) * at;
  D2 t2 = dlog((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / Q) * at;
  return p.vwn_A * (t1 - (b * x0 / X0) * t2);
}

SXC_HD D2 rs_of(D2 rho) { return dpow((3.0 / (4.0 * M_PI)) / rho, 1.0 / 3.0); }

SXC_HD D2 vwn5_c(D2 rho, const XcParams& p) { return rho * vwn5_eps(rs_of(rho), p); }

SXC_HD D2 pw92_eps(D2 rs, const XcParams& p) {
  // Perdew-Wang 1992 unpolarised correlation energy per particle (p = 1)
  D2 srs = dsqrt(rs);
  D2 den = 2.0 * p.pw_A * (p.pw_b1 * srs + p.pw_b2 * rs + p.pw_b3 * rs * srs + p.pw_b4 * rs * rs);
  return -2.0 * p.pw_A * (1.0 + p.pw_a1 * rs) * dlog(1.0 + 1.0 / den);
}

SXC_HD D2 pbe_x(D2 rho, D2 sigma, const XcParams& p, double mu) {
  // s^2 = sigma / (4 (3 pi^2)^(2/3) rho^(8/3))
  const double k = 4.0 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  D2 s2 = sigma / (k * dpow(rho, 8.0 / 3.0));
  D2 Fx = (1.0 + p.pbe_kappa) - p.pbe_kappa / (1.0 + (mu / p.pbe_kappa) * s2);
  return slater_x(rho) * Fx;
}

SXC_HD D2 pbe_c(D2 rho, D2 sigma, const XcParams& p) {
  const double gamma = (1.0 - log(2.0)) / (M_PI * M_PI);
  const double bg = p.pbe_beta / gamma;
  D2 eps = pw92_eps(rs_of(rho), p);
  // t^2 = sigma pi / (16 (3 pi^2)^(1/3) rho^(7/3))  (zeta = 0, phi = 1)
  const double kt = M_PI / (16.0 * cbrt(3.0 * M_PI * M_PI));
  D2 t2 = kt * sigma / dpow(rho, 7.0 / 3.0);
  D2 A = bg / dexpm1(-eps / gamma);
  D2 At2 = A * t2;
  D2 H = gamma * dlog(1.0 + bg * t2 * (1.0 + At2) / (1.0 + At2 * (1.0 + At2)));
  return rho * (eps + H);
}

SXC_HD D2 tf_k(D2 rho) {
  // C_F rho^(5/3), C_F = (3/10)(3 pi^2)^(2/3)
  const double cf = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  return cf * dpow(rho, 5.0 / 3.0);
}

// Functional ids as in serenity_xc.h.
constexpr int kNumFunctionals = 8;  // LDA, PBE, TF, SLATER, VWN5, PBEX, PBEC, PBE (LibXC constants)
SXC_HD bool functional_is_gga(int f) { return f == 1 || f == 5 || f == 6 || f == 7; }

SXC_HD D2 eval_functional(int f, D2 rho, D2 sigma, const XcParams& p) {
  switch (f) {
    case 0:
      return slater_x(rho) + vwn5_c(rho, p);
    case 1:
      return pbe_x(rho, sigma, p, p.pbe_mu) + pbe_c(rho, sigma, p);
    case 2:
      return tf_k(rho);
    case 3:
      return slater_x(rho);
    case 4:
      return vwn5_c(rho, p);
    case 5:
      return pbe_x(rho, sigma, p, p.pbe_mu);
    case 6:
      return pbe_c(rho, sigma, p);
    case 7:
      return pbe_x(rho, sigma, p, p.pbe_mu_libxc) + pbe_c(rho, sigma, p);
    default:
      return mk(0.0);
  }
}

// Densities below this value contribute nothing (energy and potential zero).
constexpr double kTinyDensity = 1e-14;

// One grid point: returns eps and writes v and (for GGA) g = dEps/dGradRho.
SXC_HD double xc_point(int f, double rho, double gx, double gy, double gz, const XcParams& p, double* v, double* g) {
  g[0] = g[1] = g[2] = 0.0;
  if (!(rho >= kTinyDensity)) {
    *v = 0.0;
    return 0.0;
  }
  const double sigma = gx * gx + gy * gy + gz * gz;
  D2 r = {rho, 1.0, 0.0};
  D2 s = {sigma, 0.0, 1.0};
  D2 e = eval_functional(f, r, s, p);
  *v = e.dr;
  if (functional_is_gga(f)) {
    g[0] = 2.0 * e.ds * gx;
    g[1] = 2.0 * e.ds * gy;
    g[2] = 2.0 * e.ds * gz;
  }
  return e.v;
}

}  // namespace sxc
