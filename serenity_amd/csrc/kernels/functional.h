// Exchange-correlation functionals of the XC Fock build, FP64, evaluated per
// grid point with forward-mode derivatives so the potential terms come from
// the same code as the energy.
//
// The formulas restate the basic functionals Serenity takes from XCFun (or
// LibXC) for the LDA, PBE and BP86 composites (functional_definitions.dat:44,47,48;
// basic_functionals.dat:35,78,99,113,168,226,230): Slater exchange, VWN5 correlation
// (Vosko-Wilk-Nusair 1980 with their spin interpolation), PBE exchange and
// correlation (Perdew-Burke-Ernzerhof 1996 with the Perdew-Wang 1992 LDA
// correlation), Becke 1988 exchange, Perdew 1986 correlation on Perdew-Zunger
// 1981, and Thomas-Fermi kinetic energy.
//
// Restricted inputs: total density rho and sigma = |grad rho|^2 (partials in
// that order); XCFun's XC_N_NX_NY_NZ outputs are then
//   eps (energy per volume), v = d eps/d rho, g = 2 d eps/d sigma * grad rho.
// Unrestricted inputs (XC_A_B_AX_AY_AZ_BX_BY_BZ): rho_a, rho_b and
// sigma_aa, sigma_ab, sigma_bb (partials in that order); outputs
//   v_a = d eps/d rho_a, g_a = 2 d eps/d sigma_aa grad rho_a + d eps/d sigma_ab grad rho_b
// and likewise for b.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define SXC_HD __host__ __device__ inline
#else
#define SXC_HD inline
#endif

namespace sxc {

// Value and N first partials.
template <int N>
struct Dual {
  double v;
  double d[N];
};
typedef Dual<2> D2;  // (rho, sigma)
typedef Dual<5> D5;  // (rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb)

template <int N>
SXC_HD Dual<N> mkc(double v) {
  Dual<N> r;
  r.v = v;
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N>
SXC_HD Dual<N> var(double v, int k) {
  Dual<N> r = mkc<N>(v);
  r.d[k] = 1.0;
  return r;
}
SXC_HD D2 mk(double v) { return mkc<2>(v); }

template <int N>
SXC_HD Dual<N> operator+(Dual<N> a, Dual<N> b) {
  a.v += b.v;
  for (int i = 0; i < N; ++i) a.d[i] += b.d[i];
  return a;
}
template <int N>
SXC_HD Dual<N> operator-(Dual<N> a, Dual<N> b) {
  a.v -= b.v;
  for (int i = 0; i < N; ++i) a.d[i] -= b.d[i];
  return a;
}
template <int N>
SXC_HD Dual<N> operator-(Dual<N> a) {
  a.v = -a.v;
  for (int i = 0; i < N; ++i) a.d[i] = -a.d[i];
  return a;
}
template <int N>
SXC_HD Dual<N> operator*(Dual<N> a, Dual<N> b) {
  Dual<N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
SXC_HD Dual<N> operator/(Dual<N> a, Dual<N> b) {
  const double ib = 1.0 / b.v, q = a.v * ib;
  Dual<N> r;
  r.v = q;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * ib;
  return r;
}
template <int N>
SXC_HD Dual<N> operator+(Dual<N> a, double b) {
  a.v += b;
  return a;
}
template <int N>
SXC_HD Dual<N> operator+(double b, Dual<N> a) {
  a.v += b;
  return a;
}
template <int N>
SXC_HD Dual<N> operator-(Dual<N> a, double b) {
  a.v -= b;
  return a;
}
template <int N>
SXC_HD Dual<N> operator-(double b, Dual<N> a) {
  a.v = b - a.v;
  for (int i = 0; i < N; ++i) a.d[i] = -a.d[i];
  return a;
}
template <int N>
SXC_HD Dual<N> operator*(Dual<N> a, double b) {
  a.v *= b;
  for (int i = 0; i < N; ++i) a.d[i] *= b;
  return a;
}
template <int N>
SXC_HD Dual<N> operator*(double b, Dual<N> a) {
  return a * b;
}
template <int N>
SXC_HD Dual<N> operator/(Dual<N> a, double b) {
  a.v /= b;
  for (int i = 0; i < N; ++i) a.d[i] /= b;
  return a;
}
template <int N>
SXC_HD Dual<N> operator/(double b, Dual<N> a) {
  const double q = b / a.v, dq = -q / a.v;
  Dual<N> r;
  r.v = q;
  for (int i = 0; i < N; ++i) r.d[i] = dq * a.d[i];
  return r;
}
template <int N>
SXC_HD Dual<N> chain(Dual<N> a, double f, double df) {
  Dual<N> r;
  r.v = f;
  for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i];
  return r;
}
template <int N>
SXC_HD Dual<N> dlog(Dual<N> a) {
  return chain(a, log(a.v), 1.0 / a.v);
}
template <int N>
SXC_HD Dual<N> dexpm1(Dual<N> a) {
  return chain(a, expm1(a.v), exp(a.v));
}
template <int N>
SXC_HD Dual<N> dsqrt(Dual<N> a) {
  const double s = sqrt(a.v);
  return chain(a, s, 0.5 / s);
}
template <int N>
SXC_HD Dual<N> datan(Dual<N> a) {
  return chain(a, atan(a.v), 1.0 / (1.0 + a.v * a.v));
}
template <int N>
SXC_HD Dual<N> dpow(Dual<N> a, double p) {
  const double f = pow(a.v, p);
  return chain(a, f, p * f / a.v);
}
// a^(4/3) with the derivative 4/3 a^(1/3), finite at a = 0 (spin-interpolation of a full polarisation)
template <int N>
SXC_HD Dual<N> dpow43(Dual<N> a) {
  const double c = cbrt(a.v);
  return chain(a, a.v * c, (4.0 / 3.0) * c);
}
// a^(5/3) with the derivative 5/3 a^(2/3), finite at a = 0
template <int N>
SXC_HD Dual<N> dpow53(Dual<N> a) {
  const double c = cbrt(a.v);
  return chain(a, a.v * c * c, (5.0 / 3.0) * c * c);
}
template <int N>
SXC_HD Dual<N> dexp(Dual<N> a) {
  const double e = exp(a.v);
  return chain(a, e, e);
}
template <int N>
SXC_HD Dual<N> dasinh(Dual<N> a) {
  return chain(a, asinh(a.v), 1.0 / sqrt(1.0 + a.v * a.v));
}

// Constants of the functionals (the published VWN5, PW92 and PBE values).
// The two PBE evaluators Serenity offers differ in the exchange mu: XCFun
// (its default for LDA/PBE) takes mu = 0.066725 pi^2 / 3 from the PBE paper's
// rounded beta, LibXC takes mu = beta pi^2 / 3 with the accurate beta.  Both
// use A = 0.0310907 in PW92 and the accurate beta in the correlation.  The
// full-SCF energies of Serenity's water/STO-3G PBE tests decide these values
// (qcserenity/tests/test_tasks.py:50-76: 2.8e-8 for XCFun, 2.2e-10 for LibXC).
// Spin-polarised parts: VWN5 ferromagnetic and spin-stiffness fits, PW92
// ferromagnetic and spin-stiffness fits (the A of the "modified" PW92 that
// goes with 0.0310907), f''(0) = 4 / (9 (2^(1/3) - 1)).
struct XcParams {
  double vwn_A = 0.0310907, vwn_b = 3.72744, vwn_c = 12.9352, vwn_x0 = -0.10498;
  double pw_A = 0.0310907, pw_a1 = 0.21370, pw_b1 = 7.5957, pw_b2 = 3.5876, pw_b3 = 1.6382, pw_b4 = 0.49294;
  double pbe_kappa = 0.804, pbe_mu = 0.066725 * M_PI * M_PI / 3.0, pbe_beta = 0.06672455060314922;
  double pbe_mu_libxc = 0.2195149727645171;
  // VWN5: ferromagnetic (A, b, c, x0) and spin stiffness
  double vwnF_A = 0.01554535, vwnF_b = 7.06042, vwnF_c = 18.0578, vwnF_x0 = -0.32500;
  double vwnS_A = -1.0 / (6.0 * M_PI * M_PI), vwnS_b = 1.13107, vwnS_c = 13.0045, vwnS_x0 = -0.0047584;
  // PW92: ferromagnetic and (minus) spin stiffness
  double pwF_A = 0.01554535, pwF_a1 = 0.20548, pwF_b1 = 14.1189, pwF_b2 = 6.1977, pwF_b3 = 3.3662,
         pwF_b4 = 0.62517;
  double pwS_A = 0.0168869, pwS_a1 = 0.11125, pwS_b1 = 10.357, pwS_b2 = 3.6231, pwS_b3 = 0.88026,
         pwS_b4 = 0.49671;
  // BP86: Becke 1988 exchange (beta), Perdew-Zunger 1981 correlation (paramagnetic U and
  // ferromagnetic P fits: gamma, beta1, beta2 for r_s >= 1; A, B, C, D for r_s < 1, the published
  // values) and the Perdew 1986 gradient correction (C(n) = c0 + (c1 + alpha r_s + beta r_s^2) /
  // (1 + gamma r_s + delta r_s^2 + 1e4 beta r_s^3), f~ = 0.11).  The reference's BP86 Fock
  // matrix (FuncPotential_test.cpp:148-232) decides the one choice the paper leaves open: the
  // prefactor of Phi is f~ (9 pi)^(1/6) (1.745 in the paper is its rounding), with the original
  // PZ81 constants; 1.745 f~ misses by 8e-6, the continuity-corrected PZ81 C, D by 2.7e-6.
  double b88_beta = 0.0042;
  double pzU_g = -0.1423, pzU_b1 = 1.0529, pzU_b2 = 0.3334, pzU_A = 0.0311, pzU_B = -0.048, pzU_C = 0.0020,
         pzU_D = -0.0116;
  double pzP_g = -0.0843, pzP_b1 = 1.3981, pzP_b2 = 0.2611, pzP_A = 0.01555, pzP_B = -0.0269, pzP_C = 0.0007,
         pzP_D = -0.0048;
  double p86_alpha = 0.023266, p86_beta = 7.389e-6, p86_gamma = 8.723, p86_delta = 0.472, p86_c0 = 0.001667,
         p86_c1 = 0.002568, p86_ft = 0.11;
};

// f''(0) of the spin interpolation f(zeta) = ((1+zeta)^(4/3) + (1-zeta)^(4/3) - 2) / (2^(4/3) - 2)
constexpr double kFpp0 = 1.709920934161365617563962776245;

template <class D>
SXC_HD D slater_x(D rho) {
  // -(3/4) (3/pi)^(1/3) rho^(4/3)
  const double c = -0.75 * cbrt(3.0 / M_PI);
  return c * dpow(rho, 4.0 / 3.0);
}

// VWN per-particle correlation fit as a function of x = sqrt(r_s)
template <class D>
SXC_HD D vwn_fit(D rs, double A, double b, double c, double x0) {
  const double Q = sqrt(4.0 * c - b * b);
  const double X0 = x0 * x0 + b * x0 + c;
  D x = dsqrt(rs);
  D X = x * x + b * x + c;
  D at = datan(Q / (2.0 * x + b));
  D t1 = dlog(x * x / X) + (2.0 * b / Q) * at;
  D t2 = dlog((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / Q) * at;
  return A * (t1 - (b * x0 / X0) * t2);
}

template <class D>
SXC_HD D vwn5_eps(D rs, const XcParams& p) {
  return vwn_fit(rs, p.vwn_A, p.vwn_b, p.vwn_c, p.vwn_x0);
}

template <class D>
SXC_HD D rs_of(D rho) {
  return dpow((3.0 / (4.0 * M_PI)) / rho, 1.0 / 3.0);
}

template <class D>
SXC_HD D vwn5_c(D rho, const XcParams& p) {
  return rho * vwn5_eps(rs_of(rho), p);
}

// Perdew-Wang 1992 fit G(r_s) (p = 1)
template <class D>
SXC_HD D pw_fit(D rs, double A, double a1, double b1, double b2, double b3, double b4) {
  D srs = dsqrt(rs);
  D den = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  return -2.0 * A * (1.0 + a1 * rs) * dlog(1.0 + 1.0 / den);
}

template <class D>
SXC_HD D pw92_eps(D rs, const XcParams& p) {
  // Perdew-Wang 1992 unpolarised correlation energy per particle
  return pw_fit(rs, p.pw_A, p.pw_a1, p.pw_b1, p.pw_b2, p.pw_b3, p.pw_b4);
}

template <class D>
SXC_HD D pbe_x(D rho, D sigma, const XcParams& p, double mu) {
  // s^2 = sigma / (4 (3 pi^2)^(2/3) rho^(8/3))
  const double k = 4.0 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  D s2 = sigma / (k * dpow(rho, 8.0 / 3.0));
  D Fx = (1.0 + p.pbe_kappa) - p.pbe_kappa / (1.0 + (mu / p.pbe_kappa) * s2);
  return slater_x(rho) * Fx;
}

// PBE gradient correction H(eps, t^2, phi) per particle (phi = 1 unpolarised)
template <class D>
SXC_HD D pbe_H(D eps, D t2, D phi3, const XcParams& p) {
  const double gamma = (1.0 - log(2.0)) / (M_PI * M_PI);
  const double bg = p.pbe_beta / gamma;
  D A = bg / dexpm1(-eps / (gamma * phi3));
  D At2 = A * t2;
  return gamma * phi3 * dlog(1.0 + bg * t2 * (1.0 + At2) / (1.0 + At2 * (1.0 + At2)));
}

template <class D>
SXC_HD D pbe_c(D rho, D sigma, const XcParams& p) {
  const double gamma = (1.0 - log(2.0)) / (M_PI * M_PI);
  const double bg = p.pbe_beta / gamma;
  D eps = pw92_eps(rs_of(rho), p);
  // t^2 = sigma pi / (16 (3 pi^2)^(1/3) rho^(7/3))  (zeta = 0, phi = 1)
  const double kt = M_PI / (16.0 * cbrt(3.0 * M_PI * M_PI));
  D t2 = kt * sigma / dpow(rho, 7.0 / 3.0);
  D A = bg / dexpm1(-eps / gamma);
  D At2 = A * t2;
  D H = gamma * dlog(1.0 + bg * t2 * (1.0 + At2) / (1.0 + At2 * (1.0 + At2)));
  return rho * (eps + H);
}

template <class D>
SXC_HD D tf_k(D rho) {
  // C_F rho^(5/3), C_F = (3/10)(3 pi^2)^(2/3)
  const double cf = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  return cf * dpow(rho, 5.0 / 3.0);
}

// x asinh(x) as a function of u = x^2: smooth at u = 0, where sqrt is not (series below 1e-12)
template <class D>
SXC_HD D xasinhx(D u) {
  if (u.v < 1e-12) return u - u * u / 6.0;
  D x = dsqrt(u);
  return x * dasinh(x);
}

// Becke 1988 exchange, restricted: two spins with rho_s = rho / 2, sigma_ss = sigma / 4, each
// -C rho_s^(4/3) - beta rho_s^(4/3) x^2 / (1 + 6 beta x asinh x), x^2 = sigma_ss / rho_s^(8/3)
template <class D>
SXC_HD D b88_x(D rho, D sigma, const XcParams& p) {
  D r43 = dpow(0.5 * rho, 4.0 / 3.0);
  D ss = 0.25 * sigma;
  D u = ss / (r43 * r43);
  return slater_x(rho) - 2.0 * p.b88_beta * (ss / r43) / (1.0 + 6.0 * p.b88_beta * xasinhx(u));
}

// Perdew-Zunger 1981 correlation energy per particle of one spin state
template <class D>
SXC_HD D pz81_fit(D rs, double g, double b1, double b2, double A, double B, double C, double Dd) {
  if (rs.v >= 1.0) return g / (1.0 + b1 * dsqrt(rs) + b2 * rs);
  D l = dlog(rs);
  return A * l + B + C * rs * l + Dd * rs;
}

// Perdew 1986 gradient correction exp(-Phi) C(n) sigma / (d n^(4/3)), Phi = f~ (9 pi)^(1/6)
// C(inf) / C(n) |grad n| / n^(7/6); dz = d(zeta) (1 for a closed shell)
template <class D>
SXC_HD D p86_corr(D rho, D sigma, D dz, const XcParams& p) {
  D rs = rs_of(rho);
  D Cn = p.p86_c0 + (p.p86_c1 + p.p86_alpha * rs + p.p86_beta * rs * rs) /
                        (1.0 + p.p86_gamma * rs + p.p86_delta * rs * rs + 1.0e4 * p.p86_beta * rs * rs * rs);
  const double k = p.p86_ft * pow(9.0 * M_PI, 1.0 / 6.0) * (p.p86_c0 + p.p86_c1);
  // |grad n|: at sigma = 0 the correction is C(n) sigma / n^(4/3) to first order
  D g = sigma.v > 0.0 ? dsqrt(sigma) : 0.0 * sigma;
  D Phi = k * g / (Cn * dpow(rho, 7.0 / 6.0));
  return dexp(-Phi) * Cn * sigma / (dz * dpow(rho, 4.0 / 3.0));
}

template <class D>
SXC_HD D bp86(D rho, D sigma, const XcParams& p) {
  D eU = pz81_fit(rs_of(rho), p.pzU_g, p.pzU_b1, p.pzU_b2, p.pzU_A, p.pzU_B, p.pzU_C, p.pzU_D);
  return b88_x(rho, sigma, p) + rho * eU + p86_corr(rho, sigma, 1.0 + 0.0 * rho, p);
}

// ---- spin-polarised forms ---------------------------------------------------
// Spin densities below kTinyDensity count as zero: their terms are dropped.

// spin interpolation f(zeta) from 1 + zeta = 2 rho_a / rho and 1 - zeta = 2 rho_b / rho
template <class D>
SXC_HD D spin_f(D opz, D omz) {
  return (dpow43(opz) + dpow43(omz) - 2.0) / (2.0 * cbrt(2.0) - 2.0);
}

template <class D>
SXC_HD D vwn5_c_pol(D ra, D rb, const XcParams& p) {
  D rho = ra + rb;
  D rs = rs_of(rho);
  D zeta = (ra - rb) / rho;
  D f = spin_f(1.0 + zeta, 1.0 - zeta);
  D z4 = zeta * zeta * zeta * zeta;
  D eP = vwn_fit(rs, p.vwn_A, p.vwn_b, p.vwn_c, p.vwn_x0);
  D eF = vwn_fit(rs, p.vwnF_A, p.vwnF_b, p.vwnF_c, p.vwnF_x0);
  D aS = vwn_fit(rs, p.vwnS_A, p.vwnS_b, p.vwnS_c, p.vwnS_x0);
  return rho * (eP + aS * f * (1.0 - z4) / kFpp0 + (eF - eP) * f * z4);
}

template <class D>
SXC_HD D pw92_eps_pol(D rs, D zeta, D f, const XcParams& p) {
  D z4 = zeta * zeta * zeta * zeta;
  D e0 = pw_fit(rs, p.pw_A, p.pw_a1, p.pw_b1, p.pw_b2, p.pw_b3, p.pw_b4);
  D e1 = pw_fit(rs, p.pwF_A, p.pwF_a1, p.pwF_b1, p.pwF_b2, p.pwF_b3, p.pwF_b4);
  D mac = pw_fit(rs, p.pwS_A, p.pwS_a1, p.pwS_b1, p.pwS_b2, p.pwS_b3, p.pwS_b4);  // -alpha_c
  return e0 - mac * f * (1.0 - z4) / kFpp0 + (e1 - e0) * f * z4;
}

template <class D>
SXC_HD D pbe_c_pol(D ra, D rb, D sigma, const XcParams& p) {
  D rho = ra + rb;
  D rs = rs_of(rho);
  D zeta = (ra - rb) / rho;
  D opz = 1.0 + zeta, omz = 1.0 - zeta;
  D eps = pw92_eps_pol(rs, zeta, spin_f(opz, omz), p);
  // phi = ((1+zeta)^(2/3) + (1-zeta)^(2/3)) / 2, arguments kept away from 0 (its slope diverges there)
  if (opz.v < 1e-12) opz.v = 1e-12;
  if (omz.v < 1e-12) omz.v = 1e-12;
  D phi = 0.5 * (dpow(opz, 2.0 / 3.0) + dpow(omz, 2.0 / 3.0));
  D phi3 = phi * phi * phi;
  const double kt = M_PI / (16.0 * cbrt(3.0 * M_PI * M_PI));
  D t2 = kt * sigma / (phi * phi * dpow(rho, 7.0 / 3.0));
  return rho * (eps + pbe_H(eps, t2, phi3, p));
}

// BP86 correlation, spin-polarised: PZ81 with the spin interpolation f(zeta) between its
// paramagnetic and ferromagnetic fits, P86 correction of the total density with
// d(zeta) = 2^(1/3) sqrt(((1+zeta)/2)^(5/3) + ((1-zeta)/2)^(5/3))
template <class D>
SXC_HD D bp86_c_pol(D ra, D rb, D sigma, const XcParams& p) {
  D rho = ra + rb;
  D rs = rs_of(rho);
  D zeta = (ra - rb) / rho;
  D opz = 1.0 + zeta, omz = 1.0 - zeta;
  D f = spin_f(opz, omz);
  D eU = pz81_fit(rs, p.pzU_g, p.pzU_b1, p.pzU_b2, p.pzU_A, p.pzU_B, p.pzU_C, p.pzU_D);
  D eP = pz81_fit(rs, p.pzP_g, p.pzP_b1, p.pzP_b2, p.pzP_A, p.pzP_B, p.pzP_C, p.pzP_D);
  D dz = cbrt(2.0) * dsqrt(dpow53(0.5 * opz) + dpow53(0.5 * omz));
  return rho * (eU + f * (eP - eU)) + p86_corr(rho, sigma, dz, p);
}

// Functional ids as in serenity_xc.h.
constexpr int kNumFunctionals = 9;  // LDA, PBE, TF, SLATER, VWN5, PBEX, PBEC, PBE (LibXC constants), BP86
SXC_HD bool functional_is_gga(int f) { return f == 1 || f == 5 || f == 6 || f == 7 || f == 8; }

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
    case 8:
      return bp86(rho, sigma, p);
    default:
      return mk(0.0);
  }
}

// Spin-polarised evaluation: v = (rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb) as D5
// variables; a_on / b_on say whether each spin density is above kTinyDensity.
SXC_HD D5 eval_functional_pol(int f, D5 ra, D5 rb, D5 saa, D5 sab, D5 sbb, bool a_on, bool b_on,
                              const XcParams& p) {
  D5 sigma = saa + 2.0 * sab + sbb;
  D5 e = mkc<5>(0.0);
  // exchange-type parts by spin scaling, E[rho_a, rho_b] = (E[2 rho_a] + E[2 rho_b]) / 2
  // with 4 sigma_ss as the gradient invariant
  // kind: 0 Slater, 1 Thomas-Fermi, 2 PBE exchange (XCFun mu), 3 PBE exchange (LibXC mu), 4 Becke 88
  auto one = [&](int kind, D5 r2, D5 s4) {
    switch (kind) {
      case 0:
        return slater_x(r2);
      case 1:
        return tf_k(r2);
      case 2:
        return pbe_x(r2, s4, p, p.pbe_mu);
      case 3:
        return pbe_x(r2, s4, p, p.pbe_mu_libxc);
      default:
        return b88_x(r2, s4, p);
    }
  };
  auto xpart = [&](int kind) {
    D5 out = mkc<5>(0.0);
    if (a_on) {
      D5 r2 = 2.0 * ra, s4 = 4.0 * saa;
      out = out + 0.5 * one(kind, r2, s4);
    }
    if (b_on) {
      D5 r2 = 2.0 * rb, s4 = 4.0 * sbb;
      out = out + 0.5 * one(kind, r2, s4);
    }
    return out;
  };
  // a fully polarised point (one spin density zero) keeps the exact zeta = +-1 limits
  D5 ra_c = a_on ? ra : mkc<5>(0.0), rb_c = b_on ? rb : mkc<5>(0.0);
  switch (f) {
    case 0:
      e = xpart(0) + vwn5_c_pol(ra_c, rb_c, p);
      break;
    case 1:
      e = xpart(2) + pbe_c_pol(ra_c, rb_c, sigma, p);
      break;
    case 2:
      e = xpart(1);
      break;
    case 3:
      e = xpart(0);
      break;
    case 4:
      e = vwn5_c_pol(ra_c, rb_c, p);
      break;
    case 5:
      e = xpart(2);
      break;
    case 6:
      e = pbe_c_pol(ra_c, rb_c, sigma, p);
      break;
    case 7:
      e = xpart(3) + pbe_c_pol(ra_c, rb_c, sigma, p);
      break;
    case 8:
      e = xpart(4) + bp86_c_pol(ra_c, rb_c, sigma, p);
      break;
    default:
      break;
  }
  return e;
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
  D2 r = var<2>(rho, 0);
  D2 s = var<2>(sigma, 1);
  D2 e = eval_functional(f, r, s, p);
  *v = e.d[0];
  if (functional_is_gga(f)) {
    g[0] = 2.0 * e.d[1] * gx;
    g[1] = 2.0 * e.d[1] * gy;
    g[2] = 2.0 * e.d[1] * gz;
  }
  return e.v;
}

// One grid point, spin-polarised: ga / gb are the spin density gradients (3 each);
// writes v[0..1] = d eps/d rho_a, d eps/d rho_b and (GGA) g[0..2] = d eps/d grad rho_a,
// g[3..5] = d eps/d grad rho_b.  Zero when the total density is below kTinyDensity.
SXC_HD double xc_point_pol(int f, double ra, double rb, const double* ga, const double* gb, const XcParams& p,
                           double* v, double* g) {
  for (int i = 0; i < 6; ++i) g[i] = 0.0;
  v[0] = v[1] = 0.0;
  const bool a_on = ra >= kTinyDensity, b_on = rb >= kTinyDensity;
  if (!(ra + rb >= kTinyDensity) || !(a_on || b_on)) return 0.0;
  const double saa = ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2];
  const double sbb = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2];
  const double sab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
  D5 e = eval_functional_pol(f, var<5>(a_on ? ra : 0.0, 0), var<5>(b_on ? rb : 0.0, 1), var<5>(saa, 2),
                             var<5>(sab, 3), var<5>(sbb, 4), a_on, b_on, p);
  v[0] = e.d[0];
  v[1] = e.d[1];
  if (functional_is_gga(f)) {
    for (int c = 0; c < 3; ++c) {
      g[c] = 2.0 * e.d[2] * ga[c] + e.d[3] * gb[c];
      g[3 + c] = 2.0 * e.d[4] * gb[c] + e.d[3] * ga[c];
    }
  }
  return e.v;
}

}  // namespace sxc
