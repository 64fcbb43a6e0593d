// Contracted-shell handling: Turbomole basis parsing, libint2-style contraction
// renormalisation, Cartesian component factors and real solid harmonics.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "host.h"

namespace sxc {

namespace {

double double_factorial(int k) {  // k!! with (-1)!! = 1
  double r = 1.0;
  for (int i = k; i > 1; i -= 2) r *= i;
  return r;
}

double factorial(int k) {
  double r = 1.0;
  for (int i = 2; i <= k; ++i) r *= i;
  return r;
}

double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  return factorial(n) / (factorial(k) * factorial(n - k));
}

// Index of monomial x^a y^b z^c (a+b+c = l) in Serenity's Cartesian order
// (a = l..0, b = l-a..0).
int cart_index(int l, int a, int b) {
  int idx = 0;
  for (int aa = l; aa > a; --aa) idx += l - aa + 1;
  return idx + (l - a - b);
}

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

int am_of_char(char c) {
  static const std::string letters = "spdfghi";
  auto pos = letters.find((char)std::tolower((unsigned char)c));
  if (pos == std::string::npos) throw std::runtime_error(std::string("unknown angular momentum '") + c + "'");
  return (int)pos;
}

}  // namespace

void renormalize_contraction(int l, const std::vector<double>& alpha, std::vector<double>& coef) {
  // Primitive factor N = sqrt(2^l (2a)^(l+3/2) / (pi^(3/2) (2l-1)!!)) normalises x^l exp(-a r^2).
  const double pi32 = std::pow(M_PI, 1.5);
  const double dfl = double_factorial(2 * l - 1);
  const size_t np = alpha.size();
  for (size_t p = 0; p < np; ++p) {
    if (alpha[p] == 0.0) continue;
    const double two_a = 2.0 * alpha[p];
    coef[p] *= std::sqrt(std::pow(2.0, l) * std::pow(two_a, l + 1.5) / (pi32 * dfl));
  }
  // Self-overlap of the contraction for the x^l component:
  //   <x^l g_p | x^l g_q> = (2l-1)!! pi^(3/2) / (2^l (a_p+a_q)^(l+3/2)).
  double norm = 0.0;
  for (size_t p = 0; p < np; ++p) {
    for (size_t q = 0; q <= p; ++q) {
      const double g = alpha[p] + alpha[q];
      norm += (p == q ? 1.0 : 2.0) * dfl * pi32 * coef[p] * coef[q] / (std::pow(2.0, l) * std::pow(g, l + 1.5));
    }
  }
  const double scale = 1.0 / std::sqrt(norm);
  for (auto& c : coef) c *= scale;
}

std::vector<double> cartesian_norm_factors(int l) {
  std::vector<double> f;
  const double top = double_factorial(2 * l - 1);
  for (int a = l; a >= 0; --a)
    for (int b = l - a; b >= 0; --b) {
      const int c = l - a - b;
      f.push_back(std::sqrt(top / (double_factorial(2 * a - 1) * double_factorial(2 * b - 1) * double_factorial(2 * c - 1))));
    }
  return f;
}

std::vector<double> solid_harmonic_coefficients(int l) {
  const int ncart = (l + 1) * (l + 2) / 2;
  std::vector<double> T((size_t)(2 * l + 1) * ncart, 0.0);
  for (int m = -l; m <= l; ++m) {
    const int am = std::abs(m);
    // N_lm = 1/(2^|m| l!) sqrt(2 (l+|m|)! (l-|m|)! / 2^delta_m0)
    const double nlm = 1.0 / (std::pow(2.0, am) * factorial(l)) *
                       std::sqrt(2.0 * factorial(l + am) * factorial(l - am) / (m == 0 ? 2.0 : 1.0));
    // v runs over v_m, v_m+1, ..., with v_m = 0 (m >= 0) or 1/2 (m < 0); work with 2v.
    const int twovm = (m < 0) ? 1 : 0;
    for (int t = 0; t <= (l - am) / 2; ++t) {
      for (int u = 0; u <= t; ++u) {
        const int vmax2 = 2 * ((am - twovm) / 2) + twovm;  // 2*(floor(|m|/2 - v_m) + v_m)
        for (int twov = twovm; twov <= vmax2; twov += 2) {
          const int sign_exp = t + (twov - twovm) / 2;
          const double c = ((sign_exp % 2) ? -1.0 : 1.0) * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                           binom(t, u) * binom(am, twov);
          const int ey = 2 * u + twov;
          const int ex = 2 * t + am - ey;
          const int ez = l - 2 * t - am;
          if (ex < 0 || ey < 0 || ez < 0) continue;
          T[(size_t)(m + l) * ncart + cart_index(l, ex, ey)] += nlm * c;
        }
      }
    }
  }
  return T;
}

std::vector<Shell> parse_turbomole_basis(const std::string& path, const std::string& element, bool pure) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open basis file " + path);
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(in, line)) lines.push_back(line);
  const std::string el = lower(element);
  size_t i = 0;
  bool found = false;
  for (; i < lines.size(); ++i) {
    std::istringstream ls(lines[i]);
    std::string tok;
    if (!(ls >> tok)) continue;
    if (lower(tok) == el && tok[0] != '#' && tok[0] != '*' && tok[0] != '$') {
      std::string name;
      if (ls >> name) {
        found = true;
        break;
      }
    }
  }
  if (!found) throw std::runtime_error("basis file " + path + " has no entry for element " + element);
  ++i;  // the '*' after the header
  while (i < lines.size() && lines[i].find('*') == std::string::npos) ++i;
  ++i;
  std::string body;
  for (; i < lines.size(); ++i) {
    if (lines[i].find('*') != std::string::npos) break;
    if (!lines[i].empty() && lines[i][0] == '#') continue;
    std::string s = lines[i];
    for (size_t k = 0; k + 1 < s.size(); ++k)
      if ((s[k] == 'D' || s[k] == 'd') && (s[k + 1] == '+' || s[k + 1] == '-')) s[k] = 'E';
    body += s + "\n";
  }
  std::istringstream ws(body);
  std::vector<Shell> shells;
  int nprim;
  while (ws >> nprim) {
    std::string type;
    if (!(ws >> type) || type.size() != 1) throw std::runtime_error("malformed shell header for " + element);
    Shell sh;
    sh.l = am_of_char(type[0]);
    sh.pure = pure;
    for (int p = 0; p < nprim; ++p) {
      double a, c;
      if (!(ws >> a >> c)) throw std::runtime_error("malformed primitive list for " + element);
      sh.alpha.push_back(a);
      sh.coef.push_back(c);
    }
    renormalize_contraction(sh.l, sh.alpha, sh.coef);
    shells.push_back(std::move(sh));
  }
  if (shells.empty()) throw std::runtime_error("no shells parsed for " + element);
  return shells;
}

}  // namespace sxc
