// Density-fitted Coulomb matrix on the GPU ("J engine"), the Coulomb part of
// the SCF around the XC build (Serenity's densFitJ = RI: RI_J_IntegralController,
// src/integrals/RI_J_IntegralController.cpp:41-121, with libint2's 3-centre
// integrals).  No 3-centre integral is stored.  Every significant primitive
// pair q of the orbital basis is a sum of Hermite Gaussians,
//   phi_a phi_b = K_q sum_tuv E^ab_tuv Lambda_tuv(p_q, P_q),
// and meets an auxiliary primitive (exponent a, centre C) through
//   (Lambda_tuv | Lambda_tau,nu,phi) = 2 pi^(5/2) / (p a sqrt(p + a)) (-1)^(tau+nu+phi)
//                                      R_(t+tau, u+nu, v+phi)(p a / (p + a), P - C)
// (McMurchie-Davidson Hermite Coulomb integrals, Boys function).  Per SCF
// iteration:
//   1. D_q = K_q sum_ij Pc_ij E^ij         (Hermite density of each primitive pair)
//   2. b_k = sum_q (D_q | aux k)           (fitting right-hand side, Cartesian aux)
//   3. (host) M c = b                      (metric solve)
//   4. V_q = sum_k c_k (Lambda_q | aux k)  (Hermite potential of each pair)
//   5. J_ab = K_q sum_tuv V_q,tuv E^ab_tuv (then the shells' angular transforms).
#include <hip/hip_runtime.h>

#include <cmath>

#include "xc_internal.h"

namespace sxc {

namespace {

// packed index of (t, u, v), s = t + u + v: all s' < s first, then t descending, u descending
__device__ inline int hidx(int t, int u, int v) {
  const int s = t + u + v;
  const int q = s - t;
  return s * (s + 1) * (s + 2) / 6 + q * (q + 1) / 2 + (q - u);
}
__device__ inline int hcount(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Boys function F_0..F_m (m <= kJMaxL): series with downward recursion below T = 35,
// asymptotic form with upward recursion above.
__device__ void boys(int m, double T, double* F) {
  const double e = exp(-T);
  if (T < 35.0) {
    double term = 1.0 / (2 * m + 1), sum = term;
    for (int k = 1; k < 200; ++k) {
      term *= 2.0 * T / (2 * m + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[m] = e * sum;
    for (int j = m; j > 0; --j) F[j - 1] = (2.0 * T * F[j] + e) / (2 * j - 1);
  } else {
    F[0] = 0.5 * sqrt(M_PI / T);
    for (int j = 0; j < m; ++j) F[j + 1] = ((2 * j + 1) * F[j] - e) / (2.0 * T);
  }
}

// R^(0)_tuv(a, X, Y, Z) for t + u + v <= L, packed (hidx); buf: 2 hcount(kJMaxL) doubles.
__device__ void hermite_R(int L, double a, double X, double Y, double Z, double* R, double* buf) {
  double F[kJMaxL + 1];
  boys(L, a * (X * X + Y * Y + Z * Z), F);
  double* cur = buf;
  double* nxt = buf + hcount(kJMaxL);
  double pw = 1.0;
  for (int n = 0; n < L; ++n) pw *= -2.0 * a;
  // nxt holds level n + 1, cur is built for level n (entries with s <= L - n)
  for (int n = L; n >= 0; --n) {
    cur[0] = pw * F[n];
    const int top = L - n;
    for (int s = 1; s <= top; ++s)
      for (int t = s; t >= 0; --t)
        for (int u = s - t; u >= 0; --u) {
          const int v = s - t - u;
          double val;
          if (t > 0) {
            val = X * nxt[hidx(t - 1, u, v)];
            if (t > 1) val += (t - 1) * nxt[hidx(t - 2, u, v)];
          } else if (u > 0) {
            val = Y * nxt[hidx(t, u - 1, v)];
            if (u > 1) val += (u - 1) * nxt[hidx(t, u - 2, v)];
          } else {
            val = Z * nxt[hidx(t, u, v - 1)];
            if (v > 1) val += (v - 1) * nxt[hidx(t, u, v - 2)];
          }
          cur[hidx(t, u, v)] = val;
        }
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
    if (n > 0) pw /= -2.0 * a;
  }
  const int cnt = hcount(L);
  for (int i = 0; i < cnt; ++i) R[i] = nxt[i];
}

// Hermite expansion E[i][j][t] (i <= la, j <= lb, t <= i + j) of one direction,
// stored at E[(i * (lb + 1) + j) * (la + lb + 1) + t].
__device__ void hermite_E(int la, int lb, double p, double XPA, double XPB, double* E) {
  const int nt = la + lb + 1;
  const double o2p = 0.5 / p;
  for (int k = 0; k < (la + 1) * (lb + 1) * nt; ++k) E[k] = 0.0;
  auto at = [&](int i, int j, int t) -> double& { return E[(i * (lb + 1) + j) * nt + t]; };
  at(0, 0, 0) = 1.0;
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
          v = XPA * at(i - 1, j, t);
          if (t > 0) v += o2p * at(i - 1, j, t - 1);
          if (t + 1 <= i - 1 + j) v += (t + 1) * at(i - 1, j, t + 1);
        } else {
          v = XPB * at(i, j - 1, t);
          if (t > 0) v += o2p * at(i, j - 1, t - 1);
          if (t + 1 <= i + j - 1) v += (t + 1) * at(i, j - 1, t + 1);
        }
        at(i, j, t) = v;
      }
    }
}

// Cartesian exponents of component c of a shell of angular momentum l (a = l..0, b = l-a..0)
__device__ inline void cart_of(int l, int c, int& ax, int& ay, int& az) {
  int k = 0;
  for (int a = l; a >= 0; --a)
    for (int b = l - a; b >= 0; --b, ++k)
      if (k == c) {
        ax = a;
        ay = b;
        az = l - a - b;
        return;
      }
  ax = ay = az = 0;
}

// Hermite expansion of x^n exp(-a x^2) about its own centre, n <= l: Ea[n * (l + 1) + tau].
__device__ void aux_E(int l, double a, double* Ea) {
  const int w = l + 1;
  for (int k = 0; k < w * w; ++k) Ea[k] = 0.0;
  Ea[0] = 1.0;
  const double o2a = 0.5 / a;
  for (int n = 1; n <= l; ++n)
    for (int t = 0; t <= n; ++t) {
      double v = 0.0;
      if (t > 0) v += o2a * Ea[(n - 1) * w + t - 1];
      if (t + 1 <= n - 1) v += (t + 1) * Ea[(n - 1) * w + t + 1];
      Ea[n * w + t] = v;
    }
}

// 1. Hermite density of each primitive pair
__global__ __launch_bounds__(128) void jdens_kernel(const JPair* __restrict__ pairs, int npair,
                                                    const DevShell* __restrict__ shells,
                                                    const double* __restrict__ ang, const double* __restrict__ P,
                                                    int n, double* __restrict__ D) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npair) return;
  const JPair pr = pairs[q];
  const DevShell A = shells[pr.A], B = shells[pr.B];
  const int la = A.l, lb = B.l, L = la + lb;
  double Pc[kJMaxNcart * kJMaxNcart];
  for (int i = 0; i < A.ncart; ++i)
    for (int j = 0; j < B.ncart; ++j) {
      double s = 0.0;
      for (int f = 0; f < A.nfunc; ++f) {
        const double ma = ang[A.ang_off + f * A.ncart + i];
        if (ma == 0.0) continue;
        for (int h = 0; h < B.nfunc; ++h)
          s += ma * P[(A.ao_off + f) + (int64_t)n * (B.ao_off + h)] * ang[B.ang_off + h * B.ncart + j];
      }
      Pc[i * B.ncart + j] = pr.sym * s;
    }
  double Ex[kJMaxE], Ey[kJMaxE], Ez[kJMaxE];
  hermite_E(la, lb, pr.p, pr.PA[0], pr.PB[0], Ex);
  hermite_E(la, lb, pr.p, pr.PA[1], pr.PB[1], Ey);
  hermite_E(la, lb, pr.p, pr.PA[2], pr.PB[2], Ez);
  const int nt = L + 1;
  double* out = D + pr.herm_off;
  for (int k = 0; k < hcount(L); ++k) out[k] = 0.0;
  for (int i = 0; i < A.ncart; ++i) {
    int ax, ay, az;
    cart_of(la, i, ax, ay, az);
    for (int j = 0; j < B.ncart; ++j) {
      int bx, by, bz;
      cart_of(lb, j, bx, by, bz);
      const double pc = pr.K * Pc[i * B.ncart + j];
      if (pc == 0.0) continue;
      const double* ex = Ex + (ax * (lb + 1) + bx) * nt;
      const double* ey = Ey + (ay * (lb + 1) + by) * nt;
      const double* ez = Ez + (az * (lb + 1) + bz) * nt;
      for (int t = 0; t <= ax + bx; ++t)
        for (int u = 0; u <= ay + by; ++u)
          for (int v = 0; v <= az + bz; ++v) out[hidx(t, u, v)] += pc * ex[t] * ey[u] * ez[v];
    }
  }
}

// 2. fitting right-hand side: one workgroup per (auxiliary shell, chunk of pairs)
__global__ __launch_bounds__(256) void jrhs_kernel(const JPair* __restrict__ pairs, int npair, int chunk,
                                                   const double* __restrict__ D,
                                                   const DevShell* __restrict__ aux,
                                                   const double* __restrict__ aux_alpha,
                                                   const double* __restrict__ aux_coef,
                                                   const int* __restrict__ aux_cart_off, double* __restrict__ bc) {
  __shared__ double red[kJMaxNcartAux][256];
  const int s = blockIdx.x;
  const DevShell S = aux[s];
  const int ls = S.l, ncs = S.ncart;
  double acc[kJMaxNcartAux];
  for (int k = 0; k < ncs; ++k) acc[k] = 0.0;
  double R[kJMaxH], buf[2 * kJMaxH], W[kJMaxHs], Ea[(kJMaxLAux + 1) * (kJMaxLAux + 1)];
  const int q0 = blockIdx.y * chunk, q1 = min(npair, q0 + chunk);
  for (int q = q0 + threadIdx.x; q < q1; q += 256) {
    const JPair pr = pairs[q];
    const int Lab = pr.L, L = Lab + ls;
    const double* Dq = D + pr.herm_off;
    for (int k = 0; k < S.nprim; ++k) {
      const double a = aux_alpha[S.prim_off + k];
      const double alpha = pr.p * a / (pr.p + a);
      const double pref = 2.0 * pow(M_PI, 2.5) / (pr.p * a * sqrt(pr.p + a)) * aux_coef[S.prim_off + k];
      hermite_R(L, alpha, pr.P[0] - S.cx, pr.P[1] - S.cy, pr.P[2] - S.cz, R, buf);
      // W[tau nu phi] = (-1)^(tau+nu+phi) sum_tuv D_tuv R_(t+tau, u+nu, v+phi)
      for (int sh = 0; sh <= ls; ++sh)
        for (int tt = sh; tt >= 0; --tt)
          for (int uu = sh - tt; uu >= 0; --uu) {
            const int vv = sh - tt - uu;
            double w = 0.0;
            for (int sab = 0; sab <= Lab; ++sab)
              for (int t = sab; t >= 0; --t)
                for (int u = sab - t; u >= 0; --u) {
                  const int v = sab - t - u;
                  w += Dq[hidx(t, u, v)] * R[hidx(t + tt, u + uu, v + vv)];
                }
            W[hidx(tt, uu, vv)] = (sh & 1) ? -w : w;
          }
      aux_E(ls, a, Ea);
      for (int c = 0; c < ncs; ++c) {
        int ax, ay, az;
        cart_of(ls, c, ax, ay, az);
        double val = 0.0;
        for (int tt = 0; tt <= ax; ++tt)
          for (int uu = 0; uu <= ay; ++uu)
            for (int vv = 0; vv <= az; ++vv)
              val += Ea[ax * (ls + 1) + tt] * Ea[ay * (ls + 1) + uu] * Ea[az * (ls + 1) + vv] * W[hidx(tt, uu, vv)];
        acc[c] += pref * val;
      }
    }
  }
  for (int c = 0; c < ncs; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < ncs; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < ncs) atomicAdd(&bc[aux_cart_off[s] + threadIdx.x], red[threadIdx.x][0]);
}

// 3. Hermite potential of each primitive pair from the Cartesian fitting coefficients
__global__ __launch_bounds__(128) void jpot_kernel(const JPair* __restrict__ pairs, int npair,
                                                   const DevShell* __restrict__ aux, int naux_shell,
                                                   const double* __restrict__ aux_alpha,
                                                   const double* __restrict__ aux_coef,
                                                   const int* __restrict__ aux_cart_off,
                                                   const double* __restrict__ cc, double* __restrict__ V) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npair) return;
  const JPair pr = pairs[q];
  const int Lab = pr.L;
  double acc[kJMaxHab];
  for (int i = 0; i < hcount(Lab); ++i) acc[i] = 0.0;
  double R[kJMaxH], buf[2 * kJMaxH], U[kJMaxHs], Ea[(kJMaxLAux + 1) * (kJMaxLAux + 1)];
  for (int s = 0; s < naux_shell; ++s) {
    const DevShell S = aux[s];
    const int ls = S.l;
    const double* c = cc + aux_cart_off[s];
    for (int k = 0; k < S.nprim; ++k) {
      const double a = aux_alpha[S.prim_off + k];
      aux_E(ls, a, Ea);
      // U[tau nu phi] = (-1)^(tau+nu+phi) sum_c cc_c Ea Ea Ea
      for (int i = 0; i < hcount(ls); ++i) U[i] = 0.0;
      for (int ci = 0; ci < S.ncart; ++ci) {
        int ax, ay, az;
        cart_of(ls, ci, ax, ay, az);
        const double cv = c[ci];
        if (cv == 0.0) continue;
        for (int tt = 0; tt <= ax; ++tt)
          for (int uu = 0; uu <= ay; ++uu)
            for (int vv = 0; vv <= az; ++vv)
              U[hidx(tt, uu, vv)] +=
                  cv * Ea[ax * (ls + 1) + tt] * Ea[ay * (ls + 1) + uu] * Ea[az * (ls + 1) + vv];
      }
      const double alpha = pr.p * a / (pr.p + a);
      const double pref = 2.0 * pow(M_PI, 2.5) / (pr.p * a * sqrt(pr.p + a)) * aux_coef[S.prim_off + k];
      hermite_R(Lab + ls, alpha, pr.P[0] - S.cx, pr.P[1] - S.cy, pr.P[2] - S.cz, R, buf);
      for (int sab = 0; sab <= Lab; ++sab)
        for (int t = sab; t >= 0; --t)
          for (int u = sab - t; u >= 0; --u) {
            const int v = sab - t - u;
            double w = 0.0;
            for (int sh = 0; sh <= ls; ++sh)
              for (int tt = sh; tt >= 0; --tt)
                for (int uu = sh - tt; uu >= 0; --uu) {
                  const int vv = sh - tt - uu;
                  const double r = R[hidx(t + tt, u + uu, v + vv)];
                  w += (sh & 1) ? -U[hidx(tt, uu, vv)] * r : U[hidx(tt, uu, vv)] * r;
                }
            acc[hidx(t, u, v)] += pref * w;
          }
    }
  }
  double* out = V + pr.herm_off;
  for (int i = 0; i < hcount(Lab); ++i) out[i] = acc[i];
}

// 5. J blocks of each primitive pair, angular transforms, symmetric scatter into J
__global__ __launch_bounds__(128) void jassemble_kernel(const JPair* __restrict__ pairs, int npair,
                                                        const DevShell* __restrict__ shells,
                                                        const double* __restrict__ ang,
                                                        const double* __restrict__ V, int n,
                                                        double* __restrict__ J) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npair) return;
  const JPair pr = pairs[q];
  const DevShell A = shells[pr.A], B = shells[pr.B];
  const int la = A.l, lb = B.l, nt = la + lb + 1;
  double Ex[kJMaxE], Ey[kJMaxE], Ez[kJMaxE];
  hermite_E(la, lb, pr.p, pr.PA[0], pr.PB[0], Ex);
  hermite_E(la, lb, pr.p, pr.PA[1], pr.PB[1], Ey);
  hermite_E(la, lb, pr.p, pr.PA[2], pr.PB[2], Ez);
  const double* Vq = V + pr.herm_off;
  double Jc[kJMaxNcart * kJMaxNcart];
  for (int i = 0; i < A.ncart; ++i) {
    int ax, ay, az;
    cart_of(la, i, ax, ay, az);
    for (int j = 0; j < B.ncart; ++j) {
      int bx, by, bz;
      cart_of(lb, j, bx, by, bz);
      const double* ex = Ex + (ax * (lb + 1) + bx) * nt;
      const double* ey = Ey + (ay * (lb + 1) + by) * nt;
      const double* ez = Ez + (az * (lb + 1) + bz) * nt;
      double s = 0.0;
      for (int t = 0; t <= ax + bx; ++t)
        for (int u = 0; u <= ay + by; ++u)
          for (int v = 0; v <= az + bz; ++v) s += ex[t] * ey[u] * ez[v] * Vq[hidx(t, u, v)];
      Jc[i * B.ncart + j] = pr.K * s;
    }
  }
  for (int f = 0; f < A.nfunc; ++f)
    for (int h = 0; h < B.nfunc; ++h) {
      double s = 0.0;
      for (int i = 0; i < A.ncart; ++i) {
        const double ma = ang[A.ang_off + f * A.ncart + i];
        if (ma == 0.0) continue;
        for (int j = 0; j < B.ncart; ++j) s += ma * Jc[i * B.ncart + j] * ang[B.ang_off + h * B.ncart + j];
      }
      const int mu = A.ao_off + f, nu = B.ao_off + h;
      atomicAdd(&J[mu + (int64_t)n * nu], s);
      if (pr.A != pr.B) atomicAdd(&J[nu + (int64_t)n * mu], s);
    }
}

}  // namespace

void launch_jdens(hipStream_t s, const JPair* pairs, int npair, const DevShell* shells, const double* ang,
                  const double* P, int n, double* D) {
  if (npair <= 0) return;
  hipLaunchKernelGGL(jdens_kernel, dim3((npair + 127) / 128), dim3(128), 0, s, pairs, npair, shells, ang, P, n, D);
}

void launch_jrhs(hipStream_t s, const JPair* pairs, int npair, const double* D, const DevShell* aux, int naux_shell,
                 const double* aux_alpha, const double* aux_coef, const int* aux_cart_off, double* bc) {
  if (npair <= 0 || naux_shell <= 0) return;
  const int chunk = 2048;
  const int nchunk = (npair + chunk - 1) / chunk;
  hipLaunchKernelGGL(jrhs_kernel, dim3(naux_shell, nchunk), dim3(256), 0, s, pairs, npair, chunk, D, aux, aux_alpha,
                     aux_coef, aux_cart_off, bc);
}

void launch_jpot(hipStream_t s, const JPair* pairs, int npair, const DevShell* aux, int naux_shell,
                 const double* aux_alpha, const double* aux_coef, const int* aux_cart_off, const double* cc,
                 double* V) {
  if (npair <= 0) return;
  hipLaunchKernelGGL(jpot_kernel, dim3((npair + 127) / 128), dim3(128), 0, s, pairs, npair, aux, naux_shell,
                     aux_alpha, aux_coef, aux_cart_off, cc, V);
}

void launch_jassemble(hipStream_t s, const JPair* pairs, int npair, const DevShell* shells, const double* ang,
                      const double* V, int n, double* J) {
  if (npair <= 0) return;
  hipLaunchKernelGGL(jassemble_kernel, dim3((npair + 127) / 128), dim3(128), 0, s, pairs, npair, shells, ang, V, n,
                     J);
}

}  // namespace sxc
