// HIP kernels of the XC Fock build for gfx950.
//
//   screen      block-centre shell screening (BasisFunctionOnGridController.cpp:211-255)
//   ao_eval     AO values + gradients of the significant shells into resident tiles
//               (BasisFunctionOnGridController.cpp:256-1097)
//   density     T = phi_b P_b on matrix cores, rho = rowdot(T, phi), grad rho = 2 rowdot(T, grad phi)
//               (MatrixOperatorToGridTransformer.cpp:103-165)
//   functional  per-point XC energy density and potential, block skip and E_xc partials
//               (XCFun.cpp:129-158, 751-765)
//   fock        F += phi_b^T G_b + G_b^T phi_b on matrix cores, scattered into the upper
//               triangle with FP64 atomics (ScalarOperatorToMatrixAdder.cpp:179-303)
#include <hip/hip_runtime.h>

#include "mfma.h"
#include "xc_internal.h"

namespace sxc {

namespace {

__global__ __launch_bounds__(64) void screen_kernel(const double* __restrict__ pts, int64_t npts, int bs,
                                                    const DevShell* __restrict__ shells, int nshell,
                                                    const double* __restrict__ alpha,
                                                    const double* __restrict__ coef, double thr,
                                                    unsigned char* __restrict__ flags) {
  __shared__ double c[4];
  const int b = blockIdx.x;
  const int64_t first = (int64_t)b * bs;
  const int B = (int)min((int64_t)bs, npts - first);
  if (threadIdx.x == 0) {
    // block centre (mean) and spread (largest distance to the centre),
    // accumulated point by point in order
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int p = 0; p < B; ++p) {
      sx += pts[3 * (first + p) + 0];
      sy += pts[3 * (first + p) + 1];
      sz += pts[3 * (first + p) + 2];
    }
    const double cx = sx / B, cy = sy / B, cz = sz / B;
    double sp = 0.0;
    for (int p = 0; p < B; ++p) {
      const double dx = pts[3 * (first + p) + 0] - cx;
      const double dy = pts[3 * (first + p) + 1] - cy;
      const double dz = pts[3 * (first + p) + 2] - cz;
      sp = fmax(sp, sqrt(dx * dx + dy * dy + dz * dz));
    }
    c[0] = cx;
    c[1] = cy;
    c[2] = cz;
    c[3] = sp;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nshell; s += blockDim.x) {
    const DevShell sh = shells[s];
    const double dx = c[0] - sh.cx, dy = c[1] - sh.cy, dz = c[2] - sh.cz;
    double d = sqrt(dx * dx + dy * dy + dz * dz) - c[3];
    unsigned char significant = 1;
    if (!(d < 1.0)) {
      // radial part at the closest possible distance of any block point
      d = d * d;
      double radial = 0.0;
      for (int q = 0; q < sh.nprim; ++q) radial += coef[sh.prim_off + q] * exp(-alpha[sh.prim_off + q] * d);
      if (fabs(radial) < thr) significant = 0;
    }
    flags[(int64_t)b * nshell + s] = significant;
  }
}

__global__ __launch_bounds__(128) void ao_kernel(const DevBlock* __restrict__ blocks, const int* __restrict__ sig_shell,
                                                 const int* __restrict__ sig_pos, const DevShell* __restrict__ shells,
                                                 const double* __restrict__ alpha, const double* __restrict__ coef,
                                                 const double* __restrict__ ang, const double* __restrict__ pts,
                                                 double thr, double exp_thr, int ncomp, double* __restrict__ t64,
                                                 float* __restrict__ t32) {
  const DevBlock bk = blocks[blockIdx.x];
  const int p = threadIdx.x;
  if (p >= bk.npts) return;
  const double px = pts[3 * (bk.first + p) + 0];
  const double py = pts[3 * (bk.first + p) + 1];
  const double pz = pts[3 * (bk.first + p) + 2];
  const int64_t cstride = (int64_t)bk.npad * kTile;
  for (int k = 0; k < bk.nshell; ++k) {
    const int s = sig_shell[bk.shell_off + k];
    const int pos = sig_pos[bk.shell_off + k];
    const DevShell sh = shells[s];
    const double dx = px - sh.cx, dy = py - sh.cy, dz = pz - sh.cz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    double R = 0.0, dR = 0.0;
    for (int q = 0; q < sh.nprim; ++q) {
      const double a = alpha[sh.prim_off + q];
      const double t = a * r2;
      if (t < exp_thr) {
        const double e = exp(-t);
        const double ce = coef[sh.prim_off + q] * e;
        R += ce;
        dR -= 2.0 * a * ce;
      }
    }
    const int64_t base = bk.tile_off + (int64_t)pos * kTile + p;
    if (fabs(R) < thr) {
      for (int f = 0; f < sh.nfunc; ++f)
        for (int c = 0; c < ncomp; ++c) {
          t64[base + (int64_t)f * kTile + c * cstride] = 0.0;
          if (t32) t32[base + (int64_t)f * kTile + c * cstride] = 0.0f;
        }
      continue;
    }
    double xp[kMaxL + 1], yp[kMaxL + 1], zp[kMaxL + 1];
    xp[0] = yp[0] = zp[0] = 1.0;
    for (int e = 1; e <= sh.l; ++e) {
      xp[e] = xp[e - 1] * dx;
      yp[e] = yp[e - 1] * dy;
      zp[e] = zp[e - 1] * dz;
    }
    for (int f = 0; f < sh.nfunc; ++f) {
      double Y = 0.0, Yx = 0.0, Yy = 0.0, Yz = 0.0;
      int ci = 0;
      for (int a = sh.l; a >= 0; --a)
        for (int b = sh.l - a; b >= 0; --b, ++ci) {
          const double m = ang[sh.ang_off + f * sh.ncart + ci];
          if (m == 0.0) continue;
          const int c = sh.l - a - b;
          Y += m * xp[a] * yp[b] * zp[c];
          if (a > 0) Yx += m * a * xp[a - 1] * yp[b] * zp[c];
          if (b > 0) Yy += m * b * xp[a] * yp[b - 1] * zp[c];
          if (c > 0) Yz += m * c * xp[a] * yp[b] * zp[c - 1];
        }
      const int64_t o = base + (int64_t)f * kTile;
      const double val = R * Y;
      t64[o] = val;
      if (t32) t32[o] = (float)val;
      if (ncomp > 1) {
        const double gx = R * Yx + dR * dx * Y;
        const double gy = R * Yy + dR * dy * Y;
        const double gz = R * Yz + dR * dz * Y;
        t64[o + cstride] = gx;
        t64[o + 2 * cstride] = gy;
        t64[o + 3 * cstride] = gz;
        if (t32) {
          t32[o + cstride] = (float)gx;
          t32[o + 2 * cstride] = (float)gy;
          t32[o + 3 * cstride] = (float)gz;
        }
      }
    }
  }
}

// One workgroup (4 wavefronts) per grid block.  For every group of 32 AO
// columns nu: T[p][nu] = sum_mu phi[mu][p] P[mu][nu] with wave w owning points
// 32w..32w+31 (2 x 2 MFMA tiles), phi and the gathered P chunk staged in LDS;
// then each thread dots its point's T row with phi / grad phi.
template <typename T>
__global__ __launch_bounds__(256) void density_kernel(const DevBlock* __restrict__ blocks, const int* __restrict__ ao_idx,
                                                      const T* __restrict__ tiles, const T* __restrict__ P, int n,
                                                      int want_grad, double* __restrict__ rho,
                                                      double* __restrict__ grad, int64_t npts) {
  using M = Mfma<T>;
  __shared__ T Pc[32][32];
  __shared__ T buf[32][kTile];
  __shared__ double red[4][kTile];
  const DevBlock bk = blocks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int p = tid & (kTile - 1), h = tid >> 7;
  const int npad = bk.npad;
  const T* tb = tiles + bk.tile_off;
  const int64_t cstride = (int64_t)npad * kTile;
  double racc = 0.0, gxa = 0.0, gya = 0.0, gza = 0.0;
  for (int g = 0; g < npad; g += 32) {
    typename M::acc_t acc[2][2];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) acc[r][c] = M::zero();
    for (int k0 = 0; k0 < npad; k0 += 32) {
      for (int e = tid; e < 32 * 32; e += 256) {
        const int i = e & 31, j = e >> 5;
        const int mu = ao_idx[bk.idx_off + k0 + i];
        const int nu = ao_idx[bk.idx_off + g + j];
        Pc[i][j] = (mu >= 0 && nu >= 0) ? P[mu + (int64_t)n * nu] : T(0);
      }
      const T* src = tb + (int64_t)k0 * kTile;
      for (int e = tid; e < 32 * kTile; e += 256) buf[e / kTile][e % kTile] = src[e];
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        const int kk = ks * 4 + (lane >> 4);
        const T a0 = buf[kk][32 * w + (lane & 15)];
        const T a1 = buf[kk][32 * w + 16 + (lane & 15)];
        const T b0 = Pc[kk][lane & 15];
        const T b1 = Pc[kk][16 + (lane & 15)];
        acc[0][0] = M::step(a0, b0, acc[0][0]);
        acc[0][1] = M::step(a0, b1, acc[0][1]);
        acc[1][0] = M::step(a1, b0, acc[1][0]);
        acc[1][1] = M::step(a1, b1, acc[1][1]);
      }
      __syncthreads();
    }
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c)
        for (int i = 0; i < 4; ++i) buf[16 * c + (lane & 15)][32 * w + 16 * r + M::row(lane, i)] = acc[r][c][i];
    __syncthreads();
    for (int j = 0; j < 16; ++j) {
      const int nu = 16 * h + j;
      const double tv = (double)buf[nu][p];
      const int64_t row = (int64_t)(g + nu) * kTile + p;
      racc += tv * (double)tb[row];
      if (want_grad) {
        gxa += tv * (double)tb[row + cstride];
        gya += tv * (double)tb[row + 2 * cstride];
        gza += tv * (double)tb[row + 3 * cstride];
      }
    }
    __syncthreads();
  }
  if (h == 1) {
    red[0][p] = racc;
    red[1][p] = gxa;
    red[2][p] = gya;
    red[3][p] = gza;
  }
  __syncthreads();
  if (h == 0 && p < bk.npts) {
    const int64_t q = bk.first + p;
    rho[q] = racc + red[0][p];
    if (want_grad) {
      grad[q] = 2.0 * (gxa + red[1][p]);
      grad[npts + q] = 2.0 * (gya + red[2][p]);
      grad[2 * npts + q] = 2.0 * (gza + red[3][p]);
    }
  }
}

__device__ inline double block_sum_128(double x, double* red) {
  // fixed-order tree reduction over 128 threads
  const int t = threadIdx.x;
  red[t] = x;
  __syncthreads();
  for (int s = 64; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(128) void functional_kernel(int func, XcParams prm, int64_t npts, int bs,
                                                         const double* __restrict__ rho,
                                                         const double* __restrict__ grad,
                                                         const double* __restrict__ w, double* __restrict__ eps,
                                                         double* __restrict__ v, double* __restrict__ g,
                                                         double* __restrict__ vs, double* __restrict__ gw,
                                                         double* __restrict__ e_blk) {
  __shared__ double red[128];
  const int b = blockIdx.x, t = threadIdx.x;
  const int64_t first = (int64_t)b * bs;
  const int B = (int)min((int64_t)bs, npts - first);
  const int64_t q = first + t;
  const bool in = t < B;
  const double r = in ? rho[q] : 0.0;
  const double sumabs = block_sum_128(fabs(r), red);
  const bool skip = sumabs < B * 1e-12;
  const bool gga = functional_is_gga(func);
  double e = 0.0, vv = 0.0, gg[3] = {0.0, 0.0, 0.0};
  if (in && !skip) {
    const double gx = (gga && grad) ? grad[q] : 0.0;
    const double gy = (gga && grad) ? grad[npts + q] : 0.0;
    const double gz = (gga && grad) ? grad[2 * npts + q] : 0.0;
    e = xc_point(func, r, gx, gy, gz, prm, &vv, gg);
  }
  const double wq = in ? w[q] : 0.0;
  if (in) {
    if (eps) eps[q] = e;
    if (v) v[q] = vv;
    if (g) {
      g[q] = gg[0];
      g[npts + q] = gg[1];
      g[2 * npts + q] = gg[2];
    }
    if (vs) vs[q] = wq * vv;
    if (gw) {
      gw[q] = wq * gg[0];
      gw[npts + q] = wq * gg[1];
      gw[2 * npts + q] = wq * gg[2];
    }
  }
  const double eb = block_sum_128(wq * e, red);
  if (t == 0) e_blk[b] = eb;
}

// One workgroup per grid block.  Output super-tiles of 64 x 64 compact AOs
// (I <= J); wave w owns the 32 x 32 quadrant (w / 2, w % 2) as 2 x 2 MFMA
// tiles.  S_IJ = phi_I^T G_J + G_I^T phi_J over the block's points (K = 128
// in chunks of 32), with G = v_w phi / 2 + sum_k g_w,k d_k phi built while
// staging.  S_IJ goes into the upper triangle of U with FP64 atomics.
template <typename T>
__global__ __launch_bounds__(256) void fock_kernel(const DevBlock* __restrict__ blocks, const int* __restrict__ ao_idx,
                                                   const T* __restrict__ tiles, const double* __restrict__ vs,
                                                   const double* __restrict__ gw, int gga, double thr,
                                                   double* __restrict__ U, int n, int64_t npts) {
  using M = Mfma<T>;
  __shared__ T sPI[32][64], sGI[32][64], sPJ[32][64], sGJ[32][64];
  __shared__ double spot[4][kTile];
  __shared__ double red[kTile];
  const DevBlock bk = blocks[blockIdx.x];
  if (bk.nb == 0) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int npad = bk.npad;
  const T* tb = tiles + bk.tile_off;
  const int64_t cstride = (int64_t)npad * kTile;
  double part = 0.0;
  if (tid < kTile) {
    const bool in = tid < bk.npts;
    const int64_t q = bk.first + tid;
    spot[0][tid] = in ? vs[q] : 0.0;
    spot[1][tid] = (in && gga) ? gw[q] : 0.0;
    spot[2][tid] = (in && gga) ? gw[npts + q] : 0.0;
    spot[3][tid] = (in && gga) ? gw[2 * npts + q] : 0.0;
    part = fabs(spot[0][tid]) + fabs(spot[1][tid]) + fabs(spot[2][tid]) + fabs(spot[3][tid]);
  }
  // block average of |w v| (+ |w g|) for the prescreening; all threads see the same value
  if (tid < kTile) red[tid] = part;
  __syncthreads();
  for (int s = 64; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double average = red[0] / bk.npts;
  if (average < thr) return;

  const int ngroups = (npad + 63) / 64;
  const int rowbase = 32 * (w >> 1), colbase = 32 * (w & 1);
  for (int I = 0; I < ngroups; ++I) {
    for (int J = I; J < ngroups; ++J) {
      typename M::acc_t acc[2][2];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) acc[r][c] = M::zero();
      for (int p0 = 0; p0 < kTile; p0 += 32) {
        for (int e = tid; e < 64 * 32; e += 256) {
          const int i = e >> 5, pp = e & 31;
          const int pt = p0 + pp;
          const double sv = 0.5 * spot[0][pt];
#pragma unroll
          for (int side = 0; side < 2; ++side) {
            const int mu = 64 * (side ? J : I) + i;
            T phi = T(0), G = T(0);
            if (mu < npad) {
              const T* src = tb + (int64_t)mu * kTile + pt;
              phi = src[0];
              double gv = sv * (double)phi;
              if (gga)
                gv += spot[1][pt] * (double)src[cstride] + spot[2][pt] * (double)src[2 * cstride] +
                      spot[3][pt] * (double)src[3 * cstride];
              G = (T)gv;
            }
            if (side) {
              sPJ[pp][i] = phi;
              sGJ[pp][i] = G;
            } else {
              sPI[pp][i] = phi;
              sGI[pp][i] = G;
            }
          }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
          const int kk = ks * 4 + (lane >> 4);
          T aP[2], aG[2], bP[2], bG[2];
          for (int r = 0; r < 2; ++r) {
            aP[r] = sPI[kk][rowbase + 16 * r + (lane & 15)];
            aG[r] = sGI[kk][rowbase + 16 * r + (lane & 15)];
          }
          for (int c = 0; c < 2; ++c) {
            bP[c] = sPJ[kk][colbase + 16 * c + (lane & 15)];
            bG[c] = sGJ[kk][colbase + 16 * c + (lane & 15)];
          }
          for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) {
              acc[r][c] = M::step(aP[r], bG[c], acc[r][c]);
              acc[r][c] = M::step(aG[r], bP[c], acc[r][c]);
            }
        }
        __syncthreads();
      }
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
          for (int i = 0; i < 4; ++i) {
            const int row = 64 * I + rowbase + 16 * r + M::row(lane, i);
            const int col = 64 * J + colbase + 16 * c + (lane & 15);
            if (row >= npad || col >= npad) continue;
            const int gr = ao_idx[bk.idx_off + row], gc = ao_idx[bk.idx_off + col];
            if (gr < 0 || gc < 0) continue;
            if (I == J && gr > gc) continue;
            unsafeAtomicAdd(U + gr + (int64_t)n * gc, (double)acc[r][c][i]);
          }
    }
  }
}

__global__ void weight_potential_kernel(int64_t npts, const double* __restrict__ w, const double* __restrict__ v,
                                        const double* __restrict__ g, double* __restrict__ vs,
                                        double* __restrict__ gw) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  vs[q] = w[q] * v[q];
  if (g) {
    gw[q] = w[q] * g[q];
    gw[npts + q] = w[q] * g[npts + q];
    gw[2 * npts + q] = w[q] * g[2 * npts + q];
  }
}

__global__ void fill_kernel(double* __restrict__ p, int64_t n, double value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ void to_float_kernel(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

__global__ void finalize_kernel(const double* __restrict__ U, double* __restrict__ F, int n, int accumulate) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e % n), j = (int)(e / n);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const double u = U[lo + (int64_t)n * hi];
  F[e] = accumulate ? F[e] + u : u;
}

__global__ __launch_bounds__(256) void sum_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int64_t i = t; i < n; i += 256) s += x[i];
  red[t] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) out[0] = red[0];
}

inline unsigned grid1d(int64_t n, int bsz) { return (unsigned)((n + bsz - 1) / bsz); }

}  // namespace

void launch_screen(hipStream_t s, const double* pts, int64_t npts, int bs, int nblocks, const DevShell* shells,
                   int nshell, const double* alpha, const double* coef, double radial_thr, unsigned char* flags) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(screen_kernel, dim3(nblocks), dim3(64), 0, s, pts, npts, bs, shells, nshell, alpha, coef,
                     radial_thr, flags);
}

void launch_ao_eval(hipStream_t s, const DevBlock* blocks, int nblocks, const int* sig_shell, const int* sig_pos,
                    const DevShell* shells, const double* alpha, const double* coef, const double* ang,
                    const double* pts, double radial_thr, double exp_thr, int ncomp, double* t64, float* t32) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(ao_kernel, dim3(nblocks), dim3(kTile), 0, s, blocks, sig_shell, sig_pos, shells, alpha, coef,
                     ang, pts, radial_thr, exp_thr, ncomp, t64, t32);
}

template <typename T>
void launch_density(hipStream_t s, const DevBlock* blocks, int nblocks, const int* ao_idx, const T* tiles, int ncomp,
                    const T* P, int n, int want_grad, double* rho, double* grad, int64_t npts) {
  (void)ncomp;
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(density_kernel<T>, dim3(nblocks), dim3(256), 0, s, blocks, ao_idx, tiles, P, n, want_grad, rho,
                     grad, npts);
}
template void launch_density<double>(hipStream_t, const DevBlock*, int, const int*, const double*, int, const double*,
                                     int, int, double*, double*, int64_t);
template void launch_density<float>(hipStream_t, const DevBlock*, int, const int*, const float*, int, const float*,
                                    int, int, double*, double*, int64_t);

void launch_functional(hipStream_t s, int func, const XcParams& prm, int64_t npts, int bs, const double* rho,
                       const double* grad, const double* w, double* eps, double* v, double* g, double* vs, double* gw,
                       double* e_blk) {
  const int nb = (int)((npts + bs - 1) / bs);
  if (nb <= 0) return;
  hipLaunchKernelGGL(functional_kernel, dim3(nb), dim3(128), 0, s, func, prm, npts, bs, rho, grad, w, eps, v, g, vs,
                     gw, e_blk);
}

template <typename T>
void launch_fock(hipStream_t s, const DevBlock* blocks, int nblocks, const int* ao_idx, const T* tiles, int ncomp,
                 const double* vs, const double* gw, int gga, double thr, double* U, int n, int64_t npts) {
  (void)ncomp;
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(fock_kernel<T>, dim3(nblocks), dim3(256), 0, s, blocks, ao_idx, tiles, vs, gw, gga, thr, U, n,
                     npts);
}
template void launch_fock<double>(hipStream_t, const DevBlock*, int, const int*, const double*, int, const double*,
                                  const double*, int, double, double*, int, int64_t);
template void launch_fock<float>(hipStream_t, const DevBlock*, int, const int*, const float*, int, const double*,
                                 const double*, int, double, double*, int, int64_t);

void launch_weight_potential(hipStream_t s, int64_t npts, const double* w, const double* v, const double* g,
                             double* vs, double* gw) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(weight_potential_kernel, dim3(grid1d(npts, 256)), dim3(256), 0, s, npts, w, v, g, vs, gw);
}

void launch_fill(hipStream_t s, double* p, int64_t n, double value) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, p, n, value);
}

void launch_to_float(hipStream_t s, const double* src, float* dst, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(to_float_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, src, dst, n);
}

void launch_finalize(hipStream_t s, const double* U, double* F, int n, int accumulate) {
  const int64_t nn = (int64_t)n * n;
  if (nn <= 0) return;
  hipLaunchKernelGGL(finalize_kernel, dim3(grid1d(nn, 256)), dim3(256), 0, s, U, F, n, accumulate);
}

void launch_sum(hipStream_t s, const double* x, int64_t n, double* out) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, x, n, out);
}

}  // namespace sxc
