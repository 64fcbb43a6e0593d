// XC contribution to the nuclear gradient at a fixed grid
// (FuncPotential::getGeomGradients, src/potentials/FuncPotential.cpp:113-239).
//
// For an AO nu on atom A and direction c the reference accumulates, over the
// block's points and the block's AOs mu,
//   grad[A][c] -= 2 P_mu,nu sum_p w v phi_mu d_c phi_nu
//   (GGA)      -= 2 P_mu,nu sum_p w sum_k g_k (phi_mu d_k d_c phi_nu + d_k phi_mu d_c phi_nu).
// With G_mu(p) = w v phi_mu + sum_k w g_k d_k phi_mu and H^c_nu(p) = sum_k w g_k d_k d_c phi_nu
// this is grad[A][c] = -2 sum_{nu on A} a^c_nu with
//   a^c_nu = sum_p (G P)_p,nu d_c phi_nu(p) + (phi P)_p,nu H^c_nu(p),
// i.e. two dense products per block (phi_b P_b and G_b P_b, FP64 matrix cores) and
// column dots against the gradient and Hessian tiles.  Nothing n x n is scattered:
// each work item adds 3 numbers per AO column into a[nu][c].
#include <hip/hip_runtime.h>

#include "mfma.h"
#include "xc_internal.h"

namespace sxc {

namespace {

// One workgroup of 4 wavefronts per (block, 64-column pass).  Wavefront w owns
// points 32 w .. 32 w + 31 (two 16-row tiles) and all four 16-column tiles of
// the pass: 2 x 4 accumulators of phi P and of G P.  The AO (K) dimension
// streams through LDS in chunks of 16: phi and G point-major (G formed from the
// value and gradient tiles on the way in), and the pass's 16 x 64 slice of P.
__global__ __launch_bounds__(256) void gradient_kernel(const GradItem* __restrict__ work,
                                                       const DevBlock* __restrict__ blocks,
                                                       const int* __restrict__ ao_idx,
                                                       const double* __restrict__ tiles,
                                                       const double* __restrict__ P, int n,
                                                       const double* __restrict__ vs, const double* __restrict__ gw,
                                                       int gga, int64_t npts, double* __restrict__ acc) {
  using M = Mfma<double>;
  constexpr int KC = 16, RS = KC + 1, NC = 64, PR = NC + 1;
  __shared__ double sPhi[kTile][RS];
  __shared__ double sG[kTile][RS];
  __shared__ double sP[KC][PR];
  __shared__ double sRed[4][NC][3];
  const GradItem it = work[blockIdx.x];
  const DevBlock bk = blocks[it.block];
  const int npad = bk.npad, col0 = it.col0;
  const int ncols = min(NC, npad - col0);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r16 = lane & 15, g4 = lane >> 4;
  const double* tb = tiles + bk.tile_off;
  const int64_t cs = (int64_t)npad * kTile;
  const int* idx = ao_idx + bk.idx_off;
  // staging role: point sp, AO half sh (8 AOs of the chunk)
  const int sp = tid & (kTile - 1), shf = tid >> 7;
  const bool pin = sp < bk.npts;
  const int64_t q = bk.first + sp;
  const double vsp = pin ? vs[q] : 0.0;
  const double gxp = (pin && gga) ? gw[q] : 0.0;
  const double gyp = (pin && gga) ? gw[npts + q] : 0.0;
  const double gzp = (pin && gga) ? gw[2 * npts + q] : 0.0;

  M::acc_t T[2][4], TG[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      T[a][b] = M::zero();
      TG[a][b] = M::zero();
    }

  for (int k0 = 0; k0 < npad; k0 += KC) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kk = shf * 8 + j;
      const int64_t o = (int64_t)(k0 + kk) * kTile + sp;
      const double phi = tb[o];
      double g = vsp * phi;
      if (gga) g += gxp * tb[cs + o] + gyp * tb[2 * cs + o] + gzp * tb[3 * cs + o];
      sPhi[sp][kk] = phi;
      sG[sp][kk] = g;
    }
    for (int e = tid; e < KC * NC; e += 256) {
      const int kk = e & (KC - 1), c = e >> 4;
      const int mu = idx[k0 + kk];
      const int nu = c < ncols ? idx[col0 + c] : -1;
      sP[kk][c] = (mu >= 0 && nu >= 0) ? P[mu + (int64_t)n * nu] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < KC / 4; ++t) {
      const int k = 4 * t + g4;
      double ap[2], ag[2], bp[4];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        ap[a] = sPhi[32 * wv + 16 * a + r16][k];
        ag[a] = sG[32 * wv + 16 * a + r16][k];
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) bp[b] = sP[k][16 * b + r16];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          T[a][b] = M::step(ap[a], bp[b], T[a][b]);
          TG[a][b] = M::step(ag[a], bp[b], TG[a][b]);
        }
    }
  }

  // column dots: lane (r16, g4) holds column 16 b + r16, rows 16 a + g4 + 4 r of the wavefront
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int cl = 16 * b + r16;  // column within the pass
    const int i = col0 + cl;      // compact AO index
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (cl < ncols) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int p = 32 * wv + 16 * a + g4 + 4 * r;
          if (p >= bk.npts) continue;
          const int64_t o = (int64_t)i * kTile + p;
          const double tg = TG[a][b][r], tv = T[a][b][r];
          const double dx = tb[cs + o], dy = tb[2 * cs + o], dz = tb[3 * cs + o];
          ax += tg * dx;
          ay += tg * dy;
          az += tg * dz;
          if (gga) {
            const int64_t qp = bk.first + p;
            const double wx = gw[qp], wy = gw[npts + qp], wz = gw[2 * npts + qp];
            const double hxx = tb[4 * cs + o], hxy = tb[5 * cs + o], hxz = tb[6 * cs + o];
            const double hyy = tb[7 * cs + o], hyz = tb[8 * cs + o], hzz = tb[9 * cs + o];
            ax += tv * (wx * hxx + wy * hxy + wz * hxz);
            ay += tv * (wx * hxy + wy * hyy + wz * hyz);
            az += tv * (wx * hxz + wy * hyz + wz * hzz);
          }
        }
    }
    // the four lane groups hold the same column
    ax += __shfl_xor(ax, 16);
    ay += __shfl_xor(ay, 16);
    az += __shfl_xor(az, 16);
    ax += __shfl_xor(ax, 32);
    ay += __shfl_xor(ay, 32);
    az += __shfl_xor(az, 32);
    if (g4 == 0) {
      sRed[wv][cl][0] = ax;
      sRed[wv][cl][1] = ay;
      sRed[wv][cl][2] = az;
    }
  }
  __syncthreads();
  if (tid < 3 * ncols) {
    const int cl = tid / 3, c = tid - 3 * cl;
    const int mu = idx[col0 + cl];
    if (mu >= 0) {
      const double s = sRed[0][cl][c] + sRed[1][cl][c] + sRed[2][cl][c] + sRed[3][cl][c];
      atomicAdd(&acc[3 * (int64_t)mu + c], s);
    }
  }
}

// grad[A + natoms c] = -2 sum_{nu : atom(nu) = A} acc[nu][c], AOs in order.
__global__ void gradient_atoms_kernel(const double* __restrict__ acc, const int* __restrict__ ao_atom, int n,
                                      int natoms, double* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * natoms) return;
  const int A = t % natoms, c = t / natoms;
  double s = 0.0;
  for (int nu = 0; nu < n; ++nu)
    if (ao_atom[nu] == A) s += acc[3 * (int64_t)nu + c];
  grad[t] = -2.0 * s;
}

}  // namespace

void launch_gradient(hipStream_t s, const GradItem* work, int nwork, const DevBlock* blocks, const int* ao_idx,
                     const double* tiles, const double* P, int n, const double* vs, const double* gw, int gga,
                     int64_t npts, double* acc) {
  if (nwork <= 0) return;
  hipLaunchKernelGGL(gradient_kernel, dim3(nwork), dim3(256), 0, s, work, blocks, ao_idx, tiles, P, n, vs, gw, gga,
                     npts, acc);
}

void launch_gradient_atoms(hipStream_t s, const double* acc, const int* ao_atom, int n, int natoms, double* grad) {
  const int nt = 3 * natoms;
  if (nt <= 0) return;
  hipLaunchKernelGGL(gradient_atoms_kernel, dim3((nt + 127) / 128), dim3(128), 0, s, acc, ao_atom, n, natoms, grad);
}

}  // namespace sxc
