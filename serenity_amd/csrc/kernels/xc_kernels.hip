// HIP kernels of the XC Fock build for gfx950.
//
//   screen      block-centre shell screening (BasisFunctionOnGridController.cpp:211-255)
//   ao_eval     AO values + gradients of the significant shells into resident tiles
//               (BasisFunctionOnGridController.cpp:256-1097)
//   density     T = phi_b P_b on matrix cores, rho = rowdot(T, phi), grad rho = 2 rowdot(T, grad phi)
//               (MatrixOperatorToGridTransformer.cpp:103-165)
//   functional  per-point XC energy density and potential, block skip and E_xc partials
//               (XCFun.cpp:129-158, 751-765)
//   fock        S_b = phi_b^T G_b + G_b^T phi_b on matrix cores, the upper-triangle sub-tiles
//               of every block stored to a scratch (ScalarOperatorToMatrixAdder.cpp:179-303)
//   fock_reduce F = sum over blocks of S_b scattered by the blocks' AO lists, one AO row per
//               workgroup summing its blocks in a fixed order (the reference's thread-private
//               F plus serial sum, ScalarOperatorToMatrixAdder.cpp:101-110)
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <type_traits>

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
    double R = 0.0, dR = 0.0, ddR = 0.0;
    for (int q = 0; q < sh.nprim; ++q) {
      const double a = alpha[sh.prim_off + q];
      const double t = a * r2;
      if (t < exp_thr) {
        const double e = exp(-t);
        const double ce = coef[sh.prim_off + q] * e;
        R += ce;
        dR -= 2.0 * a * ce;
        ddR += 4.0 * a * a * ce;
      }
    }
    const int64_t base = bk.tile_off + (int64_t)pos * kTile + p;
    if (fabs(R) < thr) {
      for (int f = 0; f < sh.nfunc; ++f)
        for (int c = 0; c < ncomp; ++c) {
          if (t64) t64[base + (int64_t)f * kTile + c * cstride] = 0.0;
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
    const bool hess = ncomp >= 10;
    for (int f = 0; f < sh.nfunc; ++f) {
      double Y = 0.0, Yx = 0.0, Yy = 0.0, Yz = 0.0;
      double Yxx = 0.0, Yxy = 0.0, Yxz = 0.0, Yyy = 0.0, Yyz = 0.0, Yzz = 0.0;
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
          if (hess) {
            if (a > 1) Yxx += m * a * (a - 1) * xp[a - 2] * yp[b] * zp[c];
            if (b > 1) Yyy += m * b * (b - 1) * xp[a] * yp[b - 2] * zp[c];
            if (c > 1) Yzz += m * c * (c - 1) * xp[a] * yp[b] * zp[c - 2];
            if (a > 0 && b > 0) Yxy += m * a * b * xp[a - 1] * yp[b - 1] * zp[c];
            if (a > 0 && c > 0) Yxz += m * a * c * xp[a - 1] * yp[b] * zp[c - 1];
            if (b > 0 && c > 0) Yyz += m * b * c * xp[a] * yp[b - 1] * zp[c - 1];
          }
        }
      const int64_t o = base + (int64_t)f * kTile;
      const double val = R * Y;
      if (t64) t64[o] = val;
      if (t32) t32[o] = (float)val;
      if (ncomp > 1) {
        const double gx = R * Yx + dR * dx * Y;
        const double gy = R * Yy + dR * dy * Y;
        const double gz = R * Yz + dR * dz * Y;
        if (t64) {
          t64[o + cstride] = gx;
          t64[o + 2 * cstride] = gy;
          t64[o + 3 * cstride] = gz;
        }
        if (t32) {
          t32[o + cstride] = (float)gx;
          t32[o + 2 * cstride] = (float)gy;
          t32[o + 3 * cstride] = (float)gz;
        }
      }
      if (hess) {
        // d_i d_j (R Y) = R Y_ij + R' (x_i Y_j + x_j Y_i) + d_ij R' Y + R'' x_i x_j Y,
        // R' = sum -2 a c e^(-a r^2), R'' = sum 4 a^2 c e^(-a r^2); planes xx xy xz yy yz zz
        const double h[6] = {R * Yxx + dR * (2.0 * dx * Yx + Y) + ddR * dx * dx * Y,
                             R * Yxy + dR * (dx * Yy + dy * Yx) + ddR * dx * dy * Y,
                             R * Yxz + dR * (dx * Yz + dz * Yx) + ddR * dx * dz * Y,
                             R * Yyy + dR * (2.0 * dy * Yy + Y) + ddR * dy * dy * Y,
                             R * Yyz + dR * (dy * Yz + dz * Yy) + ddR * dy * dz * Y,
                             R * Yzz + dR * (2.0 * dz * Yz + Y) + ddR * dz * dz * Y};
        for (int c = 0; c < 6; ++c) {
          if (t64) t64[o + (4 + c) * cstride] = h[c];
          if (t32) t32[o + (4 + c) * cstride] = (float)h[c];
        }
      }
    }
  }
}

// The four FP32 tile values of a lane's MFMA accumulator rows: points 4 (lane / 16) + e,
// e = 0..3, of a 16-point row segment -- one aligned 16-byte load.
__device__ inline void quad_rows(const gmem<const float>* seg, int lane, double f[4]) {
  const f32x4 v = *reinterpret_cast<const gmem<const f32x4>*>(seg + 4 * (lane >> 4));
  f[0] = v.x;
  f[1] = v.y;
  f[2] = v.z;
  f[3] = v.w;
}

// Density contraction on 16 x 16 sub-tiles.  One workgroup of 8 wavefronts
// per (block, column pass): T[p][nu] = sum_mu phi[mu][p] P[mu][nu] for the
// pass's 16 S columns, wavefront w owning points 16 w .. 16 w + 15 and all S
// column tiles (S independent accumulators).  The AO (K) dimension streams in
// chunks of KC (32 FP32 / 16 FP64) through an LDS buffer that holds phi
// point-major and the pass's P columns, both with the AO index fastest: with
// k-step t of lane group g taking AO KV g + t (KV = KC / 4) every MFMA operand
// of a chunk is one 32-byte LDS read.  The global loads of chunk k + 1 are
// issued before the MFMAs of chunk k.  The pass's partial rho and
// sum_nu T grad phi are then formed from the accumulators in registers, each
// lane dotting its T entries with the phi / grad phi values of the same
// (point, column), and summed over the columns by lane shuffles.  (FP64
// accumulates T transposed, so that its lanes read whole 128-byte point runs.)
template <typename T, int S>
__global__ __launch_bounds__(512) void density_sub_kernel(const WorkItem* __restrict__ work,
                                                          const DevBlock* __restrict__ blocks,
                                                          const int* __restrict__ ao_idx,
                                                          const T* __restrict__ tiles, const T* __restrict__ P,
                                                          int n, int want_grad, double* __restrict__ part,
                                                          int64_t npts, const PackSys* __restrict__ pack,
                                                          const int* __restrict__ nwork_dev,
                                                          const unsigned char* __restrict__ skip) {
  // a launch over a list filled on the device runs only its first *nwork_dev items
  if (nwork_dev && (int)blockIdx.x >= *nwork_dev) return;
  using M = Mfma<T>;
  constexpr int KC = 128 / (int)sizeof(T);      // AOs per chunk (128 bytes of each LDS row)
  constexpr int KV = KC / 4;
  constexpr int RS = KC + 16 / (int)sizeof(T);  // 144-byte rows
  constexpr int NC = 16 * S;
  constexpr int NB = (4 * NC + 511) / 512;      // P staging tasks per thread
  constexpr int GS = 2;
  constexpr int STAGE_BYTES = (kTile + NC) * RS * (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(KV)));
  __shared__ __attribute__((aligned(16))) unsigned char smem[STAGE_BYTES];
  __shared__ int sIdx[kDensIdxLds];
  T(*sA)[RS] = reinterpret_cast<T(*)[RS]>(smem);
  T(*sB)[RS] = reinterpret_cast<T(*)[RS]>(smem + kTile * RS * sizeof(T));

  const WorkItem wk = work[blockIdx.x];
  if (skip && skip[wk.block]) return;  // a block the caller contracts in another pass
  const DevBlock bk = blocks[wk.block];
  if (pack) {
    const PackSys& ps = pack[bk.sys];
    tiles = static_cast<const T*>(ps.tiles);
    P = static_cast<const T*>(ps.P);
    ao_idx = ps.ao_idx;
    part = ps.part;
    npts = ps.npts;
    n = ps.n;
  }
  const int col0 = wk.a, ncols = wk.b;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, g = lane >> 4;
  const int npad = bk.npad;
  const gmem<const T>* tb = to_global(tiles + bk.tile_off);
  const gmem<const int>* idx = to_global(ao_idx + bk.idx_off);
  const gmem<const T>* Pg = to_global(P);
  const int64_t cstride = (int64_t)npad * kTile;
  // staging: phi task = (point pa, AO quad qa); P tasks = (column cb, AO quad qb), 16
  // consecutive lanes taking 16 columns of one quad (their LDS stores then cover distinct banks)
  const int pa = tid & (kTile - 1), qa = tid >> 7;
  const int qb = (tid >> 4) & 3;
  auto col_of = [](int task) { return (task & 15) + 16 * (task >> 6); };
  int nu[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int task = tid + 512 * m, cb = col_of(task);
    nu[m] = (task < 4 * NC && cb < ncols) ? idx[col0 + cb] : -1;
  }
  typename M::acc_t acc[S];
#pragma unroll
  for (int c = 0; c < S; ++c) acc[c] = M::zero();
  vec_t rA, rB[NB];
  unsigned mu_ok = 0;  // bit e: AO KV qb + e of the chunk is not padding
  // chunk at AO k0 of the block, its AO indices at sIdx[j0 ..]
  auto load = [&](int k0, int j0) {
#pragma unroll
    for (int e = 0; e < KV; ++e) rA[e] = tb[(int64_t)(k0 + KV * qa + e) * kTile + pa];
    int mu[KV];
    mu_ok = 0;
#pragma unroll
    for (int e = 0; e < KV; ++e) {
      mu[e] = sIdx[j0 + KV * qb + e];
      mu_ok |= (unsigned)(mu[e] >= 0) << e;
    }
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      if (tid + 512 * m < 4 * NC) {  // whole wavefronts
        // P is symmetric: read P[nu][mu] so that the 16 lanes of a group (consecutive
        // columns nu, one mu) read consecutive addresses.  Padding (index -1) reads P[0][0]
        // and is zeroed at the LDS store by a multiply: a select at the load would let the
        // compiler move the load into a branch, and a multiply there would wait for it
        // before the chunk's MFMAs.
#pragma unroll
        for (int e = 0; e < KV; ++e) rB[m][e] = Pg[max(nu[m], 0) + (int64_t)n * max(mu[e], 0)];
      }
    }
  };
  auto store = [&]() {
    *reinterpret_cast<vec_t*>(&sA[pa][KV * qa]) = rA;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int task = tid + 512 * m;
      if (task < 4 * NC) {
        vec_t v = rB[m];
#pragma unroll
        for (int e = 0; e < KV; ++e) v[e] *= T(((mu_ok >> e) & 1u) && nu[m] >= 0);
        *reinterpret_cast<vec_t*>(&sB[col_of(task)][KV * qb]) = v;
      }
    }
  };
  // The block's AO index list goes to LDS (the P rows of a chunk are then gathered without a
  // dependent global load) in windows of kDensIdxLds entries: one window unless the block is
  // wider.  One LDS buffer; the global loads of chunk k + 1 of a window are in flight during the
  // MFMAs of chunk k.
  for (int wb = 0; wb < npad; wb += kDensIdxLds) {
    const int wl = min(kDensIdxLds, npad - wb);
    for (int i = tid; i < wl; i += 512) sIdx[i] = idx[wb + i];
    __syncthreads();
    const int nch = wl / KC;
    load(wb, 0);
    for (int ch = 0; ch < nch; ++ch) {
      store();
      __syncthreads();
      if (ch + 1 < nch) load(wb + (ch + 1) * KC, (ch + 1) * KC);
      const vec_t av = *reinterpret_cast<const vec_t*>(&sA[16 * w + r][KV * g]);
#pragma unroll
      for (int s0 = 0; s0 < S; s0 += GS) {
        vec_t bv[GS];
#pragma unroll
        for (int u = 0; u < GS; ++u)
          if (s0 + u < S) bv[u] = *reinterpret_cast<const vec_t*>(&sB[16 * (s0 + u) + r][KV * g]);
#pragma unroll
        for (int t = 0; t < KV; ++t)
#pragma unroll
          for (int u = 0; u < GS; ++u)
            if (s0 + u < S) {
              // FP64 accumulates T transposed (rows: AO columns, columns: points), so that a
              // lane's entries are one point of four AO columns in the dot phase
              if constexpr (sizeof(T) == 8)
                acc[s0 + u] = M::step(bv[u][t], av[t], acc[s0 + u]);
              else
                acc[s0 + u] = M::step(av[t], bv[u][t], acc[s0 + u]);
            }
        __builtin_amdgcn_sched_barrier(0);
      }
      __syncthreads();  // (also orders the next window's index staging after this window's reads)
    }
  }

  if constexpr (sizeof(T) == 4) {
    // Dots straight from the accumulators: lane (r, g) holds T[row(lane, e)][16 s + r] of its
    // wavefront's 16 points for every column tile s, and reads the matching phi / grad phi of
    // those 4 points from the tiles (FP32: one 16-byte load per component).  All loads of the
    // phase are independent of each other and of LDS, so they are in flight together.
    double rv[4] = {0.0, 0.0, 0.0, 0.0}, xv[4] = {0.0, 0.0, 0.0, 0.0};
    double yv[4] = {0.0, 0.0, 0.0, 0.0}, zv[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < S; ++c) {
      if (16 * c < ncols) {
        const gmem<const T>* row = tb + (int64_t)(col0 + 16 * c + r) * kTile + 16 * w;
        double f[4];
        quad_rows(row, lane, f);
#pragma unroll
        for (int e = 0; e < 4; ++e) rv[e] += (double)acc[c][e] * f[e];
        if (want_grad) {
          double fx[4], fy[4], fz[4];
          quad_rows(row + cstride, lane, fx);
          quad_rows(row + 2 * cstride, lane, fy);
          quad_rows(row + 3 * cstride, lane, fz);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            xv[e] += (double)acc[c][e] * fx[e];
            yv[e] += (double)acc[c][e] * fy[e];
            zv[e] += (double)acc[c][e] * fz[e];
          }
        }
      }
    }
    // sum over the 16 lanes of a lane group (the columns), a fixed butterfly order
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        rv[e] += __shfl_xor(rv[e], m);
        if (want_grad) {
          xv[e] += __shfl_xor(xv[e], m);
          yv[e] += __shfl_xor(yv[e], m);
          zv[e] += __shfl_xor(zv[e], m);
        }
      }
    }
    if (r == 0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int p = 16 * w + M::row(lane, e);
        if (p < bk.npts) {
          gmem<double>* dst = to_global(part + (int64_t)wk.pass * 4 * npts + bk.first + p);
          dst[0] = rv[e];
          if (want_grad) {
            dst[npts] = xv[e];
            dst[2 * npts] = yv[e];
            dst[3 * npts] = zv[e];
          }
        }
      }
    }
  } else {
    // lane (r, g) holds T[point 16 w + r][AO column 16 c + g + 4 e]: per e the dot reads 16
    // consecutive points (128 bytes) of each of four tile rows
    double rv = 0.0, xv = 0.0, yv = 0.0, zv = 0.0;
#pragma unroll
    for (int c = 0; c < S; ++c) {
      if (16 * c < ncols) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const gmem<const T>* row = tb + (int64_t)(col0 + 16 * c + g + 4 * e) * kTile + 16 * w + r;
          const double tv = acc[c][e];
          rv += tv * row[0];
          if (want_grad) {
            xv += tv * row[cstride];
            yv += tv * row[2 * cstride];
            zv += tv * row[3 * cstride];
          }
        }
      }
    }
    // sum over the four lane groups (the AO columns g + 4 e of each tile)
    rv += __shfl_xor(rv, 16);
    rv += __shfl_xor(rv, 32);
    if (want_grad) {
      xv += __shfl_xor(xv, 16);
      xv += __shfl_xor(xv, 32);
      yv += __shfl_xor(yv, 16);
      yv += __shfl_xor(yv, 32);
      zv += __shfl_xor(zv, 16);
      zv += __shfl_xor(zv, 32);
    }
    const int p = 16 * w + r;
    if (g == 0 && p < bk.npts) {
      gmem<double>* dst = to_global(part + (int64_t)wk.pass * 4 * npts + bk.first + p);
      dst[0] = rv;
      if (want_grad) {
        dst[npts] = xv;
        dst[2 * npts] = yv;
        dst[3 * npts] = zv;
      }
    }
  }
}

__global__ void reduce_partials_kernel(const DevBlock* __restrict__ blocks, int bs, int64_t npts,
                                       const double* __restrict__ part, int want_grad, double* __restrict__ rho,
                                       double* __restrict__ grad, int accumulate) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const int np = blocks[q / bs].npass;
  double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
  for (int s = 0; s < np; ++s) {
    const double* src = part + (int64_t)s * 4 * npts + q;
    r += src[0];
    if (want_grad) {
      x += src[npts];
      y += src[2 * npts];
      z += src[3 * npts];
    }
  }
  if (accumulate) {
    rho[q] += r;
    if (want_grad) {
      grad[q] += 2.0 * x;
      grad[npts + q] += 2.0 * y;
      grad[2 * npts + q] += 2.0 * z;
    }
  } else {
    rho[q] = r;
    if (want_grad) {
      grad[q] = 2.0 * x;
      grad[npts + q] = 2.0 * y;
      grad[2 * npts + q] = 2.0 * z;
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
                                                         const DevBlock* __restrict__ blocks,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ w, double* __restrict__ eps,
                                                         double* __restrict__ v, double* __restrict__ g,
                                                         double* __restrict__ vs, double* __restrict__ gw,
                                                         double* __restrict__ e_blk,
                                                         const double* __restrict__ rho_sub,
                                                         const double* __restrict__ grad_sub,
                                                         const PackSys* __restrict__ pack) {
  __shared__ double red[128];
  const int bm = blockIdx.x, t = threadIdx.x;  // bm: block in the (merged) block list
  int b = bm;                                  // block of this system's grid
  if (pack) {
    const PackSys& ps = pack[blocks[bm].sys];
    b = bm - ps.blk0;
    npts = ps.npts;
    part = ps.part;
    w = ps.w;
    vs = ps.vs;
    gw = ps.gw;
    e_blk = ps.eblk;
  }
  const int64_t first = (int64_t)b * bs;
  const int B = (int)min((int64_t)bs, npts - first);
  const int64_t q = first + t;
  const bool in = t < B;
  const bool gga = functional_is_gga(func);
  double r = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  if (in) {
    if (part) {
      // sum the density passes of this block (blocks of 128 points)
      const int np = blocks[bm].npass;
      for (int s = 0; s < np; ++s) {
        const double* src = part + (int64_t)s * 4 * npts + q;
        r += src[0];
        if (gga) {
          gx += src[npts];
          gy += src[2 * npts];
          gz += src[3 * npts];
        }
      }
      gx *= 2.0;
      gy *= 2.0;
      gz *= 2.0;
    } else {
      r = rho[q];
      if (gga && grad) {
        gx = grad[q];
        gy = grad[npts + q];
        gz = grad[2 * npts + q];
      }
    }
  }
  const double sumabs = block_sum_128(fabs(r), red);
  const bool skip = sumabs < B * 1e-12;
  double e = 0.0, vv = 0.0, gg[3] = {0.0, 0.0, 0.0};
  if (in && !skip) e = xc_point(func, r, gx, gy, gz, prm, &vv, gg);
  if (rho_sub) {
    // non-additive form: the same functional of a second density is subtracted
    // (its own block skip applies to it)
    const double r2 = in ? rho_sub[q] : 0.0;
    const bool skip2 = block_sum_128(fabs(r2), red) < B * 1e-12;
    if (in && !skip2) {
      double v2 = 0.0, g2[3] = {0.0, 0.0, 0.0};
      const bool has_g = gga && grad_sub;
      const double e2 = xc_point(func, r2, has_g ? grad_sub[q] : 0.0, has_g ? grad_sub[npts + q] : 0.0,
                                 has_g ? grad_sub[2 * npts + q] : 0.0, prm, &v2, g2);
      e -= e2;
      vv -= v2;
      gg[0] -= g2[0];
      gg[1] -= g2[1];
      gg[2] -= g2[2];
    }
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

// Unrestricted functional on 128-point blocks: spin densities and gradients in,
// w v_s and w g_s (the Fock kernel's inputs) per spin and the block energies out.
// The block is skipped when both spin densities are negligible (sum |rho_s| <
// B 1e-12 for each spin), as the reference's for_spin test does.
__global__ __launch_bounds__(128) void functional_pol_kernel(int func, XcParams prm, int64_t npts, int bs,
                                                             const double* __restrict__ ra,
                                                             const double* __restrict__ ga,
                                                             const double* __restrict__ rb,
                                                             const double* __restrict__ gb,
                                                             const double* __restrict__ w, double* __restrict__ vsa,
                                                             double* __restrict__ gwa, double* __restrict__ vsb,
                                                             double* __restrict__ gwb, double* __restrict__ e_blk,
                                                             const double* __restrict__ ra2,
                                                             const double* __restrict__ ga2,
                                                             const double* __restrict__ rb2,
                                                             const double* __restrict__ gb2) {
  __shared__ double red[128];
  const int b = blockIdx.x, t = threadIdx.x;
  const int64_t first = (int64_t)b * bs;
  const int B = (int)min((int64_t)bs, npts - first);
  const int64_t q = first + t;
  const bool in = t < B;
  const bool gga = functional_is_gga(func);
  const double a = in ? ra[q] : 0.0, bb = in ? rb[q] : 0.0;
  double xa[3] = {0.0, 0.0, 0.0}, xb[3] = {0.0, 0.0, 0.0};
  if (in && gga)
    for (int c = 0; c < 3; ++c) {
      xa[c] = ga[c * npts + q];
      xb[c] = gb[c * npts + q];
    }
  const bool skip_a = block_sum_128(fabs(a), red) < B * 1e-12;
  const bool skip_b = block_sum_128(fabs(bb), red) < B * 1e-12;
  double e = 0.0, v[2] = {0.0, 0.0}, g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (in && !(skip_a && skip_b)) e = xc_point_pol(func, a, bb, xa, xb, prm, v, g);
  if (ra2) {
    // non-additive form: the same functional of a second pair of spin densities is subtracted
    // (its own block skip applies to it)
    const double a2 = in ? ra2[q] : 0.0, b2 = in ? rb2[q] : 0.0;
    double ya[3] = {0.0, 0.0, 0.0}, yb[3] = {0.0, 0.0, 0.0};
    if (in && gga)
      for (int c = 0; c < 3; ++c) {
        ya[c] = ga2[c * npts + q];
        yb[c] = gb2[c * npts + q];
      }
    const bool skip2_a = block_sum_128(fabs(a2), red) < B * 1e-12;
    const bool skip2_b = block_sum_128(fabs(b2), red) < B * 1e-12;
    if (in && !(skip2_a && skip2_b)) {
      double v2[2] = {0.0, 0.0}, g2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      e -= xc_point_pol(func, a2, b2, ya, yb, prm, v2, g2);
      v[0] -= v2[0];
      v[1] -= v2[1];
      for (int c = 0; c < 6; ++c) g[c] -= g2[c];
    }
  }
  const double wq = in ? w[q] : 0.0;
  if (in && vsa) {
    vsa[q] = wq * v[0];
    vsb[q] = wq * v[1];
    if (gga)
      for (int c = 0; c < 3; ++c) {
        gwa[c * npts + q] = wq * g[c];
        gwb[c * npts + q] = wq * g[3 + c];
      }
  }
  const double eb = block_sum_128(wq * e, red);
  if (t == 0) e_blk[b] = eb;
}

// Sub-tile Fock contraction.  One workgroup of kFockWaves wavefronts per work item; the
// item's 16 x 16 AO sub-tiles (i, j) are dealt to the wavefronts round-robin,
// S per wavefront, each accumulating S_ij = phi_i^T G_j + G_i^T phi_j over the
// block's 128 points in registers (dummy slots of a short item compute on row
// 0 and are dropped).  The points stream in chunks of KC (32 FP32 / 16 FP64,
// one 128-byte line of each AO row): the global loads of a chunk (along the
// point axis of the AO-major tiles) are issued a chunk or two before its MFMAs,
// and a chunk is turned into phi and G = (w v / 2) phi + sum_k (w g_k) d_k phi
// rows of LDS on its way in.  Inside a chunk, k-step t of lane group g takes
// point KV g + t (KV = KC / 4), so each MFMA operand of a chunk is one 32-byte
// LDS read of KV consecutive points of one AO row.
template <typename T, int S>
__global__ __launch_bounds__(64 * kFockWaves) void fock_sub_kernel(const FockItem* __restrict__ work,
                                                       const DevBlock* __restrict__ blocks,
                                                       const int* __restrict__ ao_idx, const T* __restrict__ tiles,
                                                       const double* __restrict__ vs, const double* __restrict__ gw,
                                                       int gga, double thr, T* __restrict__ Tscr, int64_t npts,
                                                       const PackSys* __restrict__ pack) {
  using M = Mfma<T>;
  constexpr int KC = 128 / (int)sizeof(T);       // points per chunk: one 128-byte line of each AO row
  constexpr int KV = KC / 4;
  constexpr int RS = KC + 16 / (int)sizeof(T);  // 144-byte LDS rows: the 16 rows of an operand read hit distinct banks
  constexpr int NCH = kTile / KC;
  constexpr int NT = 64 * kFockWaves;            // threads
  constexpr int NTASK = kFockRows * 4 / NT;
  typedef T vec_t __attribute__((ext_vector_type(KV)));
  // Two LDS buffers, one barrier per chunk.  FP32 stages chunk c + 1 after the MFMAs of chunk c
  // (4.25 -> 4.09 ms at (H2O)64/def2-TZVP against one buffer); FP64 stages it in the instruction
  // stream of the MFMAs of slot 0 of chunk c, from registers loaded one chunk earlier (6.22 -> 5.88
  // ms; the same order measured no faster in FP32).
  constexpr int NBUF = 2;
  __shared__ __attribute__((aligned(16))) T sPhi2[NBUF][kFockRows][RS];
  __shared__ __attribute__((aligned(16))) T sG2[NBUF][kFockRows][RS];
  // w v / 2, w g_x, w g_y, w g_z of each point in the contraction type, one plane per quantity:
  // the four point runs a wavefront reads at once then start on distinct banks
  __shared__ __attribute__((aligned(16))) T spot[4][kTile];
  __shared__ double red[2];

  const FockItem it = work[blockIdx.x];
  const DevBlock bk = blocks[it.block];
  if (pack) {
    const PackSys& ps = pack[bk.sys];
    tiles = static_cast<const T*>(ps.tiles);
    vs = ps.vs;
    gw = ps.gw;
    Tscr = static_cast<T*>(ps.T);
    npts = ps.npts;
  }
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int npad = bk.npad;
  const bool diag = it.nb == 0;
  const int R = it.na + it.nb;
  const gmem<const T>* tb = to_global(tiles + bk.tile_off);
  const int64_t cstride = (int64_t)npad * kTile;
  vec_t raw[NTASK][4];
  if constexpr (sizeof(T) == 8) {  // LDA leaves the gradient components zero (the FP64 staging uses them)
#pragma unroll
    for (int m = 0; m < NTASK; ++m)
#pragma unroll
      for (int k = 1; k < 4; ++k) raw[m][k] = vec_t(T(0));
  }
  auto grow_of = [&](int lr) { return lr < it.na ? it.a0 + lr : it.b0 + (lr - it.na); };
  // staging task -> (row, quarter of the chunk): 16 consecutive lanes take 16 rows of one
  // quarter, so their 16-byte LDS stores (144-byte row stride) cover distinct banks
  const int q = (tid >> 4) & 3;
  auto row_of = [&](int task) { return (task & 15) + 16 * (task >> 6); };
  auto load = [&](int c) {
    const int p0 = c * KC;
#pragma unroll
    for (int m = 0; m < NTASK; ++m) {
      const int task = tid + NT * m;
      if (task < 4 * R) {
        const gmem<const T>* src = tb + (int64_t)grow_of(row_of(task)) * kTile + p0 + KV * q;
        raw[m][0] = *reinterpret_cast<const gmem<const vec_t>*>(src);
        if (gga) {
          raw[m][1] = *reinterpret_cast<const gmem<const vec_t>*>(src + cstride);
          raw[m][2] = *reinterpret_cast<const gmem<const vec_t>*>(src + 2 * cstride);
          raw[m][3] = *reinterpret_cast<const gmem<const vec_t>*>(src + 3 * cstride);
        }
      }
    }
  };
  load(0);  // in flight while the potentials are staged and the block skip is decided
  if (tid < kTile) {
    const bool in = tid < bk.npts;
    const int64_t q = bk.first + tid;
    const gmem<const double>* vsg = to_global(vs);
    const gmem<const double>* gwg = to_global(gw);
    const double v0 = in ? vsg[q] : 0.0;
    const double v1 = (in && gga) ? gwg[q] : 0.0;
    const double v2 = (in && gga) ? gwg[npts + q] : 0.0;
    const double v3 = (in && gga) ? gwg[2 * npts + q] : 0.0;
    spot[0][tid] = (T)(0.5 * v0);
    spot[1][tid] = (T)v1;
    spot[2][tid] = (T)v2;
    spot[3][tid] = (T)v3;
    double a = fabs(v0) + fabs(v1) + fabs(v2) + fabs(v3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if (lane == 0) red[w] = a;
  }
  __syncthreads();


  // sub-tile t of the item -> (i, j) in 16-row units (i <= j on a diagonal item)
  auto decode = [&](int t, int& i, int& j) {
    i = 0;
    if (diag) {
      const int nt = it.na / 16;
      int start = 0;
      while (start + (nt - i) <= t) {
        start += nt - i;
        ++i;
      }
      j = i + (t - start);
    } else {
      const int ntb = it.nb / 16;
      i = t / ntb;
      j = t - i * ntb;
    }
  };
  // a skipped block (every item of it takes the same decision from the same values) leaves zeros
  // in its sub-tiles, so the reduce needs no per-block test
  if ((red[0] + red[1]) / bk.npts < thr) {
    gmem<T>* out0 = to_global(Tscr + bk.toff);
    for (int t = w; t < it.nsub; t += kFockWaves) {
      int i, j;
      decode(t, i, j);
      const int I = it.a0 / 16 + i, J = (diag ? it.a0 : it.b0) / 16 + j;
      gmem<T>* sub = out0 + fock_row_start(16 * I, npad) + 16 * (J - I);
      for (int e = lane; e < 256; e += 64) sub[(int64_t)(e >> 4) * (npad - 16 * I) + (e & 15)] = T(0);
    }
    return;
  }
  // this wavefront's sub-tiles as LDS row offsets (wavefront-uniform)
  int li[S], lj[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = w + kFockWaves * s;
    int i = 0, j = 0;
    if (t < it.nsub) decode(t, i, j);
    li[s] = __builtin_amdgcn_readfirstlane(t < it.nsub ? 16 * i : 0);
    lj[s] = __builtin_amdgcn_readfirstlane(t < it.nsub ? (diag ? 16 * j : it.na + 16 * j) : 0);
  }

  auto store = [&](int c) {
    T(*sPhi)[RS] = sPhi2[NBUF == 2 ? (c & 1) : 0];
    T(*sG)[RS] = sG2[NBUF == 2 ? (c & 1) : 0];
    // the KV points of this thread's tasks are the same for every task; their potentials are
    // read as whole vectors, the GGA test taken once (a test per value branches around each
    // LDS read and waits for it)
    const int p0 = c * KC + KV * q;
    vec_t gv[NTASK];
    auto pot = [&](int k) { return *reinterpret_cast<const vec_t*>(&spot[k][p0]); };
    const vec_t w0 = pot(0);
    if (gga) {
      const vec_t w1 = pot(1), w2 = pot(2), w3 = pot(3);
#pragma unroll
      for (int m = 0; m < NTASK; ++m)
#pragma unroll
        for (int e = 0; e < KV; ++e)
          gv[m][e] = w0[e] * raw[m][0][e] + w1[e] * raw[m][1][e] + w2[e] * raw[m][2][e] + w3[e] * raw[m][3][e];
    } else {
#pragma unroll
      for (int m = 0; m < NTASK; ++m)
#pragma unroll
        for (int e = 0; e < KV; ++e) gv[m][e] = w0[e] * raw[m][0][e];
    }
#pragma unroll
    for (int m = 0; m < NTASK; ++m) {
      const int task = tid + NT * m;
      if (task < 4 * R) {
        const int lr = row_of(task);
        *reinterpret_cast<vec_t*>(&sPhi[lr][KV * q]) = raw[m][0];
        *reinterpret_cast<vec_t*>(&sG[lr][KV * q]) = gv[m];
      }
    }
  };

  typename M::acc_t acc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s] = M::zero();
  const int r = lane & 15, g = lane >> 4;
  // slots of this wavefront that hold a sub-tile (an item smaller than its class leaves the rest idle)
  const int slive = it.nsub > w ? (it.nsub - w + kFockWaves - 1) / kFockWaves : 0;
  auto slot = [&](int s, T(*sPhi)[RS], T(*sG)[RS]) {
    const vec_t aP = *reinterpret_cast<const vec_t*>(&sPhi[li[s] + r][KV * g]);
    const vec_t aG = *reinterpret_cast<const vec_t*>(&sG[li[s] + r][KV * g]);
    const vec_t bP = *reinterpret_cast<const vec_t*>(&sPhi[lj[s] + r][KV * g]);
    const vec_t bG = *reinterpret_cast<const vec_t*>(&sG[lj[s] + r][KV * g]);
#pragma unroll
    for (int t = 0; t < KV; ++t) acc[s] = M::step(aP[t], bG[t], acc[s]);
#pragma unroll
    for (int t = 0; t < KV; ++t) acc[s] = M::step(aG[t], bP[t], acc[s]);
  };
  if constexpr (sizeof(T) == 8) {
    static_assert((NCH & (NCH - 1)) == 0, "chunks per item: a power of two");
    // Staging without branches, so that the compiler can spread it between the MFMAs of slot 0:
    // G takes all four terms (LDA's gradient potentials and components are zero), tasks beyond
    // the item's rows write rows that no MFMA reads, and chunk NCH re-stages chunk 0 into the
    // buffer no later chunk reads.
    auto store_flat = [&](int c) {
      T(*sPhi)[RS] = sPhi2[c & 1];
      T(*sG)[RS] = sG2[c & 1];
      const int p0 = (c & (NCH - 1)) * KC + KV * q;
      const vec_t w0 = *reinterpret_cast<const vec_t*>(&spot[0][p0]);
      const vec_t w1 = *reinterpret_cast<const vec_t*>(&spot[1][p0]);
      const vec_t w2 = *reinterpret_cast<const vec_t*>(&spot[2][p0]);
      const vec_t w3 = *reinterpret_cast<const vec_t*>(&spot[3][p0]);
#pragma unroll
      for (int m = 0; m < NTASK; ++m) {
        vec_t gv;
#pragma unroll
        for (int e = 0; e < KV; ++e)
          gv[e] = w0[e] * raw[m][0][e] + w1[e] * raw[m][1][e] + w2[e] * raw[m][2][e] + w3[e] * raw[m][3][e];
        const int lr = row_of(tid + NT * m);
        *reinterpret_cast<vec_t*>(&sPhi[lr][KV * q]) = raw[m][0];
        *reinterpret_cast<vec_t*>(&sG[lr][KV * q]) = gv;
      }
    };
    store_flat(0);
    if (NCH > 1) load(1);
    __syncthreads();
    for (int c = 0; c < NCH; ++c) {
      T(*sPhi)[RS] = sPhi2[c & 1];
      T(*sG)[RS] = sG2[c & 1];
      slot(0, sPhi, sG);  // every wavefront (an idle one computes on row 0 and drops the result)
      store_flat(c + 1);
      if (c + 2 < NCH) load(c + 2);
#pragma unroll
      for (int s = 1; s < S; ++s) {
        if (s >= slive) break;  // wavefront-uniform
        slot(s, sPhi, sG);
        // keep one slot's operands live at a time (bounds the register file)
        __builtin_amdgcn_sched_barrier(0);
      }
      __syncthreads();
    }
  } else {
    store(0);
    __syncthreads();
    for (int c = 0; c < NCH; ++c) {
      T(*sPhi)[RS] = sPhi2[c & 1];
      T(*sG)[RS] = sG2[c & 1];
      if (c + 1 < NCH) load(c + 1);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (s >= slive) break;  // wavefront-uniform
        slot(s, sPhi, sG);
        // keep one slot's operands live at a time (bounds the register file)
        __builtin_amdgcn_sched_barrier(0);
      }
      if (c + 1 < NCH) store(c + 1);
      __syncthreads();
    }
  }

  // store the sub-tiles (block sub-tile (I, J), I <= J; 16 lanes write 16 consecutive values of a row)
  gmem<T>* tb_out = to_global(Tscr + bk.toff);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = w + kFockWaves * s;
    if (t >= it.nsub) continue;
    int i, j;
    decode(t, i, j);
    const int I = it.a0 / 16 + i, J = (diag ? it.a0 : it.b0) / 16 + j;
    gmem<T>* sub = tb_out + fock_row_start(16 * I, npad) + 16 * (J - I) + r;
    const int stride = npad - 16 * I;
#pragma unroll
    for (int e = 0; e < 4; ++e) sub[(int64_t)M::row(lane, e) * stride] = acc[s][e];
  }
}

// One AO row i of the Fock accumulator per workgroup: its entries (the blocks holding AO i, in
// block order) are dealt to the kReduceWaves wavefronts round-robin; a wavefront adds each entry's
// stored row segment into the workgroup's LDS copy of the row, and the row is written once to
// U[j + n i], j >= i.  Each lane takes an aligned group of four columns (256 per pass of the
// wavefront): one 16-byte load of their AO indices and one of their stored values, the lanes of a
// wavefront reading one contiguous run of the block's stored row (a block's AO list is padded to
// a multiple of 32 with -1 and its stored row a runs from column 16 (a / 16) to the padded end,
// so an aligned group never leaves either; the columns of the group before the segment start are
// masked).  The loads of a wavefront's next entry are issued before the current entry's LDS
// additions.  Rows longer than the LDS window are done in several passes over the entries.
// Fixed = false: FP64 additions in LDS, which the wavefronts make in an order that varies from run
// to run, so F varies at the rounding level.  Fixed = true: each term x is split into integers,
// x = h 2^-32 + l 2^-72 + e with h = x 2^32 and l = (x - h 2^-32) 2^72 rounded to the nearest
// integer, |l| <= 2^39 and |e| <= 2^-73, and h and l are added into two 64-bit integer sums per
// column; integer sums do not depend on the order, so F is the same bit for bit on every run (and
// the sum is exact up to the e terms).  The rounding adds 1.5 2^20 (1.5 2^52) in FP64, whose unit
// in the last place is then 2^-32 (1), and reads the integer from the bit pattern.  Valid for
// terms of |x| < 2^19 in columns of fewer than 2^23 terms whose absolute values sum to less than
// 2^30.
template <typename T, bool Fixed>
__global__ __launch_bounds__(64 * kReduceWaves) void fock_reduce_kernel(const RowEnt* __restrict__ ent,
                                                                        const int* __restrict__ row_ptr,
                                                                        const T* __restrict__ Tscr,
                                                                        const int* __restrict__ ao_idx,
                                                                        double* __restrict__ U, int n, int window,
                                                                        const PackSys* __restrict__ pack) {
  extern __shared__ unsigned long long racc_u[];  // window (FP64 sums) or 2 window (Fixed: h, then l sums)
  double* racc = reinterpret_cast<double*>(racc_u);
  if (pack) {
    const PackSys& ps = pack[blockIdx.y];
    ent = ps.ent;
    row_ptr = ps.row_ptr;
    Tscr = static_cast<const T*>(ps.T);
    ao_idx = ps.ao_idx;
    U = ps.U;
    n = ps.n;
  }
  const int i = blockIdx.x;
  if (i >= n) return;  // uniform over the workgroup
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const gmem<const RowEnt>* entg = to_global(ent);
  const gmem<const T>* Tg = to_global(Tscr);
  const gmem<const int>* idxg = to_global(ao_idx);
  const int e0 = to_global(row_ptr)[i], e1 = to_global(row_ptr)[i + 1];
  for (int w0 = 0; w0 < n - i; w0 += window) {
    const int wl = min(window, n - i - w0);
    for (int k = tid; k < wl; k += 64 * kReduceWaves) {
      if constexpr (Fixed) {
        racc_u[k] = 0;
        racc_u[window + k] = 0;
      } else {
        racc[k] = 0.0;
      }
    }
    __syncthreads();
    auto fetch = [&](const RowEnt& E, int c, i32x4& gi, double (&v)[4]) {
      gi = *reinterpret_cast<const gmem<const i32x4>*>(idxg + E.idx_off + c);
      const gmem<const T>* src = Tg + E.rowbase + c;
      if constexpr (sizeof(T) == 4) {
        const f32x4 f = *reinterpret_cast<const gmem<const f32x4>*>(src);
        v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
      } else {
        const f64x2 f0 = *reinterpret_cast<const gmem<const f64x2>*>(src);
        const f64x2 f1 = *reinterpret_cast<const gmem<const f64x2>*>(src + 2);
        v[0] = f0.x; v[1] = f0.y; v[2] = f1.x; v[3] = f1.y;
      }
    };
    auto add = [&](const RowEnt& E, int c, const i32x4& gi, const double (&v)[4]) {
      const int gg[4] = {gi.x, gi.y, gi.z, gi.w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int g = gg[u] - i - w0;
        if (c + u >= E.c0 && c + u < E.c1 && g >= 0 && g < wl) {
          if constexpr (Fixed) {
            const double t = v[u] + 0x1.8p20;
            const long long h = __double_as_longlong(t) - __double_as_longlong(0x1.8p20);
            const double rem = v[u] - (t - 0x1.8p20);  // exact, |rem| <= 2^-33
            const long long l = __double_as_longlong(rem * 0x1p72 + 0x1.8p52) - __double_as_longlong(0x1.8p52);
            atomicAdd(racc_u + g, (unsigned long long)h);
            atomicAdd(racc_u + window + g, (unsigned long long)l);
          } else {
            atomicAdd(racc + g, v[u]);
          }
        }
      }
    };
    // this wavefront's entries e0 + w + kReduceWaves k, k < nk: their RowEnts are read 64 at a time,
    // one per lane, and broadcast from the lanes (scalar registers), so that an entry's loads
    // depend on no load of its own; the loads of the next kAhead entries are in flight during an
    // entry's LDS additions
    constexpr int kAhead = 1;
    const int nk = e1 - e0 > w ? (e1 - e0 - w + kReduceWaves - 1) / kReduceWaves : 0;
    for (int kb = 0; kb < nk; kb += 64) {
      const int cnt = min(64, nk - kb);
      RowEnt L = {0, 0, 0, 0, 0};
      if (lane < cnt) L = load_global(entg + e0 + w + kReduceWaves * (kb + lane));
      RowEnt Eq[kAhead + 1];
      int cq[kAhead + 1];
      i32x4 gq[kAhead + 1];
      double vq[kAhead + 1][4];
      auto issue = [&](int k, int d) {  // entry k into slot d
        RowEnt& E = Eq[d];
        const int lo = __builtin_amdgcn_readlane((int)(L.rowbase & 0xffffffff), k);
        const int hi = __builtin_amdgcn_readlane((int)(L.rowbase >> 32), k);
        E.rowbase = (int64_t)(((uint64_t)(unsigned)hi << 32) | (unsigned)lo);
        E.c0 = __builtin_amdgcn_readlane(L.c0, k);
        E.c1 = __builtin_amdgcn_readlane(L.c1, k);
        E.idx_off = __builtin_amdgcn_readlane(L.idx_off, k);
        E.block = 0;
        cq[d] = (E.c0 & ~3) + 4 * lane;
        gq[d] = i32x4{-1, -1, -1, -1};
#pragma unroll
        for (int u = 0; u < 4; ++u) vq[d][u] = 0.0;
        if (cq[d] < E.c1) fetch(E, cq[d], gq[d], vq[d]);
      };
#pragma unroll
      for (int d = 0; d < kAhead; ++d)
        if (d < cnt) issue(d, d);
      for (int k = 0; k < cnt; ++k) {
        if (k + kAhead < cnt) issue(k + kAhead, kAhead);  // wavefront-uniform
        const RowEnt& E = Eq[0];
        if (cq[0] < E.c1) add(E, cq[0], gq[0], vq[0]);
        for (int c2 = cq[0] + 256; c2 < E.c1; c2 += 256) {
          i32x4 g2;
          double v2[4];
          fetch(E, c2, g2, v2);
          add(E, c2, g2, v2);
        }
#pragma unroll
        for (int d = 0; d < kAhead; ++d) {
          Eq[d] = Eq[d + 1];
          cq[d] = cq[d + 1];
          gq[d] = gq[d + 1];
#pragma unroll
          for (int u = 0; u < 4; ++u) vq[d][u] = vq[d + 1][u];
        }
      }
    }
    __syncthreads();
    for (int j = tid; j < wl; j += 64 * kReduceWaves) {
      double f;
      if constexpr (Fixed)
        f = (double)(long long)racc_u[j] * 0x1p-32 + (double)(long long)racc_u[window + j] * 0x1p-72;
      else
        f = racc[j];
      to_global(U)[(int64_t)(i + w0 + j) + (int64_t)n * i] = f;
    }
    __syncthreads();
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

// rho[q] (+)= sum_i list.rho[i][q] in list order (likewise the gradient).
__global__ void sum_densities_kernel(DensityList list, int64_t npts, int want_grad, int accumulate,
                                     double* __restrict__ rho, double* __restrict__ grad) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  double r = accumulate ? rho[q] : 0.0;
  for (int i = 0; i < list.n; ++i) r += list.rho[i][q];
  rho[q] = r;
  if (want_grad) {
    for (int c = 0; c < 3; ++c) {
      double g = accumulate ? grad[c * npts + q] : 0.0;
      for (int i = 0; i < list.n; ++i) g += list.grad[i][c * npts + q];
      grad[c * npts + q] = g;
    }
  }
}

// out[0] -= list.rho[0][0], then list.rho[1][0], ... (one thread, fixed order).
__global__ void subtract_scalars_kernel(DensityList list, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double e = out[0];
  for (int i = 0; i < list.n; ++i) e -= list.rho[i][0];
  out[0] = e;
}

__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void fill_kernel(double* __restrict__ p, int64_t n, double value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ void to_float_kernel(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

// F = mirror of the lower-triangle accumulator U (or F += it).  With P32 given, each block
// also leaves sum_e F[e] (P[e] - P32[e]) in part[blockIdx.x]: the first-order change of
// E_xc from the FP32-rounded density matrix P32 back to P, since dE_xc/dP = F.
__global__ __launch_bounds__(256) void finalize_kernel(const double* __restrict__ U, double* __restrict__ F, int n,
                                                       int accumulate, const double* __restrict__ P,
                                                       const float* __restrict__ P32, double* __restrict__ part) {
  const int t = threadIdx.x;
  const int64_t e = (int64_t)blockIdx.x * 256 + t;
  double c = 0.0;
  if (e < (int64_t)n * n) {
    const int i = (int)(e % n), j = (int)(e / n);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const double u = U[hi + (int64_t)n * lo];
    const double f = accumulate ? F[e] + u : u;
    F[e] = f;
    if (P32) c = f * (P[e] - (double)P32[e]);
  }
  if (!P32) return;  // uniform over the launch
  __shared__ double red[256];
  red[t] = c;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

// sum of x[0..n) and x2[0..n2); mode 0: out = sum; 1: out -= sum
__global__ __launch_bounds__(256) void sum_kernel(const double* __restrict__ x, int64_t n,
                                                  const double* __restrict__ x2, int64_t n2, double* __restrict__ out,
                                                  int mode) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int64_t i = t; i < n; i += 256) s += x[i];
  for (int64_t i = t; i < n2; i += 256) s += x2[i];
  red[t] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) out[0] = mode == 0 ? red[0] : out[0] - red[0];
}

// ---- packed builds: elementwise stages over the systems of a pack (blockIdx.y = system) ----
__global__ void pack_to_float_kernel(const PackSys* __restrict__ pack) {
  const PackSys& ps = pack[blockIdx.y];
  const int64_t nn = (int64_t)ps.n * ps.n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x)
    ps.P32[i] = (float)ps.P64[i];
}

// F = mirror of U; FP32 builds also leave one partial of sum F (P - P32) per workgroup
__global__ __launch_bounds__(256) void pack_finalize_kernel(const PackSys* __restrict__ pack, int fp32) {
  const PackSys& ps = pack[blockIdx.y];
  const int n = ps.n, t = threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  double c = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + t; e < nn; e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const double u = ps.U[hi + (int64_t)n * lo];
    ps.F[e] = u;
    if (fp32) c += u * (ps.P64[e] - (double)ps.P32[e]);
  }
  if (!fp32) return;  // uniform over the launch
  __shared__ double red[256];
  red[t] = c;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) ps.round[blockIdx.x] = red[0];
}

// E_xc of each system: its block energies (+ the FP32 rounding term), in a fixed order
__global__ __launch_bounds__(256) void pack_sum_kernel(const PackSys* __restrict__ pack, int fp32) {
  const PackSys& ps = pack[blockIdx.x];
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int i = t; i < ps.nxc; i += 256) s += ps.eblk[i];
  if (fp32)
    for (int i = t; i < kPackRoundBlocks; i += 256) s += ps.round[i];
  red[t] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) ps.E[0] = red[0];
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
void launch_density_sub(hipStream_t s, int S, const WorkItem* work, int nwork, const DevBlock* blocks,
                        const int* ao_idx, const T* tiles, const T* P, int n, int want_grad, double* part,
                        int64_t npts, const PackSys* pack, const int* nwork_dev, const unsigned char* skip) {
  if (nwork <= 0) return;
  switch (S) {
#define SXC_DENS_S(SV)                                                                                        \
  case SV:                                                                                                    \
    hipLaunchKernelGGL((density_sub_kernel<T, SV>), dim3(nwork), dim3(512), 0, s, work, blocks, ao_idx, tiles, \
                       P, n, want_grad, part, npts, pack, nwork_dev, skip);                                   \
    return;
    static_assert(kDensMaxS == 6, "one case per column-tile class");
    SXC_DENS_S(1)
    SXC_DENS_S(2)
    SXC_DENS_S(3)
    SXC_DENS_S(4)
    SXC_DENS_S(5)
    SXC_DENS_S(6)
#undef SXC_DENS_S
  }
  throw std::runtime_error("launch_density_sub: column-tile count out of range");
}
template void launch_density_sub<double>(hipStream_t, int, const WorkItem*, int, const DevBlock*, const int*,
                                         const double*, const double*, int, int, double*, int64_t,
                                         const PackSys*, const int*, const unsigned char*);
template void launch_density_sub<float>(hipStream_t, int, const WorkItem*, int, const DevBlock*, const int*,
                                        const float*, const float*, int, int, double*, int64_t, const PackSys*,
                                        const int*, const unsigned char*);

// Largest density of block b (128 threads), from the partials of its passes.
__device__ double block_max_density(const DevBlock& bk, int64_t npts, const double* __restrict__ part, double* red) {
  const int t = threadIdx.x;
  double r = 0.0;
  if (t < bk.npts)
    for (int s = 0; s < bk.npass; ++s) r += part[(int64_t)s * 4 * npts + bk.first + t];
  red[t] = r;
  __syncthreads();
  for (int k = 64; k > 0; k >>= 1) {
    if (t < k) red[t] = fmax(red[t], red[t + k]);
    __syncthreads();
  }
  return red[0];
}

// Blocks whose largest density (from the partials of a first pass) exceeds rho_thr have their
// density work items appended to the list of their column-tile class, for a second pass.  With a
// mask, only the blocks outside it are tested (the first pass skipped the others) and the listed
// ones join it.
__global__ __launch_bounds__(128) void mark_dense_kernel(const DevBlock* __restrict__ blocks, int64_t npts,
                                                         const double* __restrict__ part, double rho_thr,
                                                         const WorkItem* __restrict__ by_block,
                                                         const int* __restrict__ item0, const int* __restrict__ cls,
                                                         DenseLists lists, unsigned char* __restrict__ mask) {
  __shared__ double red[128];
  const int b = blockIdx.x, t = threadIdx.x;
  if (mask && mask[b]) return;  // uniform over the workgroup
  const DevBlock bk = blocks[b];
  const bool dense = bk.npass > 0 && block_max_density(bk, npts, part, red) > rho_thr;
  if (t == 0 && dense) {
    const int c = cls[b];
    const int pos = atomicAdd(lists.count + c, bk.npass);
    for (int k = 0; k < bk.npass; ++k) lists.item[c][pos + k] = by_block[item0[b] + k];
    if (mask) mask[b] = 1;
  }
}

// The density work items of the blocks in the mask, appended to the list of their class.
__global__ __launch_bounds__(64) void list_masked_kernel(const DevBlock* __restrict__ blocks, int nblocks,
                                                         const WorkItem* __restrict__ by_block,
                                                         const int* __restrict__ item0, const int* __restrict__ cls,
                                                         DenseLists lists, const unsigned char* __restrict__ mask) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= nblocks || !mask[b]) return;
  const int np = blocks[b].npass;
  if (np <= 0) return;
  const int c = cls[b];
  const int pos = atomicAdd(lists.count + c, np);
  for (int k = 0; k < np; ++k) lists.item[c][pos + k] = by_block[item0[b] + k];
}

// After the second pass: a listed block stays in the mask while its density exceeds rho_thr.
__global__ __launch_bounds__(128) void refresh_dense_kernel(const DevBlock* __restrict__ blocks, int64_t npts,
                                                            const double* __restrict__ part, double rho_thr,
                                                            unsigned char* __restrict__ mask) {
  __shared__ double red[128];
  const int b = blockIdx.x;
  if (!mask[b]) return;  // uniform over the workgroup
  const DevBlock bk = blocks[b];
  const bool dense = block_max_density(bk, npts, part, red) > rho_thr;
  if (threadIdx.x == 0) mask[b] = dense ? 1 : 0;
}

__global__ void clear_counts_kernel(int* __restrict__ c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

__global__ void to_double_kernel(const float* __restrict__ s, double* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (double)s[i];
}

void launch_mark_dense(hipStream_t s, const DevBlock* blocks, int nblocks, int64_t npts, const double* part,
                       double rho_thr, const WorkItem* by_block, const int* item0, const int* cls,
                       const DenseLists& lists, unsigned char* mask) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(clear_counts_kernel, dim3(1), dim3(64), 0, s, lists.count, kDensMaxS + 1);
  hipLaunchKernelGGL(mark_dense_kernel, dim3(nblocks), dim3(128), 0, s, blocks, npts, part, rho_thr, by_block, item0,
                     cls, lists, mask);
}

void launch_list_masked(hipStream_t s, const DevBlock* blocks, int nblocks, const WorkItem* by_block,
                        const int* item0, const int* cls, const DenseLists& lists, const unsigned char* mask) {
  hipLaunchKernelGGL(clear_counts_kernel, dim3(1), dim3(64), 0, s, lists.count, kDensMaxS + 1);
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(list_masked_kernel, dim3((nblocks + 63) / 64), dim3(64), 0, s, blocks, nblocks, by_block, item0,
                     cls, lists, mask);
}

void launch_refresh_dense(hipStream_t s, const DevBlock* blocks, int nblocks, int64_t npts, const double* part,
                          double rho_thr, unsigned char* mask) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(refresh_dense_kernel, dim3(nblocks), dim3(128), 0, s, blocks, npts, part, rho_thr, mask);
}

void launch_to_double(hipStream_t s, const float* src, double* dst, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(to_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

void launch_reduce_partials(hipStream_t s, const DevBlock* blocks, int bs, int64_t npts, const double* part,
                            int want_grad, double* rho, double* grad, int accumulate) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(grid1d(npts, 256)), dim3(256), 0, s, blocks, bs, npts, part,
                     want_grad, rho, grad, accumulate);
}

void launch_functional(hipStream_t s, int func, const XcParams& prm, int64_t npts, int bs, const double* rho,
                       const double* grad, const DevBlock* blocks, const double* part, const double* w, double* eps,
                       double* v, double* g, double* vs, double* gw, double* e_blk, const double* rho_sub,
                       const double* grad_sub) {
  const int nb = (int)((npts + bs - 1) / bs);
  if (nb <= 0) return;
  hipLaunchKernelGGL(functional_kernel, dim3(nb), dim3(128), 0, s, func, prm, npts, bs, rho, grad, blocks, part, w,
                     eps, v, g, vs, gw, e_blk, rho_sub, grad_sub, (const PackSys*)nullptr);
}

void launch_functional_packed(hipStream_t s, int func, const XcParams& prm, int nblocks, const DevBlock* blocks,
                              const PackSys* pack) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(functional_kernel, dim3(nblocks), dim3(128), 0, s, func, prm, (int64_t)0, kTile,
                     (const double*)nullptr, (const double*)nullptr, blocks, (const double*)nullptr,
                     (const double*)nullptr, (double*)nullptr, (double*)nullptr, (double*)nullptr, (double*)nullptr,
                     (double*)nullptr, (double*)nullptr, (const double*)nullptr, (const double*)nullptr, pack);
}

static unsigned pack_grid(int64_t maxnn) {
  const int64_t g = (maxnn + 255) / 256;
  return (unsigned)(g < 1 ? 1 : g > 1024 ? 1024 : g);
}

void launch_pack_to_float(hipStream_t s, const PackSys* pack, int nsys, int64_t maxnn) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(pack_to_float_kernel, dim3(pack_grid(maxnn), nsys), dim3(256), 0, s, pack);
}

void launch_pack_finalize(hipStream_t s, const PackSys* pack, int nsys, int fp32) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(pack_finalize_kernel, dim3(kPackRoundBlocks, nsys), dim3(256), 0, s, pack, fp32);
}

void launch_pack_sum(hipStream_t s, const PackSys* pack, int nsys, int fp32) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(pack_sum_kernel, dim3(nsys), dim3(256), 0, s, pack, fp32);
}

void launch_functional_pol(hipStream_t s, int func, const XcParams& prm, int64_t npts, int bs, const double* ra,
                           const double* ga, const double* rb, const double* gb, const double* w, double* vsa,
                           double* gwa, double* vsb, double* gwb, double* e_blk, const double* ra2,
                           const double* ga2, const double* rb2, const double* gb2) {
  const int nb = (int)((npts + bs - 1) / bs);
  if (nb <= 0) return;
  hipLaunchKernelGGL(functional_pol_kernel, dim3(nb), dim3(128), 0, s, func, prm, npts, bs, ra, ga, rb, gb, w, vsa,
                     gwa, vsb, gwb, e_blk, ra2, ga2, rb2, gb2);
}

template <typename T, int S>
static void launch_fock_sub_s(hipStream_t s, const FockItem* work, int nwork, const DevBlock* blocks,
                              const int* ao_idx, const T* tiles, const double* vs, const double* gw, int gga,
                              double thr, T* Tscr, int64_t npts, const PackSys* pack) {
  hipLaunchKernelGGL((fock_sub_kernel<T, S>), dim3(nwork), dim3(64 * kFockWaves), 0, s, work, blocks, ao_idx, tiles,
                     vs, gw, gga, thr, Tscr, npts, pack);
}

template <typename T>
void launch_fock_sub(hipStream_t s, int sc, const FockItem* work, int nwork, const DevBlock* blocks,
                     const int* ao_idx, const T* tiles, const double* vs, const double* gw, int gga, double thr,
                     T* Tscr, int64_t npts, const PackSys* pack) {
  if (nwork <= 0) return;
  switch (kFockS[sc]) {
#define SXC_FOCK_S(SV) \
  case SV:             \
    return launch_fock_sub_s<T, SV>(s, work, nwork, blocks, ao_idx, tiles, vs, gw, gga, thr, Tscr, npts, pack);
    SXC_FOCK_S(1)
    SXC_FOCK_S(2)
    SXC_FOCK_S(3)
    SXC_FOCK_S(4)
    SXC_FOCK_S(5)
#if SXC_FOCK_WAVES != 16
    SXC_FOCK_S(6)
    SXC_FOCK_S(7)
    SXC_FOCK_S(8)
    SXC_FOCK_S(10)
#endif
#undef SXC_FOCK_S
  }
  throw std::runtime_error("launch_fock_sub: slot class beyond the accumulator budget of this precision");
}
template void launch_fock_sub<double>(hipStream_t, int, const FockItem*, int, const DevBlock*, const int*,
                                      const double*, const double*, const double*, int, double, double*, int64_t,
                                      const PackSys*);
template void launch_fock_sub<float>(hipStream_t, int, const FockItem*, int, const DevBlock*, const int*,
                                     const float*, const double*, const double*, int, double, float*, int64_t,
                                     const PackSys*);

// Row columns per LDS pass of the reduce: kReduceWindow, or less where the device's LDS per
// workgroup cannot hold that many columns of 8 bytes (fixed = false) or 16 bytes (fixed = true).
static int reduce_window_limit(bool fixed) {
  static int cached[2][64] = {};
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= 64) dev = 0;
  if (!cached[fixed][dev]) {
    int lds = 65536;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
    const int w = lds / (fixed ? 16 : 8);
    cached[fixed][dev] = (w < kReduceWindow ? w : kReduceWindow) / 64 * 64;
  }
  return cached[fixed][dev];
}

template <typename T>
void launch_fock_reduce(hipStream_t s, const RowEnt* ent, const int* row_ptr, const T* Tscr, const int* ao_idx,
                        double* U, int n, const PackSys* pack, int nsys, bool fixed) {
  if (n <= 0 || nsys <= 0) return;
  const int lim = reduce_window_limit(fixed);
  const int window = n < lim ? n : lim;
  const size_t lds = (fixed ? 16 : 8) * (size_t)window;
  if (fixed)
    hipLaunchKernelGGL((fock_reduce_kernel<T, true>), dim3(n, nsys), dim3(64 * kReduceWaves), lds, s, ent, row_ptr,
                       Tscr, ao_idx, U, n, window, pack);
  else
    hipLaunchKernelGGL((fock_reduce_kernel<T, false>), dim3(n, nsys), dim3(64 * kReduceWaves), lds, s, ent, row_ptr,
                       Tscr, ao_idx, U, n, window, pack);
}
template void launch_fock_reduce<double>(hipStream_t, const RowEnt*, const int*, const double*, const int*, double*,
                                         int, const PackSys*, int, bool);
template void launch_fock_reduce<float>(hipStream_t, const RowEnt*, const int*, const float*, const int*, double*,
                                        int, const PackSys*, int, bool);

void launch_weight_potential(hipStream_t s, int64_t npts, const double* w, const double* v, const double* g,
                             double* vs, double* gw) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(weight_potential_kernel, dim3(grid1d(npts, 256)), dim3(256), 0, s, npts, w, v, g, vs, gw);
}

void launch_sum_densities(hipStream_t s, const DensityList& list, int64_t npts, int want_grad, int accumulate,
                          double* rho, double* grad) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(sum_densities_kernel, dim3(grid1d(npts, 256)), dim3(256), 0, s, list, npts, want_grad,
                     accumulate, rho, grad);
}

void launch_subtract_scalars(hipStream_t s, const DensityList& list, double* out) {
  hipLaunchKernelGGL(subtract_scalars_kernel, dim3(1), dim3(64), 0, s, list, out);
}

void launch_copy(hipStream_t s, const void* src, void* dst, int64_t n32) {
  if (n32 <= 0) return;
  hipLaunchKernelGGL(copy_kernel, dim3(grid1d(n32, 256)), dim3(256), 0, s, (const double4*)src, (double4*)dst, n32);
}

void launch_fill(hipStream_t s, double* p, int64_t n, double value) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, p, n, value);
}

void launch_to_float(hipStream_t s, const double* src, float* dst, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(to_float_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, src, dst, n);
}

void launch_finalize(hipStream_t s, const double* U, double* F, int n, int accumulate, const double* P,
                     const float* P32, double* part) {
  const int64_t nn = (int64_t)n * n;
  if (nn <= 0) return;
  hipLaunchKernelGGL(finalize_kernel, dim3(grid1d(nn, 256)), dim3(256), 0, s, U, F, n, accumulate, P, P32, part);
}

int finalize_blocks(int n) { return (int)grid1d((int64_t)n * n, 256); }

void launch_sum(hipStream_t s, const double* x, int64_t n, double* out, int subtract, const double* x2, int64_t n2) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, x, n, x2, n2, out, subtract);
}

}  // namespace sxc
