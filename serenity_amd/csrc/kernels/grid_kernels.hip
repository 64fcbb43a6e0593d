// Molecular grid on the GPU (GridFactory.cpp:51-300, HilbertRTreeSorting.cpp:27-216): the
// candidates (each atom's reference grid translated to the atom) are generated inside the
// partition kernel, one thread per candidate, with the same per-point arithmetic as the host loop
// (host/partition.h); the weight cut keeps the candidates above the threshold in candidate order
// (flags, exclusive scan, scatter); the Hilbert order takes the bounding box of the kept points,
// one key per point (host/hilbert.h, the function the host sort uses) and a stable radix sort on
// the descending keys, so equal keys keep their order as on the host.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <stdexcept>

#include "../host/hilbert.h"
#include "../host/partition.h"
#include "xc_internal.h"

namespace sxc {
namespace {

__global__ __launch_bounds__(256) void partition_kernel(int type, unsigned smoothing, int n,
                                                        const double* __restrict__ xyz,
                                                        const double* __restrict__ dist,
                                                        const double* __restrict__ apair,
                                                        const double* __restrict__ min_dist, int64_t ncand,
                                                        const double* __restrict__ cand_xyz,
                                                        const double* __restrict__ cand_w,
                                                        const int* __restrict__ cand_atom, double* __restrict__ w) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncand) return;
  const int k = cand_atom[c];
  w[c] = partition_weight(type, smoothing, n, xyz, dist, apair, min_dist[k], k, cand_xyz + 3 * c, cand_w[c]);
}

// candidate c -> (atom k, reference point r): cand_off holds the natoms + 1 candidate offsets
__device__ inline void candidate_of(int64_t c, int natoms, const int64_t* __restrict__ cand_off,
                                    const int64_t* __restrict__ ref_off, int& k, int64_t& r) {
  int lo = 0, hi = natoms - 1;
  while (lo < hi) {  // last k with cand_off[k] <= c
    const int mid = (lo + hi + 1) / 2;
    if (cand_off[mid] <= c)
      lo = mid;
    else
      hi = mid - 1;
  }
  k = lo;
  r = ref_off[k] + (c - cand_off[k]);
}

__device__ inline void candidate_point(const double* __restrict__ ref_xyz, int64_t r, const double* __restrict__ xyz,
                                       int k, double* p) {
#pragma clang fp contract(off)
  p[0] = ref_xyz[3 * r + 0] + xyz[3 * k + 0];
  p[1] = ref_xyz[3 * r + 1] + xyz[3 * k + 1];
  p[2] = ref_xyz[3 * r + 2] + xyz[3 * k + 2];
}

__global__ __launch_bounds__(256) void partition_gen_kernel(int type, unsigned smoothing, int n,
                                                            const double* __restrict__ xyz,
                                                            const double* __restrict__ dist,
                                                            const double* __restrict__ apair,
                                                            const double* __restrict__ min_dist, int64_t ncand,
                                                            const double* __restrict__ ref_xyz,
                                                            const double* __restrict__ ref_w,
                                                            const int64_t* __restrict__ ref_off,
                                                            const int64_t* __restrict__ cand_off, double thr,
                                                            double* __restrict__ w, int* __restrict__ keep) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncand) return;
  int k;
  int64_t r;
  candidate_of(c, n, cand_off, ref_off, k, r);
  double p[3];
  candidate_point(ref_xyz, r, xyz, k, p);
  const double wc = partition_weight(type, smoothing, n, xyz, dist, apair, min_dist[k], k, p, ref_w[r]);
  w[c] = wc;
  keep[c] = wc > thr ? 1 : 0;
}

// kept candidate c goes to position pos[c] of the cut grid
__global__ __launch_bounds__(256) void cut_scatter_kernel(int n, const double* __restrict__ xyz, int64_t ncand,
                                                          const double* __restrict__ ref_xyz,
                                                          const int64_t* __restrict__ ref_off,
                                                          const int64_t* __restrict__ cand_off,
                                                          const double* __restrict__ w,
                                                          const int* __restrict__ keep, const int* __restrict__ pos,
                                                          double* __restrict__ out_xyz, double* __restrict__ out_w) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncand || !keep[c]) return;
  int k;
  int64_t r;
  candidate_of(c, n, cand_off, ref_off, k, r);
  double p[3];
  candidate_point(ref_xyz, r, xyz, k, p);
  const int64_t o = pos[c];
  out_xyz[3 * o + 0] = p[0];
  out_xyz[3 * o + 1] = p[1];
  out_xyz[3 * o + 2] = p[2];
  out_w[o] = w[c];
}

// per-workgroup min and max of each coordinate (part[6 * blockIdx.x + {lo x,y,z, hi x,y,z}])
__global__ __launch_bounds__(256) void bbox_partial_kernel(const double* __restrict__ pts, int64_t n,
                                                           double* __restrict__ part) {
  __shared__ double red[6][256];
  const int t = threadIdx.x;
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < n; i += (int64_t)gridDim.x * 256)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], pts[3 * i + d]);
      hi[d] = fmax(hi[d], pts[3 * i + d]);
    }
  for (int d = 0; d < 3; ++d) {
    red[d][t] = lo[d];
    red[3 + d][t] = hi[d];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s)
      for (int d = 0; d < 3; ++d) {
        red[d][t] = fmin(red[d][t], red[d][t + s]);
        red[3 + d][t] = fmax(red[3 + d][t], red[3 + d][t + s]);
      }
    __syncthreads();
  }
  if (t < 6) part[6 * blockIdx.x + t] = red[t][0];
}

// one thread: the box from the partials, then mn and spread = length / (hi - lo) as the host sort
__global__ void bbox_final_kernel(const double* __restrict__ part, int nparts, int64_t length,
                                  double* __restrict__ box) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int d = 0; d < 3; ++d) {
    double lo = HUGE_VAL, hi = -HUGE_VAL;
    for (int b = 0; b < nparts; ++b) {
      lo = fmin(lo, part[6 * b + d]);
      hi = fmax(hi, part[6 * b + 3 + d]);
    }
    box[d] = lo;
    box[3 + d] = double(length) / (hi - lo);
  }
}

__global__ __launch_bounds__(256) void hilbert_keys_kernel(const double* __restrict__ pts, int64_t n,
                                                           const double* __restrict__ box, int64_t length,
                                                           int64_t* __restrict__ key, int64_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = hilbert_key(pts + 3 * i, box, box + 3, length);
  idx[i] = i;
}

__global__ __launch_bounds__(256) void gather_points_kernel(const double* __restrict__ pts,
                                                            const double* __restrict__ w, int64_t n,
                                                            const int64_t* __restrict__ order,
                                                            double* __restrict__ out_pts, double* __restrict__ out_w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = order[i];
  out_pts[3 * i + 0] = pts[3 * s + 0];
  out_pts[3 * i + 1] = pts[3 * s + 1];
  out_pts[3 * i + 2] = pts[3 * s + 2];
  out_w[i] = w[s];
}

inline unsigned blocks_of(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_partition_gen(hipStream_t s, int type, unsigned smoothing, int natoms, const double* xyz,
                          const double* dist, const double* apair, const double* min_dist, int64_t ncand,
                          const double* ref_xyz, const double* ref_w, const int64_t* ref_off,
                          const int64_t* cand_off, double thr, double* w, int* keep) {
  if (ncand <= 0) return;
  hipLaunchKernelGGL(partition_gen_kernel, dim3(blocks_of(ncand)), dim3(256), 0, s, type, smoothing, natoms, xyz,
                     dist, apair, min_dist, ncand, ref_xyz, ref_w, ref_off, cand_off, thr, w, keep);
}

size_t grid_scan_temp_bytes(int64_t n) {
  size_t bytes = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int*)nullptr, (int*)nullptr, (int)n) != hipSuccess)
    throw std::runtime_error("grid: scan temp size query failed");
  return bytes;
}

void launch_cut(hipStream_t s, int natoms, const double* xyz, int64_t ncand, const double* ref_xyz,
                const int64_t* ref_off, const int64_t* cand_off, const double* w, const int* keep, int* pos,
                void* temp, size_t temp_bytes, double* out_xyz, double* out_w) {
  if (ncand <= 0) return;
  if (hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, keep, pos, (int)ncand, s) != hipSuccess)
    throw std::runtime_error("grid: exclusive scan failed");
  hipLaunchKernelGGL(cut_scatter_kernel, dim3(blocks_of(ncand)), dim3(256), 0, s, natoms, xyz, ncand, ref_xyz,
                     ref_off, cand_off, w, keep, pos, out_xyz, out_w);
}

constexpr int kBboxParts = 256;
void launch_bbox(hipStream_t s, const double* pts, int64_t n, int64_t length, double* part, double* box) {
  hipLaunchKernelGGL(bbox_partial_kernel, dim3(kBboxParts), dim3(256), 0, s, pts, n, part);
  hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(64), 0, s, part, kBboxParts, length, box);
}
int bbox_part_count() { return kBboxParts; }

size_t grid_sort_temp_bytes(int64_t n) {
  size_t bytes = 0;
  if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const int64_t*)nullptr, (int64_t*)nullptr,
                                                   (const int64_t*)nullptr, (int64_t*)nullptr, (int)n) !=
      hipSuccess)
    throw std::runtime_error("grid: sort temp size query failed");
  return bytes;
}

void launch_hilbert_order(hipStream_t s, const double* pts, const double* w, int64_t n, const double* box,
                          int64_t length, int64_t* key, int64_t* key_sorted, int64_t* idx, int64_t* order,
                          void* temp, size_t temp_bytes, double* out_pts, double* out_w) {
  if (n <= 0) return;
  hipLaunchKernelGGL(hilbert_keys_kernel, dim3(blocks_of(n)), dim3(256), 0, s, pts, n, box, length, key, idx);
  // stable: equal keys keep their input order, as the host's stable sort
  if (hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, key, key_sorted, idx, order, (int)n, 0, 64,
                                                   s) != hipSuccess)
    throw std::runtime_error("grid: radix sort failed");
  hipLaunchKernelGGL(gather_points_kernel, dim3(blocks_of(n)), dim3(256), 0, s, pts, w, n, order, out_pts, out_w);
}

void launch_partition(hipStream_t s, int type, unsigned smoothing, int natoms, const double* xyz, const double* dist,
                      const double* apair, const double* min_dist, int64_t ncand, const double* cand_xyz,
                      const double* cand_w, const int* cand_atom, double* w) {
  if (ncand <= 0) return;
  hipLaunchKernelGGL(partition_kernel, dim3((unsigned)((ncand + 255) / 256)), dim3(256), 0, s, type, smoothing,
                     natoms, xyz, dist, apair, min_dist, ncand, cand_xyz, cand_w, cand_atom, w);
}

}  // namespace sxc
