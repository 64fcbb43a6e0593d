// Molecular-grid partition on the GPU: one thread per candidate point, the
// same per-point arithmetic as the host loop (host/partition.h).  The
// candidates' positions, raw weights and atoms come from the host; the weight
// cut and the Hilbert ordering follow on the host.
#include <hip/hip_runtime.h>

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

}  // namespace

void launch_partition(hipStream_t s, int type, unsigned smoothing, int natoms, const double* xyz, const double* dist,
                      const double* apair, const double* min_dist, int64_t ncand, const double* cand_xyz,
                      const double* cand_w, const int* cand_atom, double* w) {
  if (ncand <= 0) return;
  hipLaunchKernelGGL(partition_kernel, dim3((unsigned)((ncand + 255) / 256)), dim3(256), 0, s, type, smoothing,
                     natoms, xyz, dist, apair, min_dist, ncand, cand_xyz, cand_w, cand_atom, w);
}

}  // namespace sxc
