// Matrix-core building blocks for gfx950 (CDNA4): 16x16 output tiles with a
// K = 4 step, one A and one B element per lane and four accumulator elements
// per lane, in FP64 (v_mfma_f64_16x16x4_f64) and FP32 (v_mfma_f32_16x16x4_f32).
//
// Lane layout of one 16x16x4 step for a 64-wide wavefront (measured on gfx950):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l, element r holds D[row(l, r)][l % 16] with
//                row = l / 16 + 4 r        for the FP64 instruction,
//                row = 4 (l / 16) + r      for the FP32 instruction.
#pragma once

#include <hip/hip_runtime.h>

namespace sxc {

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef dvec4 acc_t;
  __device__ static inline acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef fvec4 acc_t;
  __device__ static inline acc_t zero() { return acc_t{0.0f, 0.0f, 0.0f, 0.0f}; }
  __device__ static inline acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

}  // namespace sxc
