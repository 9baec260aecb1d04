// MFMA 16x16x4 wrappers for FP64 and FP32 operands (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace pqg {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4v acc_t;
  // lane l: A[l%16][l/16], B[l/16][l%16]; register r holds D[4*r + l/16][l%16]
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int lane) { return 4 * r + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef f4v acc_t;
  // lane l: A[l%16][l/16], B[l/16][l%16]; register r holds D[4*(l/16) + r][l%16]
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int lane) { return 4 * (lane >> 4) + r; }
};

}  // namespace pqg
