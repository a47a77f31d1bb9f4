// MFMA wrappers for CDNA4 (gfx950): 16x16x4 FP64 and FP32 tiles on one wave64.
//   A operand: lane l holds A(l % 16, l / 16)      (16 x 4 slice)
//   B operand: lane l holds B(l / 16, l % 16)      (4 x 16 slice)
//   accumulator element r of lane l:
//     FP64: D(l / 16 + 4 r, l % 16)        FP32: D(4 (l / 16) + r, l % 16)
#pragma once

#include <hip/hip_runtime.h>

namespace dftbhip {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = v4d;
  static constexpr bool kRowMajorQuad = false;
  __device__ static inline v4d op(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator element r of lane l
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  using acc_t = v4f;
  static constexpr bool kRowMajorQuad = true;
  __device__ static inline v4f op(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) * 4 + r; }
};

}  // namespace dftbhip
