// Immersed-boundary device helpers: discrete delta kernels and stencil geometry.
//   kernels  : ib/kernels.py:4-61 (Peskin 3/4-point, cosine 4-point) + a 2-point hat
//   stencils : ib/stencil.py:7-51 (2D, idx = sx*ny + sy), ib3d/stencil.py:7-59
// Flat stencil indices outside the grid follow JAX's documented indexing
// semantics, which the reference inherits: a negative index wraps once, a
// gather clamps to the valid range and a scatter-add drops the update.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/vivsim_amd.h"

namespace vsx {

__device__ __forceinline__ float delta_kernel(int kind, float r) {
  const float a = fabsf(r);
  switch (kind) {
    case VSX_KERNEL_PESKIN_4PT:
      if (a > 2.f) return 0.f;
      if (a < 1.f) return (3.f - 2.f * a + sqrtf(1.f + 4.f * a - 4.f * (a * a))) * 0.125f;
      return (5.f - 2.f * a - sqrtf(-7.f + 12.f * a - 4.f * (a * a))) * 0.125f;
    case VSX_KERNEL_PESKIN_3PT:
      if (a > 1.5f) return 0.f;
      if (a < 0.5f) return (1.f + sqrtf(1.f - 3.f * (a * a))) / 3.f;
      return (5.f - 3.f * a - sqrtf(-2.f + 6.f * a - 3.f * (a * a))) / 6.f;
    case VSX_KERNEL_COSINE_4PT:
      if (a > 2.f) return 0.f;
      return (1.f + cosf(3.14159265358979323846f * a * 0.5f)) * 0.25f;
    default:  // 2-point hat
      return a > 1.f ? 0.f : 1.f - a;
  }
}

// d phi / d r of the kernels above (the adjoint of the stencil weights with
// respect to the marker position): d phi / d|r| times sign(r).
__device__ __forceinline__ float delta_kernel_deriv(int kind, float r) {
  const float a = fabsf(r);
  const float sg = r > 0.f ? 1.f : (r < 0.f ? -1.f : 0.f);
  float da;
  switch (kind) {
    case VSX_KERNEL_PESKIN_4PT:
      if (a > 2.f) return 0.f;
      if (a < 1.f) da = (-2.f + (4.f - 8.f * a) / (2.f * sqrtf(1.f + 4.f * a - 4.f * (a * a)))) * 0.125f;
      else da = (-2.f - (12.f - 8.f * a) / (2.f * sqrtf(-7.f + 12.f * a - 4.f * (a * a)))) * 0.125f;
      break;
    case VSX_KERNEL_PESKIN_3PT:
      if (a > 1.5f) return 0.f;
      if (a < 0.5f) da = -a / sqrtf(1.f - 3.f * (a * a));
      else da = (-3.f - (6.f - 6.f * a) / (2.f * sqrtf(-2.f + 6.f * a - 3.f * (a * a)))) / 6.f;
      break;
    case VSX_KERNEL_COSINE_4PT:
      if (a > 2.f) return 0.f;
      da = -0.125f * 3.14159265358979323846f * sinf(3.14159265358979323846f * a * 0.5f);
      break;
    default:
      if (a > 1.f) return 0.f;
      da = -1.f;
      break;
  }
  return da * sg;
}

// gather index with JAX semantics (wrap negative once, clamp)
__device__ __forceinline__ int64_t gather_index(int64_t i, int64_t n) {
  if (i < 0) i += n;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// 32-bit forms for indices into an IB window (|i| < 2 n < 2^31)
__device__ __forceinline__ int gather_index32(int i, int n) {
  if (i < 0) i += n;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
__device__ __forceinline__ int scatter_index32(int i, int n) {
  if (i < 0) i += n;
  return (i < 0 || i >= n) ? -1 : i;
}

// scatter index with JAX semantics (wrap negative once, drop out of range -> -1)
__device__ __forceinline__ int64_t scatter_index(int64_t i, int64_t n) {
  if (i < 0) i += n;
  return (i < 0 || i >= n) ? -1 : i;
}

// Separable 1D stencil of one marker along one axis: nodes ref-R+1 .. ref+R,
// ref = floor(x) (int32), weights phi(node - x) evaluated in float32.
template <int NS>
struct Axis {
  int ref;
  float w[NS];
};

template <int NS>
__device__ __forceinline__ Axis<NS> axis_stencil(int kind, float x) {
  Axis<NS> s;
  s.ref = (int)floorf(x);
#pragma unroll
  for (int k = 0; k < NS; ++k) s.w[k] = delta_kernel(kind, (float)(s.ref - NS / 2 + 1 + k) - x);
  return s;
}

}  // namespace vsx
