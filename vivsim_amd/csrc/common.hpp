// Shared host/device helpers: status codes, error text, launch checks and the
// population storage formats of the fused steps.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/vivsim_amd.h"

namespace vsx {

void set_error(const char *fmt, ...);
// event pair around a kernel launch while vsx_timer_start() is active
void timer_begin(hipStream_t s);
void timer_end(hipStream_t s);
int fail(int code, const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Checks the launch (and any earlier asynchronous error) of the kernels just queued.
#define VSX_CHECK_LAUNCH()                                                              \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess) return ::vsx::fail(VSX_ERR_HIP, "%s: %s", __func__, hipGetErrorString(e_)); \
  } while (0)

#define VSX_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return ::vsx::fail(VSX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define VSX_REQUIRE(cond, ...)                               \
  do {                                                       \
    if (!(cond)) return ::vsx::fail(VSX_ERR_VALUE, __VA_ARGS__); \
  } while (0)

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > (int64_t)1 << 20) g = (int64_t)1 << 20;   // grid-stride loops cover the rest
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ------------------------------------------------------------------ storage
// FP32 populations, or FP16 / BF16 "shifted" storage of the deviation f_i - w_i
// (all arithmetic stays FP32).  Storing the deviation from the rest
// equilibrium keeps the stored magnitude ~1e-2 instead of ~1e-1, which is
// what makes 16-bit storage usable (Lehmann et al., 2022, "FP16S").
template <int S>
struct Store;

template <>
struct Store<VSX_STORE_F32> {
  using T = float;
  __device__ static inline float ld(const T *p, uint32_t i, float) { return p[i]; }
  __device__ static inline void st(T *p, uint32_t i, float v, float) { p[i] = v; }
};

template <>
struct Store<VSX_STORE_F16S> {
  using T = _Float16;
  __device__ static inline float ld(const T *p, uint32_t i, float w) { return (float)p[i] + w; }
  __device__ static inline void st(T *p, uint32_t i, float v, float w) { p[i] = (_Float16)(v - w); }
};

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ inline uint16_t f32_to_bf16(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}

template <>
struct Store<VSX_STORE_BF16S> {
  using T = uint16_t;
  __device__ static inline float ld(const T *p, uint32_t i, float w) { return bf16_to_f32(p[i]) + w; }
  __device__ static inline void st(T *p, uint32_t i, float v, float w) { p[i] = f32_to_bf16(v - w); }
};

inline size_t store_bytes(int s) { return s == VSX_STORE_F32 ? 4 : 2; }

}  // namespace vsx
