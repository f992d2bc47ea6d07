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

// Streams of the pipelined steps (default priority, non-blocking): `s` runs the
// IB kernels; the 3D step also runs its bulk passes on `b`, forked from the
// caller's stream and joined back with the events (the 2D step keeps its bulk
// passes on the caller's stream and needs no events); the adjoint's reverse
// pass runs its IB kernels on `s` and `x` in turn and its recompute on `b`.
// One set per device, created on first use and kept for the process.  Calls
// that use them from several host threads at once on one device are not
// supported.
struct PipeRes {
  hipStream_t s = nullptr, b = nullptr, x = nullptr;
  hipEvent_t fork = nullptr, join_s = nullptr, join_b = nullptr;
};
int pipe_res(PipeRes **out);

// Largest LDS allocation of one workgroup on the current device (queried once
// per device).
int max_lds_per_block(int *out);
// Device address of pinned host memory (the last translation is kept).
int host_device_pointer(void *host, void **dev);

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
// FP32 populations, or 16-bit "shifted" storage of the deviation f_i - s_i
// from a fixed equilibrium s_i = feq_i(1, u_s) (all arithmetic stays FP32;
// `w` below is s_i).  Storing the deviation keeps the stored magnitude ~1e-2
// instead of ~1e-1, which is what makes 16-bit storage usable (Lehmann et
// al., 2022, "FP16S").
template <int S>
struct Store;

template <>
struct Store<VSX_STORE_F32> {
  using T = float;
  __device__ static inline float cvt(T r, float) { return r; }
  __device__ static inline float ld(const T *p, uint32_t i, float) { return p[i]; }
  __device__ static inline T enc(float v, float) { return v; }
  __device__ static inline void st(T *p, uint32_t i, float v, float w) { p[i] = enc(v, w); }
  __device__ static inline float rt(float v, float) { return v; }   // value as read back after a store
};

template <>
struct Store<VSX_STORE_F16S> {
  using T = _Float16;
  __device__ static inline float cvt(T r, float w) { return (float)r + w; }
  __device__ static inline float ld(const T *p, uint32_t i, float w) { return (float)p[i] + w; }
  __device__ static inline T enc(float v, float w) { return (_Float16)(v - w); }
  __device__ static inline void st(T *p, uint32_t i, float v, float w) { p[i] = enc(v, w); }
  __device__ static inline float rt(float v, float w) { return (float)(_Float16)(v - w) + w; }
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
  __device__ static inline float cvt(T r, float w) { return bf16_to_f32(r) + w; }
  __device__ static inline float ld(const T *p, uint32_t i, float w) { return bf16_to_f32(p[i]) + w; }
  __device__ static inline T enc(float v, float w) { return f32_to_bf16(v - w); }
  __device__ static inline void st(T *p, uint32_t i, float v, float w) { p[i] = enc(v, w); }
  __device__ static inline float rt(float v, float w) { return bf16_to_f32(f32_to_bf16(v - w)) + w; }
};

// E3M12: 16 bits = sign | 15-bit magnitude code r, decoded as
//   |x| = bits_to_float((r << 11) + (117 << 23)) - 2^-10,
// i.e. 8 binades of 4096 steps each: [0, 2^-10) in steps of 2^-22, then
// [2^(k-10) - 2^-10, 2^(k-9) - 2^-10) in steps of 2^(k-22), k = 1..7.  Against
// FP16 (11 significant bits) this keeps 12-13 significant bits over the range
// of shifted populations, |x| < 2^-2 - 2^-10; larger magnitudes (and NaN)
// saturate to the largest code.
__device__ inline float e3m12_to_f32(uint32_t h) {
  const float m = __uint_as_float(((h & 0x7fffu) << 11) + (117u << 23)) - 0x1p-10f;
  return __uint_as_float(__float_as_uint(m) | ((h & 0x8000u) << 16));
}
__device__ inline uint16_t f32_to_e3m12(float v) {
  const uint32_t t = __float_as_uint(fabsf(v) + 0x1p-10f);   // >= 2^-10, so the rebias below cannot wrap
  uint32_t r = (t + 0x3ffu + ((t >> 11) & 1u) - (117u << 23)) >> 11;   // round to nearest even
  r = r < 0x7fffu ? r : 0x7fffu;
  return (uint16_t)(r | ((__float_as_uint(v) >> 16) & 0x8000u));
}

template <>
struct Store<VSX_STORE_E3M12S> {
  using T = uint16_t;
  __device__ static inline float cvt(T r, float w) { return e3m12_to_f32(r) + w; }
  __device__ static inline float ld(const T *p, uint32_t i, float w) { return e3m12_to_f32(p[i]) + w; }
  __device__ static inline T enc(float v, float w) { return f32_to_e3m12(v - w); }
  __device__ static inline void st(T *p, uint32_t i, float v, float w) { p[i] = enc(v, w); }
  __device__ static inline float rt(float v, float w) { return e3m12_to_f32(f32_to_e3m12(v - w)) + w; }
};

// Storage B written with non-temporal stores: for a field larger than the
// last-level cache (MI355X: 256 MB of Infinity Cache), whose cells a step
// writes are evicted before the next step reads them.  Below that size the
// plain stores are faster: the next step reads the field from the cache.
template <class B>
struct Streamed : B {
  using T = typename B::T;
  __device__ static inline void st(T *p, uint32_t i, float v, float w) { __builtin_nontemporal_store(B::enc(v, w), p + i); }
};
constexpr size_t kLastLevelCacheBytes = size_t(256) << 20;

inline size_t store_bytes(int s) { return s == VSX_STORE_F32 ? 4 : 2; }

// ------------------------------------------------------------------ hand-off between workgroups
// Workgroups of one launch hand each other small buffers (the IB window
// velocity and force window).  Each access below is coherent at agent scope
// by itself (relaxed agent-scope atomics: the L2 of the XCD is bypassed), so
// a hand-off is: the writer's threads store with st_agent, wait until their
// stores have completed (wait_mem), meet at a barrier, and one thread sets a
// flag; the reader polls the flag and loads with ld_agent.  No whole-L2
// writeback or invalidation (what an agent-scope fence costs on gfx950,
// whose XCDs have separate L2s) is needed.
__device__ __forceinline__ void st_agent(float *p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_agent(const float *p) {
  return __hip_atomic_load(const_cast<float *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent(const double *p) {
  return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until every vector memory operation of this thread has completed (its
// agent-scope stores have reached the coherent level).  The memory clobber
// keeps the compiler from moving memory operations across the wait: a
// hand-off relies on this completion wait, not on release / acquire ordering.
__device__ __forceinline__ void wait_mem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

}  // namespace vsx
