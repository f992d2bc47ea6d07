// Library-level entry points: version, error text, device sync.
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace vsx {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

// ------------------------------------------------------------------ bulk-kernel timer
namespace {
struct Timer {
  std::vector<hipEvent_t> a, b;
  int n = 0;
  bool on = false;
};
Timer g_timer;
}  // namespace

// no event pairs while a stream is being captured into a graph
static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

void timer_begin(hipStream_t s) {
  if (g_timer.on && g_timer.n < (int)g_timer.a.size() && !capturing(s)) (void)hipEventRecord(g_timer.a[g_timer.n], s);
}

void timer_end(hipStream_t s) {
  if (g_timer.on && g_timer.n < (int)g_timer.a.size() && !capturing(s)) {
    (void)hipEventRecord(g_timer.b[g_timer.n], s);
    ++g_timer.n;
  }
}

constexpr int kMaxDevices = 64;
PipeRes g_pipe[kMaxDevices];
std::mutex g_pipe_mu;

int max_lds_per_block(int *out) {
  static int cache[64] = {0};
  int dev = 0;
  VSX_HIP(hipGetDevice(&dev));
  VSX_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
  if (cache[dev] == 0) {
    int v = 0;
    VSX_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cache[dev] = v;
  }
  *out = cache[dev];
  return VSX_OK;
}

int host_device_pointer(void *host, void **dev_ptr) {
  static thread_local void *last_host = nullptr, *last_dev = nullptr;
  if (host != last_host || !last_dev) {
    void *d = nullptr;
    VSX_HIP(hipHostGetDevicePointer(&d, host, 0));
    last_host = host;
    last_dev = d;
  }
  *dev_ptr = last_dev;
  return VSX_OK;
}

int pipe_res(PipeRes **out) {
  int dev = 0;
  VSX_HIP(hipGetDevice(&dev));
  VSX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_pipe_mu);
  PipeRes &r = g_pipe[dev];
  if (!r.s) {
    VSX_HIP(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
    VSX_HIP(hipStreamCreateWithFlags(&r.b, hipStreamNonBlocking));
    VSX_HIP(hipStreamCreateWithFlags(&r.x, hipStreamNonBlocking));
    // ordering only: every kernel launch already acquires at its start and
    // releases at its end, so the events need no cache fence of their own
    const unsigned ev = hipEventDisableTiming | hipEventDisableSystemFence;
    VSX_HIP(hipEventCreateWithFlags(&r.fork, ev));
    VSX_HIP(hipEventCreateWithFlags(&r.join_s, ev));
    VSX_HIP(hipEventCreateWithFlags(&r.join_b, ev));
  }
  *out = &r;
  return VSX_OK;
}

}  // namespace vsx

extern "C" int vsx_timer_start(int32_t max_pairs) {
  using vsx::g_timer;
  VSX_REQUIRE(max_pairs > 0 && max_pairs <= (1 << 20), "max_pairs out of range");
  while ((int)g_timer.a.size() < max_pairs) {
    hipEvent_t x, y;
    VSX_HIP(hipEventCreate(&x));
    VSX_HIP(hipEventCreate(&y));
    g_timer.a.push_back(x);
    g_timer.b.push_back(y);
  }
  g_timer.n = 0;
  g_timer.on = true;
  return VSX_OK;
}

extern "C" int vsx_timer_stop(float *ms, int32_t max, int32_t *n_out) {
  using vsx::g_timer;
  g_timer.on = false;
  const int n = g_timer.n < max ? g_timer.n : max;
  if (n > 0) VSX_HIP(hipEventSynchronize(g_timer.b[n - 1]));
  for (int i = 0; i < n; ++i) VSX_HIP(hipEventElapsedTime(&ms[i], g_timer.a[i], g_timer.b[i]));
  if (n_out) *n_out = n;
  return VSX_OK;
}

extern "C" const char *vsx_version(void) { return "vivsim_amd 0.1.0 (gfx950)"; }

extern "C" const char *vsx_last_error(void) { return vsx::g_last_error.c_str(); }

extern "C" int vsx_device_sync(void) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return vsx::fail(VSX_ERR_HIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  return VSX_OK;
}
