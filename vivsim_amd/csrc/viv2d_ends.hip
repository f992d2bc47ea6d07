// The first and last steps of a fused 2D VIV run: the first reads the
// caller's post-streaming FP32 field in place, the last pushes the collided
// populations into the caller's FP32 output (viv_fused.hpp).  A separate
// translation unit so that the build compiles them beside step2d.hip.
#include "step_impl.hpp"
#include "viv_fused.hpp"

namespace vsx {

int fused_first(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  return fused_mode_any<kModeFirst>(c, ib, src, dst, step, s);
}

int fused_last(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  return fused_mode_any<kModeLast>(c, ib, src, dst, step, s);
}

}  // namespace vsx
