// D3Q19 fused steps.
#include "step_impl.hpp"

using namespace vsx;

extern "C" int vsx_d3q19_step(const VsxLbm *p, const void *src, int32_t src_f32, void *dst, int32_t dst_f32,
                              int32_t phase, void *stream) {
  return step_entry<D3Q19>(p, src, src_f32, dst, dst_f32, phase, stream);
}

extern "C" int vsx_d3q19_run(const VsxLbm *p, const float *f_in, float *f_out, void *work_a, void *work_b,
                             int32_t n_steps, void *stream) {
  return run_entry<D3Q19>(p, f_in, f_out, work_a, work_b, n_steps, stream);
}
