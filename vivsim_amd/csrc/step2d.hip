// D2Q9 fused steps and the 2D IB-LBM vortex-induced-vibration driver.
#include "step_impl.hpp"
#include "viv_fused.hpp"

namespace vsx {

// Stand-alone IB kernel (the first step of a run, and the non-fused path):
// the fixed-point LDS spread of the fused steps for up to one marker per
// thread, the binned spread for more markers (MPT > 1); TPM: threads per
// marker (ib_tpm) of the former.
template <class SRC, bool PULL, int MPT, int TPM = 1>
__global__ __launch_bounds__(kIbThreads) void k_viv_ib(StepCtx c_arg, VsxIB2D ib,
                                                        const typename SRC::T *__restrict__ src, int step) {
  const StepCtx &c = kernel_ctx();
  extern __shared__ float lds[];
  const int gstep = ib.step0 + step, slot = gstep & 1;
  if constexpr (MPT == 1)
    ib_atomic<SRC, PULL ? kWinPull : kWinLocal, 2, 3, true, TPM>(c, ib, src, nullptr, step, gstep, nullptr,
                                                            ib.body + 6 * slot, ib.body + 6 * (slot ^ 1), lds);
  else
    ib_block<SRC, PULL, MPT, 4, 0, false, 4>(c, ib, src, step, gstep, 0, ib.body + 6 * slot,
                                          ib.body + 6 * (slot ^ 1), lds);
  publish_ready(ib, gstep + 1);
}

// ------------------------------------------------------------------ host side

template <class SRC, bool PULL>
int launch_viv_ib(const StepCtx &c, const VsxIB2D &ib, const void *src, int step, hipStream_t s) {
  auto S = static_cast<const typename SRC::T *>(src);
  int max_lds = 0;
  if (int rc = max_lds_per_block(&max_lds)) return rc;
  const size_t lds_atomic = 4 * (size_t)ib_split_lds_words(ib.win);
  if (ib.n_markers <= kIbThreads && lds_atomic + 2048 <= (size_t)max_lds) {
    switch (ib_tpm(ib.n_markers, kIbThreads)) {
      case 4: hipLaunchKernelGGL((k_viv_ib<SRC, PULL, 1, 4>), dim3(1), dim3(kIbThreads), lds_atomic, s, c, ib, S, step); break;
      case 2: hipLaunchKernelGGL((k_viv_ib<SRC, PULL, 1, 2>), dim3(1), dim3(kIbThreads), lds_atomic, s, c, ib, S, step); break;
      default: hipLaunchKernelGGL((k_viv_ib<SRC, PULL, 1, 1>), dim3(1), dim3(kIbThreads), lds_atomic, s, c, ib, S, step); break;
    }
  } else {
    const size_t lds = 4 * (size_t)ib_lds_words(ib.win, ib.n_markers, false);
    hipLaunchKernelGGL((k_viv_ib<SRC, PULL, kMarkersPerThread>), dim3(1), dim3(kIbThreads), lds, s, c, ib, S, step);
  }
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

int viv_ib_any(const StepCtx &c, const VsxIB2D &ib, const void *src, bool first, int step, hipStream_t s) {
  if (first) return launch_viv_ib<Store<VSX_STORE_F32>, false>(c, ib, src, step, s);
  switch (c.p.storage) {
    case VSX_STORE_F32: return launch_viv_ib<Store<VSX_STORE_F32>, true>(c, ib, src, step, s);
    case VSX_STORE_F16S: return launch_viv_ib<Store<VSX_STORE_F16S>, true>(c, ib, src, step, s);
    case VSX_STORE_BF16S: return launch_viv_ib<Store<VSX_STORE_BF16S>, true>(c, ib, src, step, s);
    default: return launch_viv_ib<Store<VSX_STORE_E3M12S>, true>(c, ib, src, step, s);
  }
}

Tiles make_tiles(const StepCtx &c) {
  Tiles T;
  int ty = ((c.dims[1] + 63) / 64) * 64;
  T.TY = ty > 256 ? 256 : ty;
  T.TX = kTileCells / T.TY;
  T.ntx = (c.dims[0] + T.TX - 1) / T.TX;
  T.nty = (c.dims[1] + T.TY - 1) / T.TY;
  return T;
}

int fused_full(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  return fused_mode_any<kModeFull>(c, ib, src, dst, step, s);
}


int viv_run(const VsxLbm *p, const VsxIB2D *ib_in, const float *f_in, float *f_out, void *wa, void *wb, int n,
            void *stream) {
  int rc = validate_lbm<D2Q9>(p);
  if (rc) return rc;
  rc = validate_ib2d(p, ib_in);
  if (rc) return rc;
  // the kernels' view of the IB parameters: the status word as a device address
  VsxIB2D ib_dev = *ib_in;
  if (ib_dev.status) {
    void *d = nullptr;
    if ((rc = host_device_pointer(ib_in->status, &d))) return rc;
    ib_dev.status = static_cast<int32_t *>(d);
  }
  const VsxIB2D *ib = &ib_dev;
  VSX_REQUIRE(f_in && f_out && wa, "null buffer");
  VSX_REQUIRE(n >= 0, "negative step count");
  VSX_REQUIRE(n <= 1 || wb, "need two work buffers for more than one step");
  hipStream_t s = as_stream(stream);
  const StepCtx c = make_ctx<D2Q9>(*p, Window{ib->win_origin, ib->g_win, {ib->win, ib->win, 1}});
  if (n == 0) {
    if (f_out != f_in) VSX_HIP(hipMemcpyAsync(f_out, f_in, sizeof(float) * 9 * c.n, hipMemcpyDeviceToDevice, s));
    return VSX_OK;
  }
  // fused launches need one marker per thread of the IB workgroup, and walls
  // without corner cells (every window cell is then finished by its tile)
  const bool fused = ib->fused && ib->n_markers <= kIbThreads && !(p->n_walls > 0 && walls_need_edge<D2Q9>(*p));
  // the last step pushes into f_out (then the walls are applied in place) unless
  // a wall needs the pre-streaming populations
  const bool push = fused && n > 1 && walls_allow_push(*p);
  const void *cur = f_in;
  void *bufs[2] = {wa, wb};
  if (fused && ib->fused == 2) {
    // pipelined: bulk passes on the caller's stream, IB kernels on a second
    // stream.  No stream events: the first IB kernel waits (in the kernel) for
    // the first bulk pass to start, each later one for the previous pass's
    // hand-off, and every bulk pass for its IB kernel's force window, so the
    // caller's stream is complete only after every IB kernel of the run has
    // written its last result (IB kernels of consecutive runs keep their order
    // on the second stream).
    VSX_REQUIRE(ib->halo, "the pipelined VIV step needs the hand-off buffer ib->halo");
    PipeRes *r = nullptr;
    if ((rc = pipe_res(&r))) return rc;
    for (int t = 0; t < n; ++t) {
      if ((rc = launch_ib_pipe(c, *ib, f_in, t == 0, t, r->s))) return rc;
      void *dst = (push && t == n - 1) ? f_out : bufs[t & 1];
      if (t == 0) rc = fused_first(c, *ib, cur, dst, t, s);
      else if (push && t == n - 1) rc = fused_last(c, *ib, cur, dst, t, s);
      else rc = fused_full(c, *ib, cur, dst, t, s);
      if (rc) return rc;
      cur = dst;
    }
    if (push) return walls_in_place<D2Q9>(c, f_out, s);
    const StepCtx cs = make_ctx<D2Q9>(*p, Window{nullptr, nullptr, {0, 0, 0}});
    return step_any<D2Q9>(cs, cur, false, f_out, true, VSX_PHASE_STREAM, s);
  }
  for (int t = 0; t < n; ++t) {
    void *dst = bufs[t & 1];
    if (!fused) {
      rc = viv_ib_any(c, *ib, cur, t == 0, t, s);
      if (rc) return rc;
      rc = step_any<D2Q9>(c, cur, t == 0, dst, false, t == 0 ? VSX_PHASE_COLLIDE : VSX_PHASE_FULL, s);
    } else if (t == 0) {
      rc = fused_first(c, *ib, cur, dst, t, s);
    } else if (push && t == n - 1) {
      rc = fused_last(c, *ib, cur, f_out, t, s);
      if (rc) return rc;
      return walls_in_place<D2Q9>(c, f_out, s);
    } else {
      rc = fused_full(c, *ib, cur, dst, t, s);
    }
    if (rc) return rc;
    cur = dst;
  }
  const StepCtx cs = make_ctx<D2Q9>(*p, Window{nullptr, nullptr, {0, 0, 0}});
  return step_any<D2Q9>(cs, cur, false, f_out, true, VSX_PHASE_STREAM, s);
}

}  // namespace vsx

using namespace vsx;

extern "C" int vsx_d2q9_step(const VsxLbm *p, const void *src, int32_t src_f32, void *dst, int32_t dst_f32,
                             int32_t phase, void *stream) {
  return step_entry<D2Q9>(p, src, src_f32, dst, dst_f32, phase, stream);
}

extern "C" int vsx_d2q9_run(const VsxLbm *p, const float *f_in, float *f_out, void *work_a, void *work_b,
                            int32_t n_steps, void *stream) {
  return run_entry<D2Q9>(p, f_in, f_out, work_a, work_b, n_steps, stream);
}

extern "C" int vsx_viv2d_run(const VsxLbm *p, const VsxIB2D *ib, const float *f_in, float *f_out, void *work_a,
                             void *work_b, int32_t n_steps, void *stream) {
  return viv_run(p, ib, f_in, f_out, work_a, work_b, n_steps, stream);
}
