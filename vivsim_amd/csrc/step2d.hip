// D2Q9 fused steps and the 2D IB-LBM vortex-induced-vibration driver.
#include <cmath>

#include "ib_common.hpp"
#include "step_impl.hpp"

namespace vsx {

constexpr int kIbThreads = 1024;
constexpr int kMarkersPerThread = 4;

// ------------------------------------------------------------------ VIV immersed-boundary kernel
// One workgroup per body (examples/2d/vortex_induced_vibration.py:99-137):
//   window origin  int32(IB_X0 + d) (truncation, :104-105), clamped like lax.dynamic_slice
//   window u       macroscopic velocity of the post-streaming state, kept in LDS
//   stencil        ib/stencil.py:7-51 at marker coordinates (x0 + d) - origin, 4x4 Peskin/cosine
//   MDF            ib/mdf.py:35-64: interpolate, n_iter x (spread dF into LDS with LDS
//                  atomics, interpolate 0.5 * spread), final spread of the total force
//   body force     h = sum_m (-F_m) + a * area in FP64 (wave shuffles + LDS), dyn.py:126-136
//   structure      Newmark-beta in FP64 on lane 0, dyn.py:5-51
// The spread force window is written to global memory for the bulk kernel.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <class SRC, bool PULL, int MPT>
__global__ __launch_bounds__(kIbThreads) void k_viv_ib(StepCtx c, VsxIB2D ib, const typename SRC::T *__restrict__ src,
                                                        int step) {
  extern __shared__ float lds[];
  const int S = ib.win, SS = S * S;
  float *ul = lds;           // [2][SS]
  float *gl = lds + 2 * SS;  // [2][SS]
  __shared__ double s_body[6];
  __shared__ int s_org[4];
  __shared__ double s_red[2][kIbThreads / 64];
  const int tid = threadIdx.x, nt = blockDim.x;

  if (tid < 6) s_body[tid] = ib.body[tid];
  __syncthreads();
  if (tid == 0) {
    const int x0 = (int)trunc(ib.win_x0 + s_body[4]);
    const int y0 = (int)trunc(ib.win_y0 + s_body[5]);
    const int xc = min(max(x0, 0), c.dims[0] - S);
    const int yc = min(max(y0, 0), c.dims[1] - S);
    s_org[0] = x0;
    s_org[1] = y0;
    s_org[2] = xc;
    s_org[3] = yc;
    ib.win_origin[0] = xc;
    ib.win_origin[1] = yc;
  }
  __syncthreads();
  const int x0 = s_org[0], y0 = s_org[1], xc = s_org[2], yc = s_org[3];

  // window velocity: kUW cells per thread per pass, all loads issued before use
  constexpr int kUW = 4;
  for (int base = 0; base < SS; base += kUW * nt) {
    float f[kUW][9];
#pragma unroll
    for (int j = 0; j < kUW; ++j) {
      const int k = base + j * nt + tid;
      if (k < SS) {
        const int co[3] = {xc + k / S, yc + k % S, 0};
        if constexpr (!PULL) local<D2Q9, SRC>(c, src, cell_of<D2Q9>(c, co[0], co[1], 0), f[j]);
        else cell_state<D2Q9, SRC>(c, src, co, f[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < kUW; ++j) {
      const int k = base + j * nt + tid;
      if (k < SS) {
        float rho, u[3];
        macroscopic<D2Q9>(f[j], rho, u);
        ul[k] = u[0];
        ul[SS + k] = u[1];
      }
    }
  }

  // marker stencils and initial interpolation
  const float Ux = (float)s_body[2], Uy = (float)s_body[3];
  Axis<4> ax[MPT], ay[MPT];
  float ubx[MPT], uby[MPT], Fx[MPT], Fy[MPT];
  float dFx[MPT], dFy[MPT], ds2[MPT];
  __syncthreads();
  // Marker q of this thread: consecutive lanes of a wavefront hold markers nt/64
  // apart along the body, so the 64 lanes of one LDS atomic (spread) or read
  // (interpolate) touch different cells instead of serialising on shared ones.
  int mk[MPT];
  const int nwave = nt >> 6;
#pragma unroll
  for (int q = 0; q < MPT; ++q) mk[q] = q * nt + (tid & 63) * nwave + (tid >> 6);
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    const int m = mk[q];
    ubx[q] = uby[q] = Fx[q] = Fy[q] = dFx[q] = dFy[q] = ds2[q] = 0.f;
    if (m < ib.n_markers) {
      const float rx = (float)(ib.marker_x[m] + s_body[4] - (double)x0);
      const float ry = (float)(ib.marker_y[m] + s_body[5] - (double)y0);
      ax[q] = axis_stencil<4>(ib.kernel, rx);
      ay[q] = axis_stencil<4>(ib.kernel, ry);
      ds2[q] = (ib.ds_arr ? ib.ds_arr[m] : ib.ds) * 2.0f;
      float sx = 0.f, sy = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t gi = gather_index((int64_t)(ax[q].ref - 1 + i) * S + (ay[q].ref - 1 + j), SS);
          const float w = ax[q].w[i] * ay[q].w[j];
          sx += w * ul[gi];
          sy += w * ul[SS + gi];
        }
      ubx[q] = sx;
      uby[q] = sy;
    }
  }

  for (int it = 0; it <= ib.n_iter; ++it) {
    const bool last = it == ib.n_iter;
    // dF = (U - ub) * 2 ds ; F += dF   (the final pass spreads F itself)
#pragma unroll
    for (int q = 0; q < MPT; ++q) {
      if (!last) {
        dFx[q] = (Ux - ubx[q]) * ds2[q];
        dFy[q] = (Uy - uby[q]) * ds2[q];
        Fx[q] += dFx[q];
        Fy[q] += dFy[q];
      } else {
        dFx[q] = Fx[q];
        dFy[q] = Fy[q];
      }
    }
    for (int k = tid; k < 2 * SS; k += nt) gl[k] = 0.f;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MPT; ++q) {
      if (mk[q] < ib.n_markers) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int64_t si = scatter_index((int64_t)(ax[q].ref - 1 + i) * S + (ay[q].ref - 1 + j), SS);
            if (si >= 0) {
              const float w = ax[q].w[i] * ay[q].w[j];
              atomicAdd(&gl[si], dFx[q] * w);
              atomicAdd(&gl[SS + si], dFy[q] * w);
            }
          }
      }
    }
    __syncthreads();
    if (last) break;
#pragma unroll
    for (int q = 0; q < MPT; ++q) {
      if (mk[q] < ib.n_markers) {
        float sx = 0.f, sy = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int64_t gi = gather_index((int64_t)(ax[q].ref - 1 + i) * S + (ay[q].ref - 1 + j), SS);
            const float w = ax[q].w[i] * ay[q].w[j];
            sx += w * (gl[gi] * 0.5f);
            sy += w * (gl[SS + gi] * 0.5f);
          }
        ubx[q] += sx;
        uby[q] += sy;
      }
    }
    __syncthreads();
  }

  // force window -> global
  for (int k = tid; k < 2 * SS; k += nt) ib.g_win[k] = gl[k];

  // body force h = sum(-F) (FP64)
  double hx = 0.0, hy = 0.0;
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    hx -= (double)Fx[q];
    hy -= (double)Fy[q];
  }
  hx = wave_sum(hx);
  hy = wave_sum(hy);
  if ((tid & 63) == 0) {
    s_red[0][tid >> 6] = hx;
    s_red[1][tid >> 6] = hy;
  }
  __syncthreads();
  if (tid == 0) {
    double h[2] = {0.0, 0.0};
    for (int wv = 0; wv < (nt + 63) / 64; ++wv) {
      h[0] += s_red[0][wv];
      h[1] += s_red[1][wv];
    }
    const double c1 = ib.gamma * ib.dt, c2 = ib.beta * ib.dt * ib.dt;
    double out[6];
    for (int k = 0; k < 2; ++k) {
      const double a = s_body[k], v = s_body[2 + k], d = s_body[4 + k];
      h[k] += a * ib.area;
      const double v1 = v + ib.dt * (1.0 - ib.gamma) * a;
      const double v2 = d + ib.dt * v + ib.dt * ib.dt * (0.5 - ib.beta) * a;
      const double an = (h[k] - ib.damping * v1 - ib.stiffness * v2) / (ib.mass + c1 * ib.damping + c2 * ib.stiffness);
      const bool on = (ib.dof_mask >> k) & 1;
      out[k] = on ? an : 0.0;
      out[2 + k] = on ? c1 * an + v1 : 0.0;
      out[4 + k] = on ? c2 * an + v2 : 0.0;
    }
    for (int k = 0; k < 6; ++k) ib.body[k] = out[k];
    if (ib.hist_d) {
      ib.hist_d[2 * step] = out[4];
      ib.hist_d[2 * step + 1] = out[5];
    }
    if (ib.hist_h) {
      ib.hist_h[2 * step] = h[0];
      ib.hist_h[2 * step + 1] = h[1];
    }
  }
}

template <class SRC, bool PULL>
int launch_viv_ib(const StepCtx &c, const VsxIB2D &ib, const void *src, int step, hipStream_t s) {
  const size_t lds = sizeof(float) * 4 * (size_t)ib.win * ib.win;
  auto S = static_cast<const typename SRC::T *>(src);
  if (ib.n_markers <= kIbThreads)
    hipLaunchKernelGGL((k_viv_ib<SRC, PULL, 1>), dim3(1), dim3(kIbThreads), lds, s, c, ib, S, step);
  else
    hipLaunchKernelGGL((k_viv_ib<SRC, PULL, kMarkersPerThread>), dim3(1), dim3(kIbThreads), lds, s, c, ib, S, step);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

int viv_ib_any(const StepCtx &c, const VsxIB2D &ib, const void *src, bool first, int step, hipStream_t s) {
  if (first) return launch_viv_ib<Store<VSX_STORE_F32>, false>(c, ib, src, step, s);
  switch (c.p.storage) {
    case VSX_STORE_F32: return launch_viv_ib<Store<VSX_STORE_F32>, true>(c, ib, src, step, s);
    case VSX_STORE_F16S: return launch_viv_ib<Store<VSX_STORE_F16S>, true>(c, ib, src, step, s);
    default: return launch_viv_ib<Store<VSX_STORE_BF16S>, true>(c, ib, src, step, s);
  }
}

int validate_ib2d(const VsxLbm *p, const VsxIB2D *ib) {
  if (!ib) return fail(VSX_ERR_VALUE, "null IB parameters");
  VSX_REQUIRE(ib->n_markers > 0 && ib->n_markers <= kIbThreads * kMarkersPerThread, "marker count %d out of range (1..%d)",
              ib->n_markers, kIbThreads * kMarkersPerThread);
  VSX_REQUIRE(ib->win >= 4 && ib->win <= p->nx && ib->win <= p->ny, "IB window %d does not fit the grid", ib->win);
  VSX_REQUIRE(ib->kernel >= VSX_KERNEL_PESKIN_2PT && ib->kernel <= VSX_KERNEL_COSINE_4PT, "unknown kernel %d", ib->kernel);
  VSX_REQUIRE(ib->n_iter >= 0, "negative n_iter");
  VSX_REQUIRE(ib->marker_x && ib->marker_y && ib->body && ib->g_win && ib->win_origin, "null IB buffer");
  VSX_REQUIRE(ib->mass + ib->gamma * ib->dt * ib->damping + ib->beta * ib->dt * ib->dt * ib->stiffness != 0.0,
              "singular Newmark system");
  int dev = 0, max_lds = 0;
  VSX_HIP(hipGetDevice(&dev));
  VSX_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t need = sizeof(float) * 4 * (size_t)ib->win * ib->win + 1024;
  VSX_REQUIRE(need <= (size_t)max_lds, "IB window %d needs %zu B of LDS, device allows %d", ib->win, need, max_lds);
  return VSX_OK;
}

int viv_run(const VsxLbm *p, const VsxIB2D *ib, const float *f_in, float *f_out, void *wa, void *wb, int n,
            void *stream) {
  int rc = validate_lbm<D2Q9>(p);
  if (rc) return rc;
  rc = validate_ib2d(p, ib);
  if (rc) return rc;
  VSX_REQUIRE(f_in && f_out && wa, "null buffer");
  VSX_REQUIRE(n >= 0, "negative step count");
  VSX_REQUIRE(n <= 1 || wb, "need two work buffers for more than one step");
  hipStream_t s = as_stream(stream);
  const StepCtx c = make_ctx<D2Q9>(*p, Window{ib->win_origin, ib->g_win, {ib->win, ib->win, 1}});
  if (n == 0) {
    if (f_out != f_in) VSX_HIP(hipMemcpyAsync(f_out, f_in, sizeof(float) * 9 * c.n, hipMemcpyDeviceToDevice, s));
    return VSX_OK;
  }
  const void *cur = f_in;
  void *bufs[2] = {wa, wb};
  for (int t = 0; t < n; ++t) {
    void *dst = bufs[t & 1];
    rc = viv_ib_any(c, *ib, cur, t == 0, t, s);
    if (rc) return rc;
    rc = step_any<D2Q9>(c, cur, t == 0, dst, false, t == 0 ? VSX_PHASE_COLLIDE : VSX_PHASE_FULL, s);
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
