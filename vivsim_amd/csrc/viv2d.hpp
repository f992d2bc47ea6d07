// The 2D IB workgroup (window velocity, multi-direct forcing, FP64 body force
// and Newmark) and the tile mapping shared by the single-case and batched
// fused VIV steps.
#pragma once
#include <cmath>

#include "ib_common.hpp"
#include "step.hpp"

namespace vsx {

constexpr int kIbThreads = 1024;
constexpr int kMarkersPerThread = 4;
constexpr int kSpinLimit = 1 << 22;   // bounded wait of a window tile for the IB workgroup

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Window origin of step t from the body state before the step:
// int32(IB_X0 + d) truncated (examples/2d/vortex_induced_vibration.py:104-105),
// and that origin clamped to the grid as lax.dynamic_slice clamps it.
__device__ __forceinline__ void window_origin(const StepCtx &c, const VsxIB2D &ib, const double *body, int (&o)[4]) {
  const int x0 = (int)trunc(ib.win_x0 + body[4]);
  const int y0 = (int)trunc(ib.win_y0 + body[5]);
  o[0] = x0;
  o[1] = y0;
  o[2] = min(max(x0, 0), c.dims[0] - ib.win);
  o[3] = min(max(y0, 0), c.dims[1] - ib.win);
}

// Window velocity of a cell that lies on a domain wall (rare: only when the IB
// window touches a wall).  Kept out of line and returning by value, so the
// register arrays of the interior path are never address-taken.
template <class SRC, int TAG>
__device__ __noinline__ float2 window_wall_u(const StepCtx &c, const typename SRC::T *__restrict__ src, int x, int y) {
  const int co[3] = {x, y, 0};
  float f[9], rho, u[3];
  cell_state<D2Q9, SRC, TAG>(c, src, co, f);
  macroscopic<D2Q9>(f, rho, u);
  return make_float2(u[0], u[1]);
}

// ------------------------------------------------------------------ the IB workgroup
// One workgroup per body (examples/2d/vortex_induced_vibration.py:99-137):
//   window u     macroscopic velocity of the post-streaming state (pull + walls), in LDS
//   stencil      ib/stencil.py:7-51 at marker coordinates (x_m + d) - origin (4x4)
//   MDF          ib/mdf.py:35-64: interpolate, n_iter x (spread dF into LDS with LDS
//                atomics, interpolate 0.5 * spread), final spread of the total force
//   body force   h = sum_m (-F_m) + a * area in FP64 (wave shuffles + LDS), dyn.py:126-136
//   structure    Newmark-beta in FP64 on lane 0, dyn.py:5-51
// `lds` holds 2 * win^2 floats: the window velocity, reused for the spread
// force once the initial interpolation is done.  Writes the force window to
// ib.g_win and the new body state to body_out.
template <class SRC, bool PULL, int MPT, int UW, int TAG>
__device__ __forceinline__ void ib_block(const StepCtx &c, const VsxIB2D &ib, const typename SRC::T *__restrict__ src,
                                         int hidx, const double *body_in, double *body_out, float *lds) {
  const int S = ib.win, SS = S * S;
  float *ul = lds;   // [2][SS] window velocity, then the spread force
  float *gl = lds;
  __shared__ double s_body[6];
  __shared__ int s_org[4];
  __shared__ double s_red[2][kIbThreads / 64];
  const int tid = threadIdx.x, nt = blockDim.x;

  if (tid < 6) s_body[tid] = body_in[tid];
  __syncthreads();
  if (tid == 0) {
    int o[4];
    window_origin(c, ib, s_body, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) s_org[k] = o[k];
    ib.win_origin[0] = o[2];
    ib.win_origin[1] = o[3];
  }
  __syncthreads();
  const int x0 = s_org[0], y0 = s_org[1], xc = s_org[2], yc = s_org[3];

  // window velocity: UW cells per thread per pass, all loads issued before use
  for (int base = 0; base < SS; base += UW * nt) {
    float f[UW][9];
    bool on_w[UW];
#pragma unroll
    for (int j = 0; j < UW; ++j) {
      const int k = base + j * nt + tid;
      on_w[j] = false;
      if (k < SS) {
        const int co[3] = {xc + k / S, yc + k % S, 0};
        if constexpr (!PULL) {
          local<D2Q9, SRC>(c, src, cell_of<D2Q9>(c, co[0], co[1], 0), f[j]);
        } else {
          on_w[j] = c.p.n_walls > 0 && wall_class<D2Q9>(c, co) != -1;
          if (!on_w[j]) pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f[j]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < UW; ++j) {
      const int k = base + j * nt + tid;
      if (k < SS) {
        if (on_w[j]) {
          const float2 uw = window_wall_u<SRC, TAG>(c, src, xc + k / S, yc + k % S);
          ul[k] = uw.x;
          ul[SS + k] = uw.y;
        } else {
          float rho, u[3];
          macroscopic<D2Q9>(f[j], rho, u);
          ul[k] = u[0];
          ul[SS + k] = u[1];
        }
      }
    }
  }

  // Marker q of this thread: consecutive lanes of a wavefront hold markers nt/64
  // apart along the body, so the 64 lanes of one LDS atomic (spread) or read
  // (interpolate) touch different cells instead of serialising on shared ones.
  const float Ux = (float)s_body[2], Uy = (float)s_body[3];
  int mk[MPT], base_idx[MPT];
  float wx[MPT][4], wy[MPT][4];
  float ubx[MPT], uby[MPT], Fx[MPT], Fy[MPT], ds2[MPT];
  const int nwave = nt >> 6;
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    mk[q] = q * nt + (tid & 63) * nwave + (tid >> 6);
    ubx[q] = uby[q] = Fx[q] = Fy[q] = ds2[q] = 0.f;
    base_idx[q] = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) wx[q][i] = wy[q][i] = 0.f;
    if (mk[q] < ib.n_markers) {
      const int m = mk[q];
      const float rx = (float)(ib.marker_x[m] + s_body[4] - (double)x0);
      const float ry = (float)(ib.marker_y[m] + s_body[5] - (double)y0);
      const Axis<4> ax = axis_stencil<4>(ib.kernel, rx), ay = axis_stencil<4>(ib.kernel, ry);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        wx[q][i] = ax.w[i];
        wy[q][i] = ay.w[i];
      }
      base_idx[q] = (ax.ref - 1) * S + (ay.ref - 1);   // flat window index of stencil point (0, 0)
      ds2[q] = (ib.ds_arr ? ib.ds_arr[m] : ib.ds) * 2.0f;
    }
  }
  __syncthreads();

  // initial interpolation of the window velocity
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    if (mk[q] < ib.n_markers) {
      float sx = 0.f, sy = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = (int)gather_index(base_idx[q] + i * S + j, SS);
          const float w = wx[q][i] * wy[q][j];
          sx += w * ul[gi];
          sy += w * ul[SS + gi];
        }
      ubx[q] = sx;
      uby[q] = sy;
    }
  }
  __syncthreads();   // the window velocity is dead from here; its LDS holds the spread force

  // multi-direct forcing (ib/mdf.py:45-64); pass n_iter spreads the total force
  for (int it = 0; it <= ib.n_iter; ++it) {
    const bool last = it == ib.n_iter;
    float sfx[MPT], sfy[MPT];
#pragma unroll
    for (int q = 0; q < MPT; ++q) {
      if (!last) {
        const float dfx = (Ux - ubx[q]) * ds2[q], dfy = (Uy - uby[q]) * ds2[q];
        Fx[q] += dfx;
        Fy[q] += dfy;
        sfx[q] = dfx;
        sfy[q] = dfy;
      } else {
        sfx[q] = Fx[q];
        sfy[q] = Fy[q];
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
            const int si = (int)scatter_index(base_idx[q] + i * S + j, SS);
            if (si >= 0) {
              const float w = wx[q][i] * wy[q][j];
              atomicAdd(&gl[si], sfx[q] * w);
              atomicAdd(&gl[SS + si], sfy[q] * w);
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
            const int gi = (int)gather_index(base_idx[q] + i * S + j, SS);
            const float w = wx[q][i] * wy[q][j];
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
    for (int k = 0; k < 6; ++k) body_out[k] = out[k];
    if (ib.hist_d && hidx >= 0) {
      ib.hist_d[2 * hidx] = out[4];
      ib.hist_d[2 * hidx + 1] = out[5];
    }
    if (ib.hist_h && hidx >= 0) {
      ib.hist_h[2 * hidx] = h[0];
      ib.hist_h[2 * hidx + 1] = h[1];
    }
  }
}

// Publish "the force window of this step is written" by storing `value` (the
// global step index + 1) to the ready flag with an agent-scope release, so the
// window tiles on other XCDs see the stores.
__device__ __forceinline__ void publish_ready(const VsxIB2D &ib, int value) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_store(&ib.win_origin[2], value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Bounded wait of a window tile (thread 0) until the ready flag reaches
// `target`; on time-out sets the error flag and proceeds.  Then every thread
// of the tile acquires at agent scope.
__device__ __forceinline__ void wait_ready(const VsxIB2D &ib, int target) {
  if (threadIdx.x == 0) {
    int spins = 0;
    while (__hip_atomic_load(&ib.win_origin[2], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kSpinLimit) {
        __hip_atomic_store(&ib.win_origin[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// ------------------------------------------------------------------ fused VIV step
// One launch per step.  Block 0 runs the IB workgroup; blocks 1.. are bulk
// tiles of kTileCells cells (TX rows x TY cells).  Every tile computes the
// step's window origin from the body state of the previous step; the tiles
// that intersect the force window are scheduled last and wait (bounded) for
// the IB workgroup's ready flag, all other tiles start immediately, so the IB
// work overlaps the streaming of the rest of the grid.
constexpr int kTileCells = 1024;

struct Tiles {
  int TX, TY, ntx, nty;
};

__device__ __forceinline__ void tile_of_block(int p, const Tiles &T, int wx0, int wx1, int wy0, int wy1, int &tx,
                                              int &ty, bool &in_win) {
  const int nwx = wx1 - wx0 + 1, nwy = wy1 - wy0 + 1;
  const int nn = T.ntx * T.nty - nwx * nwy;
  in_win = p >= nn;
  if (in_win) {
    const int q = p - nn;
    tx = wx0 + q / nwy;
    ty = wy0 + q % nwy;
    return;
  }
  const int A = wx0 * T.nty, rw = T.nty - nwy, B = nwx * rw;
  if (p < A) {
    tx = p / T.nty;
    ty = p % T.nty;
  } else if (p < A + B) {
    const int q = p - A, r = q % rw;
    tx = wx0 + q / rw;
    ty = r < wy0 ? r : r + nwy;
  } else {
    const int q = p - A - B;
    tx = wx1 + 1 + q / T.nty;
    ty = q % T.nty;
  }
}

// Host-side checks of a VsxIB2D against its grid and the device limits.
inline int validate_ib2d(const VsxLbm *p, const VsxIB2D *ib) {
  if (!ib) return fail(VSX_ERR_VALUE, "null IB parameters");
  VSX_REQUIRE(ib->n_markers > 0 && ib->n_markers <= kIbThreads * kMarkersPerThread, "marker count %d out of range (1..%d)",
              ib->n_markers, kIbThreads * kMarkersPerThread);
  VSX_REQUIRE(ib->win >= 4 && ib->win <= p->nx && ib->win <= p->ny, "IB window %d does not fit the grid", ib->win);
  VSX_REQUIRE(ib->kernel >= VSX_KERNEL_PESKIN_2PT && ib->kernel <= VSX_KERNEL_COSINE_4PT, "unknown kernel %d", ib->kernel);
  VSX_REQUIRE(ib->n_iter >= 0, "negative n_iter");
  VSX_REQUIRE(ib->step0 >= 0, "negative step0");
  VSX_REQUIRE(ib->marker_x && ib->marker_y && ib->body && ib->g_win && ib->win_origin, "null IB buffer");
  VSX_REQUIRE(ib->mass + ib->gamma * ib->dt * ib->damping + ib->beta * ib->dt * ib->dt * ib->stiffness != 0.0,
              "singular Newmark system");
  int dev = 0, max_lds = 0;
  VSX_HIP(hipGetDevice(&dev));
  VSX_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t need = sizeof(float) * 2 * (size_t)ib->win * ib->win + 1024;
  VSX_REQUIRE(need <= (size_t)max_lds, "IB window %d needs %zu B of LDS, device allows %d", ib->win, need, max_lds);
  return VSX_OK;
}

}  // namespace vsx
