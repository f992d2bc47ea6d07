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
// LDS of the IB workgroup in 4-byte words (S = window side, M = markers):
//   ug    [2][S*S]   window velocity, then the spread force
//   with BINS (one marker per thread), from a 16-byte aligned offset:
//   rec   [M][12]    base, -, wx0..3, wy0..3, fx, fy of the markers in bin order
//                    (base: flat window index of stencil point (0, 0); fx, fy:
//                    the force being spread)
//   bptr  [S*S + 1]  markers binned by base (CSR: after placement bptr[b] = end of bin b)
//   blist [M]        markers by bin (placement order)
//   pos   [M]        bin-ordered position of each marker
__host__ __device__ inline int ib_rec_offset(int S) { return (2 * S * S + 3) & ~3; }
__host__ __device__ inline int ib_lds_words(int S, int M, bool bins) {
  return bins ? ib_rec_offset(S) + 14 * M + S * S + 1 : 2 * S * S;
}

// Exclusive prefix sum of a[0..n) in place by the whole workgroup.
__device__ __forceinline__ void block_exclusive_scan(int *a, int n, int *s_wsum) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int per = (n + nt - 1) / nt;
  const int b0 = min(tid * per, n), b1 = min(b0 + per, n);
  int sum = 0;
  for (int b = b0; b < b1; ++b) sum += a[b];
  int v = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if ((tid & 63) >= o) v += t;
  }
  if ((tid & 63) == 63) s_wsum[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int w = 0; w < (nt + 63) / 64; ++w) {
      const int x = s_wsum[w];
      s_wsum[w] = run;
      run += x;
    }
  }
  __syncthreads();
  int run = s_wsum[tid >> 6] + v - sum;
  for (int b = b0; b < b1; ++b) {
    const int x = a[b];
    a[b] = run;
    run += x;
  }
  __syncthreads();
}

// Spread of the binned markers onto the window as a gather: window cell k
// receives from the markers whose base lies in the 4 runs of 4 bins
// k - a S - 3 .. k - a S (a = 0..3), in ascending bin and marker order, with
// weight wx[a] * wy[k - a S - base].  Every cell is written exactly once.
__device__ __forceinline__ void bin_spread(float *ug, const float *rec, const int *bptr, int S) {
  const int SS = S * S;
  for (int k = threadIdx.x; k < SS; k += blockDim.x) {
    int e0[4], e1[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int hi = k - a * S, lo = hi - 3;
      e0[a] = hi >= 0 && lo > 0 ? bptr[lo - 1] : 0;
      e1[a] = hi >= 0 ? bptr[hi] : 0;
    }
    float gx = 0.f, gy = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int hi = k - a * S;
#pragma unroll 2
      for (int e = e0[a]; e < e1[a]; ++e) {
        const float4 q0 = *reinterpret_cast<const float4 *>(rec + 12 * e);
        const float4 q1 = *reinterpret_cast<const float4 *>(rec + 12 * e + 4);
        const float4 q2 = *reinterpret_cast<const float4 *>(rec + 12 * e + 8);
        const int b = hi - __float_as_int(q0.x);
        const float wxa = a == 0 ? q0.z : a == 1 ? q0.w : a == 2 ? q1.x : q1.y;
        float wyb = q1.z;   // flat selects, no branches
        wyb = b == 1 ? q1.w : wyb;
        wyb = b == 2 ? q2.x : wyb;
        wyb = b == 3 ? q2.y : wyb;
        const float w = wxa * wyb;
        gx += q2.z * w;
        gy += q2.w * w;
      }
    }
    ug[k] = gx;
    ug[SS + k] = gy;
  }
}

// Bounded wait of thread 0 until *flag >= target (agent-scope acquire); on
// time-out the error flag ib.win_origin[3] is set and the wait ends.
__device__ __forceinline__ void spin_until(const VsxIB2D &ib, int32_t *flag, int target) {
  int spins = 0;
  while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > kSpinLimit) {
      __hip_atomic_store(&ib.win_origin[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
  }
}

// One workgroup per body (examples/2d/vortex_induced_vibration.py:99-137):
//   window u     macroscopic velocity of the post-streaming state (pull + walls);
//                U_GLOBAL: published in ib.g_win by the n_wtiles tiles that
//                cover the window (counter ib.win_origin[4 + (gstep & 1)])
//   stencil      ib/stencil.py:7-51 at marker coordinates (x_m + d) - origin (4x4)
//   MDF          ib/mdf.py:35-64: interpolate, n_iter x (spread dF, interpolate
//                0.5 * spread), final spread of the total force.  BINS: spread is
//                a gather over the markers binned by stencil base cell (see
//                bin_spread): no atomics, deterministic.
//                Markers whose stencil leaves the window, and the path with
//                several markers per thread, use LDS atomics with the scatter
//                rule of ib/stencil.py:106-110
//   body force   h = sum_m (-F_m) + a * area in FP64 (wave shuffles + LDS), dyn.py:126-136
//   structure    Newmark-beta in FP64 on lane 0, dyn.py:5-51
// Writes the force window to ib.g_win and the new body state to body_out.
template <class SRC, bool PULL, int MPT, int UW, int TAG, bool U_GLOBAL>
__device__ __forceinline__ void ib_block(const StepCtx &c, const VsxIB2D &ib, const typename SRC::T *__restrict__ src,
                                         int hidx, int gstep, int n_wtiles, const double *body_in, double *body_out,
                                         float *lds) {
  const int S = ib.win, SS = S * S, M = ib.n_markers;
  constexpr bool BINS = MPT == 1;
  float *ug = lds;
  float *rec = lds + ib_rec_offset(S);
  int *bptr = reinterpret_cast<int *>(rec + 12 * M);
  int *blist = bptr + SS + 1;
  int *pos = blist + M;
  __shared__ double s_body[6];
  __shared__ int s_org[4];
  __shared__ int s_irregular;
  __shared__ int s_wsum[kIbThreads / 64];
  __shared__ double s_red[2][kIbThreads / 64];
  const int tid = threadIdx.x, nt = blockDim.x;

  if (tid < 6) s_body[tid] = body_in[tid];
  if (tid == 0) s_irregular = 0;
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

  // Markers first (they need no field data).  Marker q of this thread:
  // consecutive lanes of a wavefront hold markers nt/64 apart along the body,
  // so the lanes of one LDS access touch different cells.
  const float Ux = (float)s_body[2], Uy = (float)s_body[3];
  int mk[MPT], base_idx[MPT], bx[MPT], by[MPT];
  bool regular[MPT];
  float wx[MPT][4], wy[MPT][4];
  float ubx[MPT], uby[MPT], Fx[MPT], Fy[MPT], ds2[MPT];
  const int nwave = nt >> 6;
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    mk[q] = q * nt + (tid & 63) * nwave + (tid >> 6);
    ubx[q] = uby[q] = Fx[q] = Fy[q] = ds2[q] = 0.f;
    base_idx[q] = bx[q] = by[q] = 0;
    regular[q] = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) wx[q][i] = wy[q][i] = 0.f;
    if (mk[q] < M) {
      const int m = mk[q];
      const float rx = (float)(ib.marker_x[m] + s_body[4] - (double)x0);
      const float ry = (float)(ib.marker_y[m] + s_body[5] - (double)y0);
      const Axis<4> ax = axis_stencil<4>(ib.kernel, rx), ay = axis_stencil<4>(ib.kernel, ry);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        wx[q][i] = ax.w[i];
        wy[q][i] = ay.w[i];
      }
      bx[q] = ax.ref - 1;   // window row / column of stencil point (0, 0)
      by[q] = ay.ref - 1;
      base_idx[q] = bx[q] * S + by[q];   // its flat window index
      ds2[q] = (ib.ds_arr ? ib.ds_arr[m] : ib.ds) * 2.0f;
      regular[q] = BINS && bx[q] >= 0 && bx[q] <= S - 4 && by[q] >= 0 && by[q] <= S - 4;
      if (!regular[q]) s_irregular = 1;
    }
  }
  if constexpr (BINS) {
    // regular markers binned by base cell: count, scan, place; a marker's
    // position = start of its bin + its rank by marker index within the bin
    for (int k = tid; k <= SS; k += nt) bptr[k] = 0;
    __syncthreads();
    if (mk[0] < M && regular[0]) atomicAdd(&bptr[base_idx[0]], 1);
    __syncthreads();
    block_exclusive_scan(bptr, SS, s_wsum);
    if (mk[0] < M && regular[0]) blist[atomicAdd(&bptr[base_idx[0]], 1)] = mk[0];
    __syncthreads();
    if (mk[0] < M && regular[0]) {
      const int s0 = base_idx[0] > 0 ? bptr[base_idx[0] - 1] : 0, e0 = bptr[base_idx[0]];
      int rank = 0;
      for (int i = s0; i < e0; ++i) rank += blist[i] < mk[0];
      pos[mk[0]] = s0 + rank;
    }
    __syncthreads();
    if (mk[0] < M && regular[0]) {
      const int e = pos[mk[0]];
      float *r = rec + 12 * e;
      r[0] = __int_as_float(base_idx[0]);
      r[1] = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        r[2 + i] = wx[0][i];
        r[6 + i] = wy[0][i];
      }
    }
  }

  // window velocity
  if constexpr (U_GLOBAL) {
    if (tid == 0) {
      int32_t *cnt = &ib.win_origin[4 + (gstep & 1)];
      spin_until(ib, cnt, n_wtiles);
      __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);   // reused two steps later
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int k = tid; k < 2 * SS; k += nt) ug[k] = ib.g_win[k];
  } else {
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
            ug[k] = uw.x;
            ug[SS + k] = uw.y;
          } else {
            float rho, u[3];
            macroscopic<D2Q9>(f[j], rho, u);
            ug[k] = u[0];
            ug[SS + k] = u[1];
          }
        }
      }
    }
  }
  __syncthreads();

  // initial interpolation of the window velocity
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    if (mk[q] < M) {
      float sx = 0.f, sy = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = (int)gather_index(base_idx[q] + i * S + j, SS);
          const float w = wx[q][i] * wy[q][j];
          sx += w * ug[gi];
          sy += w * ug[SS + gi];
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
    if constexpr (BINS) {
      if (mk[0] < M && regular[0]) {
        float *r = rec + 12 * pos[mk[0]];
        r[10] = sfx[0];
        r[11] = sfy[0];
      }
      __syncthreads();
      bin_spread(ug, rec, bptr, S);
      __syncthreads();
    } else {
      for (int k = tid; k < 2 * SS; k += nt) ug[k] = 0.f;
      __syncthreads();
    }
    if (s_irregular) {
#pragma unroll
      for (int q = 0; q < MPT; ++q) {
        if (mk[q] < M && !regular[q]) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int si = (int)scatter_index(base_idx[q] + i * S + j, SS);
              if (si >= 0) {
                const float w = wx[q][i] * wy[q][j];
                atomicAdd(&ug[si], sfx[q] * w);
                atomicAdd(&ug[SS + si], sfy[q] * w);
              }
            }
        }
      }
      __syncthreads();
    }
    if (last) break;
#pragma unroll
    for (int q = 0; q < MPT; ++q) {
      if (mk[q] < M) {
        float sx = 0.f, sy = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int gi = (int)gather_index(base_idx[q] + i * S + j, SS);
            const float w = wx[q][i] * wy[q][j];
            sx += w * (ug[gi] * 0.5f);
            sy += w * (ug[SS + gi] * 0.5f);
          }
        ubx[q] += sx;
        uby[q] += sy;
      }
    }
    __syncthreads();
  }

  // force window -> global
  for (int k = tid; k < 2 * SS; k += nt) ib.g_win[k] = ug[k];

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

// Wait of a window tile until the ready flag reaches `target` (bounded, see
// spin_until); then every thread of the tile acquires at agent scope.
__device__ __forceinline__ void wait_ready(const VsxIB2D &ib, int target) {
  if (threadIdx.x == 0) spin_until(ib, &ib.win_origin[2], target);
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Window tile: publish the post-streaming velocity of its cells that lie in
// the force window (into ib.g_win, read by the IB workgroup), then count the
// tile in ib.win_origin[4 + (gstep & 1)] with an agent-scope release.
__device__ __forceinline__ void publish_window_u(const VsxIB2D &ib, int gstep, bool valid, const int (&co)[3],
                                                 const int (&org)[3], const float (&f)[9]) {
  const int S = ib.win;
  const int r0 = co[0] - org[0], r1 = co[1] - org[1];
  if (valid && r0 >= 0 && r0 < S && r1 >= 0 && r1 < S) {
    float rho, u[3];
    macroscopic<D2Q9>(f, rho, u);
    ib.g_win[r0 * S + r1] = u[0];
    ib.g_win[S * S + r0 * S + r1] = u[1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(&ib.win_origin[4 + (gstep & 1)], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------------ fused VIV step
// One launch per step.  Block 0 runs the IB workgroup; blocks 1.. are bulk
// tiles of kTileCells cells (TX rows x TY cells).  Every tile computes the
// step's window origin from the body state before the step.  The tiles that
// intersect the force window come first: they pull their cells, hand the
// window velocity to the IB workgroup, wait (bounded) for its force window and
// then collide; all other tiles stream and collide at once, so the IB work
// overlaps the streaming of the rest of the grid.
constexpr int kTileCells = 1024;

struct Tiles {
  int TX, TY, ntx, nty;
};

// Window tile rectangle [wx0, wx1] x [wy0, wy1] (tile coordinates) of a window at org.
__host__ __device__ inline void window_tiles(const Tiles &T, int org0, int org1, int S, int &wx0, int &wx1, int &wy0,
                                             int &wy1) {
  wx0 = org0 / T.TX;
  wx1 = (org0 + S - 1) / T.TX;
  wy0 = org1 / T.TY;
  wy1 = (org1 + S - 1) / T.TY;
}

// Block p (0-based among the tile blocks) -> tile (tx, ty): the window
// rectangle's tiles first, then the others in row-major order.
__host__ __device__ inline void tile_of_block(int p, const Tiles &T, int wx0, int wx1, int wy0, int wy1, int &tx,
                                              int &ty, bool &in_win) {
  const int nwx = wx1 - wx0 + 1, nwy = wy1 - wy0 + 1;
  in_win = p < nwx * nwy;
  if (in_win) {
    tx = wx0 + p / nwy;
    ty = wy0 + p % nwy;
    return;
  }
  const int q = p - nwx * nwy;
  const int A = wx0 * T.nty, rw = T.nty - nwy, B = nwx * rw;
  if (q < A) {
    tx = q / T.nty;
    ty = q % T.nty;
  } else if (q < A + B) {
    const int r = (q - A) % rw;
    tx = wx0 + (q - A) / rw;
    ty = r < wy0 ? r : r + nwy;
  } else {
    tx = wx1 + 1 + (q - A - B) / T.nty;
    ty = (q - A - B) % T.nty;
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
  const size_t need = 4 * (size_t)ib_lds_words(ib->win, ib->n_markers, ib->n_markers <= kIbThreads) + 2048;
  VSX_REQUIRE(need <= (size_t)max_lds, "IB window %d needs %zu B of LDS, device allows %d", ib->win, need, max_lds);
  return VSX_OK;
}

}  // namespace vsx
