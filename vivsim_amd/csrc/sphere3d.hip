// 3D IB-LBM step past a fixed body (flow past a sphere), D3Q19.
//
// Per step (examples/3d/flow_past_sphere.py:139-171, ib/mdf.py:10-64):
//   k_win_u3d      window velocity of the post-streaming state          (cells)
//   k_mark3d(0)    ub = interp(u); dF = (U - ub) 2 ds; F = dF    (a wavefront per marker)
//   k_spread3d     A = spread(dF)                                        (cells)
//   k_mark3d(k)    ub += interp(0.5 A); dF = ...; F += dF                (markers)
//   ...            (n_iter - 1 times)
//   k_spread3d     g_win = spread(F); total marker force -F in FP64      (cells)
//   bulk           collision + forcing with g_win, streaming, walls      (cells)
// The body is fixed, so its stencil (computed once, relative to the window
// origin) is transposed once: spread becomes a gather over the markers that
// touch each window cell, in ascending marker order -- deterministic and free
// of atomics.  The last spread of the reference loop (its interpolation only
// updates ub, which is not used afterwards) is not evaluated.
//
// Pipelined form (ib.pipelined): the IB kernels of step t + 1 run on a second
// stream while the bulk pass of step t streams.  The bulk blocks that cover
// the window +- 1 cell write their new populations (as read back from
// storage) into the FP32 hand-off buffer ib.halo (two halves by step parity)
// with st_agent and count themselves; the last one raises the hand-off flag,
// which the first IB kernel of the next step waits for before it pulls the
// window velocity from the buffer.  The last IB kernel writes the force
// window with st_agent and a one-thread kernel raises the ready flag; the bulk
// blocks that intersect the window wait for it and read the force window with
// ld_agent.  Flags: work words [4] ready, [5] hand-off counter, [6] hand-off
// flag, [7] error (1: a bounded wait ran out).
#include "ib_common.hpp"
#include "step_impl.hpp"

namespace vsx {

constexpr int kS3 = 64;   // stencil points per marker (4 x 4 x 4)

struct IB3DDev {
  int32_t *origin;   // [4] window origin (read by the bulk kernel)
  int32_t *flags;    // [4] pipelined form: ready, hand-off counter, hand-off flag, error
  float *u_win, *acc, *ub, *F, *dF, *g_win;
  const float *W;
  const int32_t *I, *cell_ptr, *cell_slot;
  const float *u_target, *ds_arr;
  float ds;
  int32_t M, n_iter;
  int32_t S[3];
  int32_t org[3];
  int64_t cells;
};

int64_t work_floats(int32_t m, int64_t cells) { return 8 + 2 * 3 * cells + 3 * 3 * (int64_t)m; }

IB3DDev make_dev(const VsxIB3D &ib) {
  IB3DDev d;
  d.cells = (int64_t)ib.win[0] * ib.win[1] * ib.win[2];
  float *w = ib.work;
  d.origin = reinterpret_cast<int32_t *>(w);
  d.flags = reinterpret_cast<int32_t *>(w + 4);
  d.u_win = w + 8;
  d.acc = d.u_win + 3 * d.cells;
  d.ub = d.acc + 3 * d.cells;
  d.F = d.ub + 3 * (int64_t)ib.n_markers;
  d.dF = d.F + 3 * (int64_t)ib.n_markers;
  d.g_win = ib.g_win;
  d.W = ib.W;
  d.I = ib.I;
  d.cell_ptr = ib.cell_ptr;
  d.cell_slot = ib.cell_slot;
  d.u_target = ib.u_target;
  d.ds_arr = ib.ds_arr;
  d.ds = ib.ds;
  d.M = ib.n_markers;
  d.n_iter = ib.n_iter;
  for (int a = 0; a < 3; ++a) {
    d.S[a] = ib.win[a];
    d.org[a] = ib.origin[a];
  }
  return d;
}

// Window velocity (3 components) of the post-streaming state: pulled from the
// post-collision state with the walls applied (PULL), or read in place (the
// first step of a run reads the caller's post-streaming state).
template <class SRC, bool PULL>
__global__ __launch_bounds__(256) void k_win_u3d(StepCtx c_arg, IB3DDev d, const typename SRC::T *__restrict__ src) {
  const StepCtx &c = kernel_ctx();
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k == 0) {
    for (int a = 0; a < 3; ++a) d.origin[a] = d.org[a];
  }
  if (k >= d.cells) return;
  const int r2 = (int)(k % d.S[2]);
  const int r1 = (int)((k / d.S[2]) % d.S[1]);
  const int r0 = (int)(k / ((int64_t)d.S[2] * d.S[1]));
  const int co[3] = {d.org[0] + r0, d.org[1] + r1, d.org[2] + r2};
  float f[19], rho, u[3];
  if constexpr (PULL) cell_state<D3Q19, SRC>(c, src, co, f);
  else local<D3Q19, SRC>(c, src, cell_of<D3Q19>(c, co[0], co[1], co[2]), f);
  macroscopic<D3Q19>(f, rho, u);
  for (int a = 0; a < 3; ++a) d.u_win[a * d.cells + k] = u[a];
}

// Marker phase `it` of the multi-direct forcing loop (ib/mdf.py:45-57).
// it = 0: ub = interp(u_win), F = 0; it > 0: ub += interp(0.5 * acc).
// Then, when it < n_iter: dF = (U - ub) * 2 ds, F += dF.
// One wavefront per marker, one lane per stencil point (kS3 = 64): the
// stencil and window loads of a marker are issued at once and summed by a
// fixed butterfly over the lanes (deterministic).
static_assert(kS3 == 64, "one lane per stencil point of a wavefront");
__global__ __launch_bounds__(256) void k_mark3d(IB3DDev d, int it) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (m >= d.M) return;   // whole wavefronts
  const float *grid = it == 0 ? d.u_win : d.acc;
  const int64_t gi = gather_index(d.I[(int64_t)m * kS3 + lane], d.cells);
  const float w = d.W[(int64_t)m * kS3 + lane];
  float v0 = grid[gi], v1 = grid[d.cells + gi], v2 = grid[2 * d.cells + gi];
  if (it != 0) {
    v0 *= 0.5f;
    v1 *= 0.5f;
    v2 *= 0.5f;
  }
  float s0 = w * v0, s1 = w * v1, s2 = w * v2;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o, 64);
    s1 += __shfl_xor(s1, o, 64);
    s2 += __shfl_xor(s2, o, 64);
  }
  if (lane != 0) return;
  float ub[3], F[3];
  if (it == 0) {
    ub[0] = s0, ub[1] = s1, ub[2] = s2;
    F[0] = F[1] = F[2] = 0.f;
  } else {
    ub[0] = d.ub[3 * m] + s0, ub[1] = d.ub[3 * m + 1] + s1, ub[2] = d.ub[3 * m + 2] + s2;
    F[0] = d.F[3 * m], F[1] = d.F[3 * m + 1], F[2] = d.F[3 * m + 2];
  }
  if (it < d.n_iter) {
    const float ds2 = (d.ds_arr ? d.ds_arr[m] : d.ds) * 2.0f;
    for (int a = 0; a < 3; ++a) {
      const float df = (d.u_target[3 * m + a] - ub[a]) * ds2;
      F[a] += df;
      d.dF[3 * m + a] = df;
    }
  }
  for (int a = 0; a < 3; ++a) {
    d.ub[3 * m + a] = ub[a];
    d.F[3 * m + a] = F[a];
  }
}

__device__ __forceinline__ double wave_sum3(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// out[a][k] = sum over the stencil slots that land on window cell k of
// mf[m][a] * W[slot] (ib/stencil.py:81-110 as a gather).  With hist, block 0
// also writes the total marker force sum_m (-F_m) in FP64 (dyn.py:126-136).
__global__ __launch_bounds__(256) void k_spread3d(IB3DDev d, const float *__restrict__ mf, float *__restrict__ out,
                                                  double *hist, bool agent) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < d.cells) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    const int e1 = d.cell_ptr[k + 1];
    int e = d.cell_ptr[k];
    for (; e + 4 <= e1; e += 4) {   // four slots' loads in flight, summed in slot order
      int slot[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) slot[u] = d.cell_slot[e + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int m = slot[u] / kS3;
        const float w = d.W[slot[u]];
        s0 += mf[3 * m] * w;
        s1 += mf[3 * m + 1] * w;
        s2 += mf[3 * m + 2] * w;
      }
    }
    for (; e < e1; ++e) {
      const int slot = d.cell_slot[e];
      const int m = slot / kS3;
      const float w = d.W[slot];
      s0 += mf[3 * m] * w;
      s1 += mf[3 * m + 1] * w;
      s2 += mf[3 * m + 2] * w;
    }
    if (agent) {   // read by the bulk pass of another stream
      st_agent(&out[k], s0);
      st_agent(&out[d.cells + k], s1);
      st_agent(&out[2 * d.cells + k], s2);
    } else {
      out[k] = s0;
      out[d.cells + k] = s1;
      out[2 * d.cells + k] = s2;
    }
  }
  if (hist && blockIdx.x == 0) {
    __shared__ double red[3][4];
    double h[3] = {0.0, 0.0, 0.0};
    for (int m = threadIdx.x; m < d.M; m += blockDim.x)
      for (int a = 0; a < 3; ++a) h[a] -= (double)d.F[3 * m + a];
    for (int a = 0; a < 3; ++a) h[a] = wave_sum3(h[a]);
    if ((threadIdx.x & 63) == 0)
      for (int a = 0; a < 3; ++a) red[a][threadIdx.x >> 6] = h[a];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int a = 0; a < 3; ++a) hist[a] = red[a][0] + red[a][1] + red[a][2] + red[a][3];
  }
}

// ------------------------------------------------------------------ pipelined form
// Bounded wait of thread 0 until flags[k] >= target; on time-out flags[3] = 1.
__device__ __forceinline__ void spin3d(int32_t *flags, int k, int target) {
  int spins = 0;
  while (__hip_atomic_load(&flags[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > (1 << 22)) {
      __hip_atomic_store(&flags[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
  }
}

// Window velocity of step `step` pulled from the hand-off buffer written by the
// bulk pass of step - 1 (the box of the window +- 1 cell, origin org - 1).
__global__ __launch_bounds__(256) void k_win_u3d_halo(IB3DDev d, const float *__restrict__ halo, int step) {
  if (threadIdx.x == 0) spin3d(d.flags, 2, step);
  __syncthreads();
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= d.cells) return;
  const int E0 = d.S[0] + 2, E1 = d.S[1] + 2, E2 = d.S[2] + 2;
  const int64_t EE = (int64_t)E0 * E1 * E2;
  const float *h = halo + (size_t)((step - 1) & 1) * 19 * EE;
  const int r2 = (int)(k % d.S[2]) + 1;
  const int r1 = (int)((k / d.S[2]) % d.S[1]) + 1;
  const int r0 = (int)(k / ((int64_t)d.S[2] * d.S[1])) + 1;
  float f[19], rho, u[3];
#pragma unroll
  for (int q = 0; q < 19; ++q)
    f[q] = ld_agent(&h[q * EE + ((int64_t)(r0 - D3Q19::cx[q]) * E1 + (r1 - D3Q19::cy[q])) * E2 + (r2 - D3Q19::cz[q])]);
  macroscopic<D3Q19>(f, rho, u);
  for (int a = 0; a < 3; ++a) d.u_win[a * d.cells + k] = u[a];
}

// ready flag of step `step`: the force window (written with st_agent by the
// last spread of this stream) is complete
__global__ void k_publish3d(IB3DDev d, int step) {
  __hip_atomic_store(&d.flags[0], step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// body force of a cell in the window of the pipelined form (the force window
// comes from the other stream: agent-coherent loads)
__device__ __forceinline__ void window_force3d(const IB3DDev &d, const int (&co)[3], float (&g)[3]) {
  const int r0 = co[0] - d.org[0], r1 = co[1] - d.org[1], r2 = co[2] - d.org[2];
  g[0] = g[1] = g[2] = 0.f;
  if (r0 >= 0 && r0 < d.S[0] && r1 >= 0 && r1 < d.S[1] && r2 >= 0 && r2 < d.S[2]) {
    const int64_t k = ((int64_t)r0 * d.S[1] + r1) * d.S[2] + r2;
#pragma unroll
    for (int a = 0; a < 3; ++a) g[a] = ld_agent(&d.g_win[a * d.cells + k]);
  }
}

// Bulk pass of the pipelined form: the 3D pull-scheme step (k_bulk) where the
// blocks that intersect the window wait for the ready flag of this step, and
// the blocks that touch the window +- 1 cell hand their cells to the next
// step's IB (except in the last step of a run, LAST).  Grid as k_bulk: x =
// blockIdx.z, y = blockIdx.y, z from blockIdx.x; PHASE COLLIDE reads the
// caller's post-streaming field in place (no walls).
template <int COLL, int PHASE, class SRC, class DST>
__global__ __launch_bounds__(256) void k_bulk3d_pipe(StepCtx c_arg, IB3DDev d, const typename SRC::T *__restrict__ src,
                                                     typename DST::T *__restrict__ dst, float *__restrict__ halo,
                                                     int step, int n_halo_blocks, bool last) {
  const StepCtx &c = kernel_ctx();
  const int z0 = blockIdx.x * blockDim.x, z1 = z0 + blockDim.x;
  const int co[3] = {(int)blockIdx.z, (int)blockIdx.y, z0 + (int)threadIdx.x};
  const bool valid = co[2] < c.dims[2];
  auto in_box = [&](int lo_pad, int hi_pad) {
    return co[0] >= d.org[0] - lo_pad && co[0] < d.org[0] + d.S[0] + hi_pad && co[1] >= d.org[1] - lo_pad &&
           co[1] < d.org[1] + d.S[1] + hi_pad && z0 < d.org[2] + d.S[2] + hi_pad && z1 > d.org[2] - lo_pad;
  };
  const bool win_block = in_box(0, 0), halo_block = in_box(1, 1);
  if (win_block) {
    if (threadIdx.x == 0) spin3d(d.flags, 0, step + 1);
    __syncthreads();
  }
  float f[19];
  const uint32_t cell = valid ? cell_of<D3Q19>(c, co[0], co[1], co[2]) : 0;
  if (valid) {
    if constexpr (PHASE == VSX_PHASE_COLLIDE) {
      local<D3Q19, SRC>(c, src, cell, f);
    } else {
      const int cls = c.p.n_walls > 0 ? wall_class<D3Q19>(c, co) : -1;
      if (cls >= 0) single_wall_state<D3Q19, SRC>(c, src, co, cls, f);
      else pull<D3Q19, SRC>(c, src, co[0], co[1], co[2], f);
    }
    float rho, u[3], g[3];
    macroscopic<D3Q19>(f, rho, u);
    window_force3d(d, co, g);
#pragma unroll
    for (int a = 0; a < 3; ++a) g[a] += c.p.g_uniform[a] + (c.p.g ? c.p.g[(size_t)a * c.n + cell] : 0.f);
    collide_force<D3Q19, COLL>(f, rho, u, g, c.p.forcing, c.p.omega, c.p.mrt_rates);
    store_cell<D3Q19, DST>(c, dst, cell, f);
  }
  if (halo_block && !last) {
    const int E0 = d.S[0] + 2, E1 = d.S[1] + 2, E2 = d.S[2] + 2;
    const int64_t EE = (int64_t)E0 * E1 * E2;
    const int a = co[0] - (d.org[0] - 1), b = co[1] - (d.org[1] - 1), e = co[2] - (d.org[2] - 1);
    float *h = halo + (size_t)(step & 1) * 19 * EE;
    if (valid && a >= 0 && a < E0 && b >= 0 && b < E1 && e >= 0 && e < E2) {
      const int64_t k = ((int64_t)a * E1 + b) * E2 + e;
#pragma unroll
      for (int q = 0; q < 19; ++q) st_agent(&h[q * EE + k], DST::rt(f[q], c.shift[q]));
    }
    wait_mem();
    __syncthreads();
    if (threadIdx.x == 0) {
      const int before = __hip_atomic_fetch_add(&d.flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (before == n_halo_blocks - 1) {
        __hip_atomic_store(&d.flags[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&d.flags[2], step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// ------------------------------------------------------------------ stencil transpose

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_count(const int32_t *I, int64_t n_slots, int64_t n_grid, int32_t *cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_slots; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = scatter_index(I[i], n_grid);
    if (k >= 0) atomicAdd(&cnt[k], 1);
  }
}

// exclusive scan of cnt[0..n) into ptr[0..n] by one workgroup of 1024 threads
// (each thread scans a contiguous chunk; chunk totals are scanned in LDS)
__global__ __launch_bounds__(1024) void k_scan(const int32_t *cnt, int64_t n, int32_t *ptr) {
  __shared__ int32_t tot[1024];
  const int t = threadIdx.x;
  const int64_t chunk = (n + 1023) / 1024;
  const int64_t b = t * chunk, e = b + chunk < n ? b + chunk : n;
  int32_t s = 0;
  for (int64_t i = b; i < e; ++i) s += cnt[i];
  tot[t] = s;
  __syncthreads();
  if (t == 0) {
    int32_t run = 0;
    for (int j = 0; j < 1024; ++j) {
      const int32_t v = tot[j];
      tot[j] = run;
      run += v;
    }
  }
  __syncthreads();
  int32_t run = tot[t];
  for (int64_t i = b; i < e; ++i) {
    ptr[i] = run;
    run += cnt[i];
  }
  if (t == 1023) ptr[n] = run;
}

__global__ void k_place(const int32_t *I, int64_t n_slots, int64_t n_grid, const int32_t *ptr, int32_t *cursor,
                        int32_t *slot_out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_slots; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = scatter_index(I[i], n_grid);
    if (k >= 0) slot_out[ptr[k] + atomicAdd(&cursor[k], 1)] = (int32_t)i;
  }
}

// each cell's list in ascending slot order (insertion sort: lists are short)
__global__ void k_sort_lists(const int32_t *ptr, int64_t n_grid, int32_t *slot) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n_grid; k += (int64_t)gridDim.x * blockDim.x) {
    const int b = ptr[k], e = ptr[k + 1];
    for (int i = b + 1; i < e; ++i) {
      const int32_t v = slot[i];
      int j = i - 1;
      while (j >= b && slot[j] > v) {
        slot[j + 1] = slot[j];
        --j;
      }
      slot[j + 1] = v;
    }
  }
}

// ------------------------------------------------------------------ host side

template <class SRC, bool PULL>
void launch_win_u(const StepCtx &c, const IB3DDev &d, const void *src, hipStream_t s) {
  hipLaunchKernelGGL((k_win_u3d<SRC, PULL>), dim3((unsigned)((d.cells + 255) / 256)), dim3(256), 0, s, c, d,
                     static_cast<const typename SRC::T *>(src));
}

void win_u_any(const StepCtx &c, const IB3DDev &d, const void *src, bool first, hipStream_t s) {
  if (first) return launch_win_u<Store<VSX_STORE_F32>, false>(c, d, src, s);
  switch (c.p.storage) {
    case VSX_STORE_F32: return launch_win_u<Store<VSX_STORE_F32>, true>(c, d, src, s);
    case VSX_STORE_F16S: return launch_win_u<Store<VSX_STORE_F16S>, true>(c, d, src, s);
    case VSX_STORE_BF16S: return launch_win_u<Store<VSX_STORE_BF16S>, true>(c, d, src, s);
    default: return launch_win_u<Store<VSX_STORE_E3M12S>, true>(c, d, src, s);
  }
}

// The IB of one step.  pipe_step >= 0 (pipelined form): the window velocity of a
// step after the first comes from the hand-off buffer, the force window is
// written with st_agent and the ready flag raised for step pipe_step.
int ib3d_phase(const StepCtx &c, const IB3DDev &d, const void *src, bool first, double *hist, hipStream_t s,
               const float *halo = nullptr, int pipe_step = -1) {
  const unsigned gm = (unsigned)((d.M + 3) / 4), gc = (unsigned)((d.cells + 255) / 256);   // k_mark3d: 4 markers / block
  if (pipe_step > 0 && !first) hipLaunchKernelGGL(k_win_u3d_halo, dim3(gc), dim3(256), 0, s, d, halo, pipe_step);
  else win_u_any(c, d, src, first, s);
  hipLaunchKernelGGL(k_mark3d, dim3(gm), dim3(256), 0, s, d, 0);
  for (int it = 1; it < d.n_iter; ++it) {
    hipLaunchKernelGGL(k_spread3d, dim3(gc), dim3(256), 0, s, d, (const float *)d.dF, d.acc, (double *)nullptr, false);
    hipLaunchKernelGGL(k_mark3d, dim3(gm), dim3(256), 0, s, d, it);
  }
  hipLaunchKernelGGL(k_spread3d, dim3(gc), dim3(256), 0, s, d, (const float *)d.F, d.g_win, hist, pipe_step >= 0);
  if (pipe_step >= 0) hipLaunchKernelGGL(k_publish3d, dim3(1), dim3(1), 0, s, d, pipe_step);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

template <int COLL, int PHASE, class SRC, class DST>
void launch_bulk3d_pipe(const StepCtx &c, const IB3DDev &d, const void *src, void *dst, float *halo, int step,
                        int n_halo, bool last, hipStream_t s) {
  dim3 grid, block;
  bulk_grid<D3Q19>(c, grid, block);
  timer_begin(s);
  hipLaunchKernelGGL((k_bulk3d_pipe<COLL, PHASE, SRC, DST>), grid, block, 0, s, c, d,
                     static_cast<const typename SRC::T *>(src), static_cast<typename DST::T *>(dst), halo, step,
                     n_halo, last);
  timer_end(s);
}

template <int PHASE, class SRC, class DST>
void bulk3d_pipe_coll(const StepCtx &c, const IB3DDev &d, const void *src, void *dst, float *halo, int step,
                      int n_halo, bool last, hipStream_t s) {
  switch (c.p.collision) {
    case VSX_COLL_BGK: return launch_bulk3d_pipe<VSX_COLL_BGK, PHASE, SRC, DST>(c, d, src, dst, halo, step, n_halo, last, s);
    case VSX_COLL_MRT: return launch_bulk3d_pipe<VSX_COLL_MRT, PHASE, SRC, DST>(c, d, src, dst, halo, step, n_halo, last, s);
    case VSX_COLL_KBC: return launch_bulk3d_pipe<VSX_COLL_KBC, PHASE, SRC, DST>(c, d, src, dst, halo, step, n_halo, last, s);
    default: return launch_bulk3d_pipe<VSX_COLL_REG, PHASE, SRC, DST>(c, d, src, dst, halo, step, n_halo, last, s);
  }
}

template <class ST>
void bulk3d_pipe_any(const StepCtx &c, const IB3DDev &d, const void *src, void *dst, float *halo, int step,
                     int n_halo, bool first, bool last, hipStream_t s) {
  using F32 = Store<VSX_STORE_F32>;
  if (first) bulk3d_pipe_coll<VSX_PHASE_COLLIDE, F32, ST>(c, d, src, dst, halo, step, n_halo, last, s);
  else bulk3d_pipe_coll<VSX_PHASE_FULL, ST, ST>(c, d, src, dst, halo, step, n_halo, last, s);
}

// blocks of the bulk grid that touch the window +- 1 cell
int64_t halo_blocks(const StepCtx &c, const IB3DDev &d) {
  dim3 grid, block;
  bulk_grid<D3Q19>(c, grid, block);
  const int64_t bd = block.x;
  const int64_t zb0 = (d.org[2] - 1) / bd, zb1 = (d.org[2] + d.S[2]) / bd;
  return (int64_t)(d.S[0] + 2) * (d.S[1] + 2) * (zb1 - zb0 + 1);
}

int validate_ib3d(const VsxLbm *p, const VsxIB3D *ib) {
  if (!ib) return fail(VSX_ERR_VALUE, "null IB parameters");
  VSX_REQUIRE(ib->n_markers > 0, "marker count %d out of range", ib->n_markers);
  VSX_REQUIRE(ib->n_iter >= 0, "negative n_iter");
  const int dims[3] = {p->nx, p->ny, p->nz};
  for (int a = 0; a < 3; ++a)
    VSX_REQUIRE(ib->win[a] >= 1 && ib->origin[a] >= 0 && ib->origin[a] + ib->win[a] <= dims[a],
                "IB window axis %d (origin %d, size %d) not inside the grid (%d)", a, ib->origin[a], ib->win[a], dims[a]);
  VSX_REQUIRE(ib->W && ib->I && ib->cell_ptr && ib->cell_slot && ib->u_target && ib->g_win && ib->work,
              "null IB buffer");
  VSX_REQUIRE(p->forcing != VSX_FORCE_NONE, "the IB step needs a forcing scheme (guo or edm)");
  return VSX_OK;
}

int sphere_run(const VsxLbm *p, const VsxIB3D *ib, const float *f_in, float *f_out, void *wa, void *wb, int n,
               void *stream) {
  int rc = validate_lbm<D3Q19>(p);
  if (rc) return rc;
  rc = validate_ib3d(p, ib);
  if (rc) return rc;
  VSX_REQUIRE(f_in && f_out && wa, "null buffer");
  VSX_REQUIRE(n >= 0, "negative step count");
  VSX_REQUIRE(n <= 1 || wb, "need two work buffers for more than one step");
  hipStream_t s = as_stream(stream);
  const IB3DDev d = make_dev(*ib);
  const StepCtx c = make_ctx<D3Q19>(*p, Window{d.origin, d.g_win, {ib->win[0], ib->win[1], ib->win[2]}});
  if (n == 0) {
    if (f_out != f_in) VSX_HIP(hipMemcpyAsync(f_out, f_in, sizeof(float) * 19 * c.n, hipMemcpyDeviceToDevice, s));
    return VSX_OK;
  }
  const void *cur = f_in;
  void *bufs[2] = {wa, wb};
  if (ib->pipelined) {
    // window 2+ cells inside the grid, walls without shared corners
    for (int a = 0; a < 3; ++a)
      VSX_REQUIRE(ib->origin[a] >= 2 && ib->origin[a] + ib->win[a] + 2 <= c.dims[a],
                  "the pipelined sphere step needs the IB window 2+ cells inside the grid (axis %d)", a);
    VSX_REQUIRE(!(p->n_walls > 0 && walls_need_edge<D3Q19>(*p)), "the pipelined sphere step needs walls without "
                "shared corners");
    VSX_REQUIRE(ib->halo, "the pipelined sphere step needs the hand-off buffer ib->halo");
    PipeRes *r = nullptr;
    if ((rc = pipe_res(&r))) return rc;
    const int n_halo = (int)halo_blocks(c, d);
    // flags restart with the run (a fill kernel, not a memset: the run may be captured into a graph)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d.flags, (int64_t)4, 0);
    VSX_HIP(hipEventRecord(r->fork, s));
    VSX_HIP(hipStreamWaitEvent(r->s, r->fork, 0));
    VSX_HIP(hipStreamWaitEvent(r->b, r->fork, 0));
    for (int t = 0; t < n; ++t) {
      rc = ib3d_phase(c, d, f_in, t == 0, ib->hist_h ? ib->hist_h + 3 * (int64_t)t : nullptr, r->s, ib->halo, t);
      if (rc) return rc;
      void *dst = bufs[t & 1];
      switch (c.p.storage) {
        case VSX_STORE_F32: bulk3d_pipe_any<Store<VSX_STORE_F32>>(c, d, cur, dst, ib->halo, t, n_halo, t == 0, t == n - 1, r->b); break;
        case VSX_STORE_F16S: bulk3d_pipe_any<Store<VSX_STORE_F16S>>(c, d, cur, dst, ib->halo, t, n_halo, t == 0, t == n - 1, r->b); break;
        case VSX_STORE_BF16S: bulk3d_pipe_any<Store<VSX_STORE_BF16S>>(c, d, cur, dst, ib->halo, t, n_halo, t == 0, t == n - 1, r->b); break;
        default: bulk3d_pipe_any<Store<VSX_STORE_E3M12S>>(c, d, cur, dst, ib->halo, t, n_halo, t == 0, t == n - 1, r->b); break;
      }
      VSX_CHECK_LAUNCH();
      cur = dst;
    }
    VSX_HIP(hipEventRecord(r->join_s, r->s));
    VSX_HIP(hipEventRecord(r->join_b, r->b));
    VSX_HIP(hipStreamWaitEvent(s, r->join_s, 0));
    VSX_HIP(hipStreamWaitEvent(s, r->join_b, 0));
    const StepCtx cs = make_ctx<D3Q19>(*p, Window{nullptr, nullptr, {0, 0, 0}});
    return step_any<D3Q19>(cs, cur, false, f_out, true, VSX_PHASE_STREAM, s);
  }
  for (int t = 0; t < n; ++t) {
    void *dst = bufs[t & 1];
    rc = ib3d_phase(c, d, cur, t == 0, ib->hist_h ? ib->hist_h + 3 * (int64_t)t : nullptr, s);
    if (rc) return rc;
    rc = step_any<D3Q19>(c, cur, t == 0, dst, false, t == 0 ? VSX_PHASE_COLLIDE : VSX_PHASE_FULL, s);
    if (rc) return rc;
    cur = dst;
  }
  const StepCtx cs = make_ctx<D3Q19>(*p, Window{nullptr, nullptr, {0, 0, 0}});
  return step_any<D3Q19>(cs, cur, false, f_out, true, VSX_PHASE_STREAM, s);
}

}  // namespace vsx

using namespace vsx;

extern "C" int64_t vsx_ib3d_work_floats(int32_t n_markers, int64_t cells) { return work_floats(n_markers, cells); }

extern "C" int vsx_sphere3d_run(const VsxLbm *p, const VsxIB3D *ib, const float *f_in, float *f_out, void *work_a,
                                void *work_b, int32_t n_steps, void *stream) {
  return sphere_run(p, ib, f_in, f_out, work_a, work_b, n_steps, stream);
}

extern "C" int vsx_ib3d_transpose_stencil(const int32_t *I, int32_t m, int32_t s, int64_t n_grid, int32_t *cell_ptr,
                                          int32_t *cell_slot, int32_t *tmp, void *stream) {
  VSX_REQUIRE(I && cell_ptr && cell_slot && tmp, "null buffer");
  VSX_REQUIRE(m > 0 && s > 0 && n_grid > 0 && n_grid < ((int64_t)1 << 31), "bad sizes m=%d s=%d n_grid=%lld", m, s,
              (long long)n_grid);
  VSX_REQUIRE((int64_t)m * s < ((int64_t)1 << 31), "too many stencil slots");
  hipStream_t st = as_stream(stream);
  const int64_t slots = (int64_t)m * s;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n_grid + 1, 256)), dim3(256), 0, st, tmp, n_grid + 1, 0);
  hipLaunchKernelGGL(k_count, dim3(grid_for(slots, 256)), dim3(256), 0, st, I, slots, n_grid, tmp);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, (const int32_t *)tmp, n_grid, cell_ptr);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n_grid + 1, 256)), dim3(256), 0, st, tmp, n_grid + 1, 0);
  hipLaunchKernelGGL(k_place, dim3(grid_for(slots, 256)), dim3(256), 0, st, I, slots, n_grid,
                     (const int32_t *)cell_ptr, tmp, cell_slot);
  hipLaunchKernelGGL(k_sort_lists, dim3(grid_for(n_grid, 256)), dim3(256), 0, st, (const int32_t *)cell_ptr, n_grid,
                     cell_slot);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}
