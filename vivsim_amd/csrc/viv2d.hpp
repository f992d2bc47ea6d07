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

// Lane exchange of a 32-bit value from lane `src` (ds_bpermute).  The source
// lane index is recomputed at each use: shared between distant exchanges it
// would hold a register across the whole IB workgroup (or a spill slot).
__device__ __forceinline__ float lane_get(float v, int src) {
  asm volatile("" : "+v"(src));
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ int lane_get(int v, int src) {
  asm volatile("" : "+v"(src));
  return __builtin_amdgcn_ds_bpermute(src << 2, v);
}

// Exchange inside each group of four lanes (DPP quad_perm, no LDS round trip):
// lane 4q + i reads lane 4q + ((CTRL >> 2i) & 3).
template <int CTRL>
__device__ __forceinline__ float quad_get(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
constexpr int kQuadXor1 = 0xB1;   // [1, 0, 3, 2]
constexpr int kQuadXor2 = 0x4E;   // [2, 3, 0, 1]

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
__host__ __device__ inline int imax(int a, int b) { return a > b ? a : b; }

// LDS of the IB workgroup, offsets in 4-byte words (S = window side, SS = S * S,
// M = markers):
//   grid  float2 [SS]          window velocity, then the accumulated spread force
//                              (x and y of a cell side by side); while the
//                              segmented spread is built it holds the uint16
//                              entry offsets of the active cells
//   with BINS (one marker per thread), markers at their bin-order position p
//   (markers whose stencil leaves the window at the end):
//   rec   [max(10.5 M, SS + 1)] 16-byte aligned: wx [4][M], wy [4][M] stencil
//                              weights, force [M] float2 (the force being
//                              spread), bin [M] uint16 (base cell of each
//                              binned marker).  While the bins are built it
//                              holds the int32 bin counters.
//   bend  uint16 [3 S + 4 + SS + 1]  end of each bin in bin order, after 3 S + 4
//                              zeros (bins with a negative index are empty)
//   list  [max(M, SS / 2)]     int32 markers by bin while the bins are built, then
//                              the uint16 list of active window cells
//   seg   (over bend and list once the segmented spread is built) first and
//         last cell (uint16 [kIbThreads] each) and head partial sum (float2
//         [kIbThreads]) of each thread's spread entries
struct IbLds {
  int rec, bend, list, seg, total;
};
// zeros in front of the bin ends: bend[i] = 0 for -bend_pad(S) <= i < 0
__host__ __device__ inline int bend_pad(int S) { return 3 * S + 4; }
__host__ __device__ inline IbLds ib_lds_layout(int S, int M) {
  const int SS = S * S;
  IbLds L;
  L.rec = (2 * SS + 3) & ~3;
  L.bend = (L.rec + imax(10 * M + (M + 1) / 2, SS + 1) + 3) & ~3;
  L.list = L.bend + (bend_pad(S) + SS + 2) / 2;
  L.seg = L.bend;
  L.total = L.bend + imax(L.list - L.bend + imax(M, (SS + 1) / 2), 3 * kIbThreads);
  return L;
}
__host__ __device__ inline int ib_lds_words(int S, int M, bool bins) {
  return bins ? ib_lds_layout(S, M).total : 2 * S * S;
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

// Bin run a of window cell k: the bins lo..lo+3 with lo = k - a S - 3 hold the
// markers [e0, e0 + n) in bin order, and marker e lies in bin lo + j with
// j = (e >= b1) + (e >= b2) + (e >= b3) (the ends of bins lo..lo+2; bins with a
// negative index are empty).  Stencil point (a, 3 - j) of that marker is cell k.
__device__ __forceinline__ void bin_run(const uint16_t *bend, int k, int a, int S, int &e0, int &n, int &b1, int &b2,
                                        int &b3) {
  const int lo = k - a * S - 3;
  e0 = bend[lo - 1];
  n = bend[lo + 3] - e0;
  b1 = bend[lo];
  b2 = bend[lo + 1];
  b3 = bend[lo + 2];
}

// Window cells that receive force from a binned marker (a non-empty bin run),
// listed once per step; returns their number (the list order does not matter).
__device__ __forceinline__ int active_cells(const uint16_t *bend, uint16_t *act, int S, int *s_count) {
  const int SS = S * S, tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) *s_count = 0;
  __syncthreads();
  for (int k0 = 0; k0 < SS; k0 += blockDim.x) {
    const int k = k0 + tid;
    bool on = false;
    if (k < SS) {
      int tot = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int hi = k - a * S;
        tot += bend[hi] - bend[hi - 4];
      }
      on = tot != 0;
    }
    const unsigned long long mask = __ballot(on);
    int first = 0;
    if (lane == 0 && mask) first = atomicAdd(s_count, (int)__popcll(mask));
    first = __shfl(first, 0, 64);
    if (on) act[first + __popcll(mask & ((1ull << lane) - 1ull))] = (uint16_t)k;
  }
  __syncthreads();
  return *s_count;
}

// Spread of the binned markers' forces added onto the active window cells, as
// a gather: cell k receives from the markers of its 4 bin runs (bin_run) with
// weight wx[a] * wy[3 - j].  R of the runs are walked side by side (their
// loads issued together), each summed in bin and marker order, and the run
// sums are added in run order.  One thread per active cell: no atomics,
// deterministic.
template <int R>
__device__ __forceinline__ void bin_spread(float2 *grid, const float *rec, int M, const uint16_t *bend,
                                           const uint16_t *act, int n_act, int S) {
  const float2 *force = reinterpret_cast<const float2 *>(rec + 8 * M);
  for (int q = threadIdx.x; q < n_act; q += blockDim.x) {
    const int k = act[q];
    float gx[4], gy[4];
#pragma unroll
    for (int h = 0; h < 4; h += R) {
      int e0[R], n[R], b1[R], b2[R], b3[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        bin_run(bend, k, h + r, S, e0[r], n[r], b1[r], b2[r], b3[r]);
        gx[h + r] = gy[h + r] = 0.f;
      }
      int len = n[0];
#pragma unroll
      for (int r = 1; r < R; ++r) len = max(len, n[r]);
      for (int t = 0; t < len; ++t) {
        // a finished run reads marker 0 (present whenever a run is non-empty)
        // with weight 0
        float w[R];
        float2 fm[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int e = t < n[r] ? e0[r] + t : 0;
          const int j = (e >= b1[r]) + (e >= b2[r]) + (e >= b3[r]);
          w[r] = rec[(h + r) * M + e] * rec[(7 - j) * M + e];   // wx[a] * wy[3 - j]
          w[r] = t < n[r] ? w[r] : 0.f;
          fm[r] = force[e];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          gx[h + r] += fm[r].x * w[r];
          gy[h + r] += fm[r].y * w[r];
        }
      }
    }
    float2 g = grid[k];
    g.x += (gx[0] + gx[1]) + (gx[2] + gx[3]);
    g.y += (gy[0] + gy[1]) + (gy[2] + gy[3]);
    grid[k] = g;
  }
}

// ------------------------------------------------------------------ segmented spread
// The 16 stencil entries of every binned marker, ordered by window cell (cells
// ascending; within a cell runs a = 0..3, markers in bin order, as bin_spread
// sums them), kSpreadEPT consecutive entries per thread in registers, 16 bits
// each: marker position e (10 bits) and stencil point (a, b) (2 + 2 bits),
// with their weights.  A spread reads the force and the base cell of each
// entry's marker (the cell is base + a S + b) and adds the products as a
// segmented sum: the thread holding the first
// entry of a cell owns it and adds the partial sums of the following threads
// whose entries continue that cell, in thread order.  Deterministic, no
// atomics, balanced over the threads.
constexpr int kSpreadEPT = 8;

struct SpreadSeg {
  uint32_t eab[kSpreadEPT / 2];   // entry t in bits 16 (t & 1) .. 16 (t & 1) + 15
  float w[kSpreadEPT];            // stencil weight wx[a] * wy[b] (0 for padding)
  int nthr;                       // threads holding entries
};

__device__ __forceinline__ int seg_entry(const SpreadSeg &sg, int t) {
  return (int)((sg.eab[t >> 1] >> (16 * (t & 1))) & 0xffffu);
}

// Entry count of window cell k (its 4 bin runs).
__device__ __forceinline__ int cell_entries(const uint16_t *bend, int k, int S) {
  int tot = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int hi = k - a * S;
    tot += bend[hi] - bend[hi - 4];
  }
  return tot;
}

// Builds the entries (every thread calls it; ends with a barrier).  Returns
// false (before any barrier, the same on every thread) when the binned markers
// have more entries than the workgroup holds (16 n > kSpreadEPT nt) or than
// the staging area takes; the caller then uses bin_spread.  tmp: uint16
// scratch of 4 SS (first entry of each cell, then the staged entries);
// fcell, lcell: first and last cell of each thread's entries (written after
// the entries are built, so they may overlay bend).  rec holds the weights
// and, after the forces, the base cell of each binned marker (uint16).
// binned / pos / base: this thread's marker, its bin-order position and base
// cell.  Order of the entries: cells ascending; within a cell runs a = 0..3,
// markers in bin order, as bin_spread sums them.  Each marker writes its 16
// entries at their ranks: cell k = base + a S + b gets from run a the
// markers of bins k - a S - 3 .. k - a S, so the rank is the cell's first
// entry + the entries of its runs 0..a-1 + (pos - first marker of run a).
__device__ __forceinline__ bool seg_build(const float *rec, int M, const uint16_t *bend, int S, uint16_t *tmp,
                                          uint16_t *fcell, uint16_t *lcell, int *s_wsum, SpreadSeg &sg, bool binned,
                                          int pos, int base) {
  const int SS = S * S, tid = threadIdx.x, nt = blockDim.x;
  const uint16_t *binof = reinterpret_cast<const uint16_t *>(rec + 10 * M);
  const int T = 16 * (int)bend[SS - 1];
  const int st0 = (SS + 8) & ~7;   // staged entries, 16-byte aligned
  sg.nthr = (T + kSpreadEPT - 1) / kSpreadEPT;
  if (T > kSpreadEPT * nt || st0 + kSpreadEPT * sg.nthr > 4 * SS) return false;
  uint16_t *cstart = tmp, *stage = tmp + st0;
  // first entry of every cell: per-thread chunks of consecutive cells, one
  // block scan of the entry counts
  const int chunk = (SS + nt - 1) / nt, k0 = min(tid * chunk, SS), k1 = min(k0 + chunk, SS);
  auto cell_count = [&](int k) {
    int e = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) e += bend[k - a * S] - bend[k - a * S - 4];
    return e;
  };
  int mine = 0;
  for (int k = k0; k < k1; ++k) mine += cell_count(k);
  int v = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if ((tid & 63) >= o) v += t;
  }
  if ((tid & 63) == 63) s_wsum[tid >> 6] = v;
  __syncthreads();
  if (tid < 64) {   // exclusive scan of the wave totals by the first wavefront
    const int nw = (nt + 63) / 64;
    const int x = tid < nw ? s_wsum[tid] : 0;
    int y = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(y, o, 64);
      if (tid >= o) y += t;
    }
    if (tid < nw) s_wsum[tid] = y - x;
  }
  __syncthreads();
  int r = s_wsum[tid >> 6] + v - mine;
  for (int k = k0; k < k1; ++k) {
    cstart[k] = (uint16_t)r;
    r += cell_count(k);
  }
  __syncthreads();
  // every binned marker writes its 16 entries at their ranks
  if (binned) {
    int e0b[4], pre[4] = {0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < 4; ++b) e0b[b] = bend[base + b - 4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int k = base + a * S + b;
        if (a > 0) pre[b] += bend[base + a * S + b] - bend[base + a * S + b - 4];
        stage[cstart[k] + pre[b] + pos - e0b[b]] = (uint16_t)(pos | (a << 10) | (b << 12));
      }
    }
  }
  __syncthreads();   // bend and cstart are dead from here
  // this thread's entries [8 tid, 8 tid + 8): past the end, the last entry
  // again with weight 0
  const int g0 = tid * kSpreadEPT;
  int kfirst = 0, klast = 0;
#pragma unroll
  for (int t = 0; t < kSpreadEPT / 2; ++t) sg.eab[t] = 0;
  if (g0 < T) {
    const uint4 raw = *reinterpret_cast<const uint4 *>(stage + g0);
    const uint32_t w4[4] = {raw.x, raw.y, raw.z, raw.w};
    uint32_t last = 0;
#pragma unroll
    for (int t = 0; t < kSpreadEPT; ++t) {
      uint32_t v16 = (w4[t >> 1] >> (16 * (t & 1))) & 0xffffu;
      v16 = g0 + t < T ? v16 : last;
      last = v16;
      sg.eab[t >> 1] |= v16 << (16 * (t & 1));
      const int e = v16 & 1023, a = (v16 >> 10) & 3, b = (v16 >> 12) & 3;
      sg.w[t] = g0 + t < T ? rec[a * M + e] * rec[(4 + b) * M + e] : 0.f;
      const int k = (int)binof[e] + a * S + b;
      if (t == 0) kfirst = k;
      klast = k;
    }
  } else {
#pragma unroll
    for (int t = 0; t < kSpreadEPT; ++t) sg.w[t] = 0.f;
  }
  if (tid < sg.nthr) {
    fcell[tid] = (uint16_t)kfirst;
    lcell[tid] = (uint16_t)klast;
  }
  __syncthreads();
  return true;
}

// One spread of the forces (rec force [M] float2) added onto the grid.
__device__ __forceinline__ void seg_spread(float2 *grid, const float *rec, int M, int S, const SpreadSeg &sg,
                                           const uint16_t *fcell, const uint16_t *lcell, float2 *headp) {
  const int tid = threadIdx.x;
  const float2 *force = reinterpret_cast<const float2 *>(rec + 8 * M);
  const uint16_t *binof = reinterpret_cast<const uint16_t *>(rec + 10 * M);
  int tail_cell = -1;
  float2 tail = make_float2(0.f, 0.f);
  if (tid < sg.nthr) {
    const bool cont_prev = tid > 0 && fcell[tid] == lcell[tid - 1];
    const bool cont_next = tid + 1 < sg.nthr && lcell[tid] == fcell[tid + 1];
    float sx = 0.f, sy = 0.f;
    bool first = true;
    int cur = fcell[tid];
    // the entries are made opaque here: what is derived from them (addresses,
    // cells) would otherwise be kept across the MDF loop in registers
    SpreadSeg es = sg;
#pragma unroll
    for (int t = 0; t < kSpreadEPT / 2; ++t) asm volatile("" : "+v"(es.eab[t]));
    // half the entries' loads in flight at a time (register budget)
#pragma unroll
    for (int h = 0; h < kSpreadEPT; h += kSpreadEPT / 2) {
      int cell[kSpreadEPT / 2];
      float w[kSpreadEPT / 2];
      float2 fm[kSpreadEPT / 2];
#pragma unroll
      for (int t = 0; t < kSpreadEPT / 2; ++t) {
        const int v = seg_entry(es, h + t);
        const int e = v & 1023, a = (v >> 10) & 3, b = (v >> 12) & 3;
        fm[t] = force[e];
        w[t] = sg.w[h + t];
        cell[t] = (int)binof[e] + a * S + b;
      }
#pragma unroll
      for (int t = 0; t < kSpreadEPT / 2; ++t) {
        if (cell[t] != cur) {   // the segment of cur ends inside this thread
          if (first && cont_prev) {
            headp[tid] = make_float2(sx, sy);
          } else {
            float2 g = grid[cur];
            g.x += sx;
            g.y += sy;
            grid[cur] = g;
          }
          first = false;
          sx = sy = 0.f;
          cur = cell[t];
        }
        sx += w[t] * fm[t].x;
        sy += w[t] * fm[t].y;
      }
      asm volatile("" : "+v"(sx), "+v"(sy));
    }
    if (first && cont_prev) {
      headp[tid] = make_float2(sx, sy);
    } else if (cont_next) {
      tail_cell = cur;
      tail = make_float2(sx, sy);
    } else {
      float2 g = grid[cur];
      g.x += sx;
      g.y += sy;
      grid[cur] = g;
    }
  }
  __syncthreads();
  if (tail_cell >= 0) {
    for (int j = tid + 1; j < sg.nthr && fcell[j] == tail_cell; ++j) {
      const float2 h = headp[j];
      tail.x += h.x;
      tail.y += h.y;
      if (lcell[j] != tail_cell) break;
    }
    float2 g = grid[tail_cell];
    g.x += tail.x;
    g.y += tail.y;
    grid[tail_cell] = g;
  }
}

// Error code of a fused step: into win_origin[3] and, when the caller gave
// one, the pinned host word ib.status.
__device__ __forceinline__ void set_ib_error(const VsxIB2D &ib, int code) {
  __hip_atomic_store(&ib.win_origin[3], code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (ib.status) __hip_atomic_store(ib.status, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Bounded wait of thread 0 until *flag >= target, polling with relaxed
// agent-scope loads every SLEEP x 64 cycles (an acquire per poll would
// invalidate the L2 every time; data published with the flag is read with
// ld_agent, see the hand-off rules of common.hpp).  Waits that are long anyway
// poll less often, which keeps hundreds of pollers off the flag's memory
// channel.  On time-out the error flag ib.win_origin[3] is set and the wait ends.
template <int SLEEP = 2>
__device__ __forceinline__ void spin_until(const VsxIB2D &ib, int32_t *flag, int target) {
  int spins = 0;
  while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(SLEEP);
    if (++spins > kSpinLimit) {
      set_ib_error(ib, 1);
      break;
    }
  }
}

// Interpolation of the window field at a marker (ib/stencil.py:54-78): 4 x 4
// stencil from flat base index, gather rule for indices outside the window.
__device__ __forceinline__ float2 interp_marker(const float2 *grid, int base, int S, const float (&wx)[4],
                                                const float (&wy)[4]) {
  const int SS = S * S;
  // the indices are recomputed at every call: kept across the MDF loop they
  // would take 16 registers
  asm volatile("" : "+v"(base));
  float sx = 0.f, sy = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float2 v = grid[gather_index32(base + i * S + j, SS)];
      const float w = wx[i] * wy[j];
      sx += w * v.x;
      sy += w * v.y;
    }
    // one stencil row of loads in flight at a time (register budget)
    asm volatile("" : "+v"(sx), "+v"(sy));
  }
  return make_float2(sx, sy);
}

// Spread of one marker's force with LDS atomics (scatter rule of
// ib/stencil.py:106-110: indices outside the window are dropped).
__device__ __forceinline__ void scatter_marker(float *gflat, int base, int S, const float (&wx)[4], const float (&wy)[4],
                                               float fx, float fy) {
  const int SS = S * S;
  asm volatile("" : "+v"(base));   // indices recomputed at every call, as in interp_marker
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int si = scatter_index32(base + i * S + j, SS);
      if (si >= 0) {
        const float w = wx[i] * wy[j];
        atomicAdd(&gflat[2 * si], fx * w);
        atomicAdd(&gflat[2 * si + 1], fy * w);
      }
    }
}

// Window velocity by the IB workgroup, in rounds of UW cells per thread: the
// loads of window cells base + j nt + tid (j < UW) off the domain walls ...
template <class SRC, bool PULL, int UW>
__device__ __forceinline__ void window_pull(const StepCtx &c, const typename SRC::T *__restrict__ src, int base, int S,
                                            int xc, int yc, float (&f)[UW][9], bool (&on)[UW]) {
  const int SS = S * S, nt = blockDim.x;
#pragma unroll
  for (int j = 0; j < UW; ++j) {
    const int k = base + j * nt + (int)threadIdx.x;
    const int co[3] = {xc + k / S, yc + k % S, 0};
    on[j] = k < SS && !(PULL && c.p.n_walls > 0 && wall_class<D2Q9>(c, co) != -1);
    if (on[j]) {
      if constexpr (PULL) pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f[j]);
      else local<D2Q9, SRC>(c, src, cell_of<D2Q9>(c, co[0], co[1], 0), f[j]);
    }
  }
}

// ... and their macroscopic velocity into the grid.
template <int UW>
__device__ __forceinline__ void window_store(float2 *grid, int base, const float (&f)[UW][9], const bool (&on)[UW]) {
#pragma unroll
  for (int j = 0; j < UW; ++j) {
    if (on[j]) {
      float rho, u[3];
      macroscopic<D2Q9>(f[j], rho, u);
      grid[base + j * (int)blockDim.x + (int)threadIdx.x] = make_float2(u[0], u[1]);
    }
  }
}

// Wait (thread 0, bounded) until the n_wtiles window tiles of step gstep have
// published the window velocity in ib.g_win (read it with ld_agent), and
// reset their counter for step gstep + 2.
__device__ __forceinline__ void wait_window_u(const VsxIB2D &ib, int gstep, int n_wtiles) {
  if (threadIdx.x == 0) {
    int32_t *cnt_u = &ib.win_origin[4 + (gstep & 1)];
    spin_until(ib, cnt_u, n_wtiles);
    __hip_atomic_store(cnt_u, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// window values per thread held in registers between the wait for the window
// velocity and its store into LDS (windows with 2 S^2 <= kUPrefetch threads)
constexpr int kUPrefetch = 8;

// One workgroup per body (examples/2d/vortex_induced_vibration.py:99-137):
//   window u     macroscopic velocity of the post-streaming state (pull + walls);
//                U_GLOBAL: published in ib.g_win by the n_wtiles tiles that
//                cover the window (counter ib.win_origin[4 + (gstep & 1)])
//   stencil      ib/stencil.py:7-51 at marker coordinates (x_m + d) - origin (4x4)
//   MDF          ib/mdf.py:35-64: ub = interp(u); n_iter x (dF = (U - ub) 2 ds,
//                F += dF, ub += interp(0.5 spread(dF))); force window spread(F).
//                By linearity of spread and interp the spreads accumulate in one
//                grid G = sum spread(dF), ub = ub0 + 0.5 interp(G), and the
//                force window is the final G (n_iter spreads, no separate
//                spread of F).  BINS: the spread is a gather over the markers
//                binned by stencil base cell (bin_spread): no atomics,
//                deterministic; the stencil weights live in LDS.  Markers whose
//                stencil leaves the window, and the path with several markers
//                per thread, use LDS atomics with the scatter rule
//   body force   h = sum_m (-F_m) + a * area in FP64 (wave shuffles + LDS), dyn.py:126-136
//   structure    Newmark-beta in FP64 on lane 0, dyn.py:5-51
// Writes the force window to ib.g_win and the new body state to body_out.
template <class SRC, bool PULL, int MPT, int UW, int TAG, bool U_GLOBAL, int SPREAD_RUNS = 2, bool SEG = true>
__device__ __forceinline__ void ib_block(const StepCtx &c, const VsxIB2D &ib, const typename SRC::T *__restrict__ src,
                                         int hidx, int gstep, int n_wtiles, const double *body_in, double *body_out,
                                         float *lds) {
  const int S = ib.win, SS = S * S, M = ib.n_markers;
  constexpr bool BINS = MPT == 1;
  constexpr int WR = BINS ? 1 : MPT;   // markers whose weights stay in registers
  float2 *grid = reinterpret_cast<float2 *>(lds);
  float *gflat = lds;
  const IbLds lay = ib_lds_layout(S, M);
  float *rec = lds + lay.rec;
  float2 *rforce = reinterpret_cast<float2 *>(rec + 8 * M);
  int *cnt = reinterpret_cast<int *>(rec);
  uint16_t *bend = reinterpret_cast<uint16_t *>(lds + lay.bend) + bend_pad(S);
  int *blist = reinterpret_cast<int *>(lds + lay.list);
  uint16_t *act = reinterpret_cast<uint16_t *>(lds + lay.list);
  __shared__ double s_body[6];
  __shared__ int s_org[4];
  __shared__ int s_irregular, s_nact, s_nirr;
  __shared__ int s_wsum[kIbThreads / 64];
  __shared__ double s_red[2][kIbThreads / 64];
  const int tid = threadIdx.x, nt = blockDim.x;

  // Marker q of this thread: consecutive lanes of a wavefront hold markers
  // nt/64 apart along the body, so the lanes of one LDS access touch different
  // cells.  The marker inputs are loaded first, beside the body state.
  const int nwave = nt >> 6;
  int mk[MPT];
  double mx[MPT], my[MPT];
  float ds2[MPT];
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    mk[q] = q * nt + (tid & 63) * nwave + (tid >> 6);
    mx[q] = my[q] = 0.0;
    ds2[q] = 0.f;
    if (mk[q] < M) {
      mx[q] = ib.marker_x[mk[q]];
      my[q] = ib.marker_y[mk[q]];
      ds2[q] = (ib.ds_arr ? ib.ds_arr[mk[q]] : ib.ds) * 2.0f;
    }
  }
  if (tid < 6) s_body[tid] = body_in[tid];
  if (tid == 0) s_irregular = s_nirr = 0;
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

  // stencils of the markers (weights in registers until they are stored)
  int base_idx[MPT];
  bool regular[MPT];
  float wx[WR][4], wy[WR][4];
  float wx_t[4], wy_t[4];
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    base_idx[q] = 0;
    regular[q] = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) wx_t[i] = wy_t[i] = 0.f;
    if (mk[q] < M) {
      const float rx = (float)(mx[q] + s_body[4] - (double)x0);
      const float ry = (float)(my[q] + s_body[5] - (double)y0);
      const Axis<4> ax = axis_stencil<4>(ib.kernel, rx), ay = axis_stencil<4>(ib.kernel, ry);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        wx_t[i] = ax.w[i];
        wy_t[i] = ay.w[i];
      }
      const int bx = ax.ref - 1, by = ay.ref - 1;   // window row / column of stencil point (0, 0)
      base_idx[q] = bx * S + by;                    // its flat window index
      regular[q] = BINS && bx >= 0 && bx <= S - 4 && by >= 0 && by <= S - 4;
      if (!regular[q]) s_irregular = 1;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      wx[BINS ? 0 : q][i] = wx_t[i];
      wy[BINS ? 0 : q][i] = wy_t[i];
    }
  }

  int my_pos = 0, n_act = 0;
  bool use_seg = false, u_pref = false;
  float upf[kUPrefetch];
  SpreadSeg sg;
  uint16_t *fcell = reinterpret_cast<uint16_t *>(lds + lay.seg), *lcell = fcell + kIbThreads;
  float2 *headp = reinterpret_cast<float2 *>(lds + lay.seg + kIbThreads);
  if constexpr (BINS) {
    // regular markers binned by base cell: count, scan, place; a marker's
    // position = start of its bin + its rank by marker index within the bin;
    // the other markers take the positions from M - 1 down
    const bool live = mk[0] < M, binned = live && regular[0];
    for (int k = tid; k <= SS; k += nt) cnt[k] = 0;
    __syncthreads();
    if (binned) atomicAdd(&cnt[base_idx[0]], 1);
    __syncthreads();
    block_exclusive_scan(cnt, SS, s_wsum);
    if (binned) blist[atomicAdd(&cnt[base_idx[0]], 1)] = mk[0];
    __syncthreads();
    if (binned) {
      const int s0 = base_idx[0] > 0 ? cnt[base_idx[0] - 1] : 0, e0 = cnt[base_idx[0]];
      int rank = 0;
      for (int i = s0; i < e0; ++i) rank += blist[i] < mk[0];
      my_pos = s0 + rank;
    } else if (live) {
      my_pos = M - 1 - atomicAdd(&s_nirr, 1);
    }
    for (int k = tid - bend_pad(S); k < SS; k += nt) bend[k] = k < 0 ? 0 : (uint16_t)cnt[k];
    __syncthreads();   // counters and marker list are dead from here
    if (live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        rec[i * M + my_pos] = wx[0][i];
        rec[(4 + i) * M + my_pos] = wy[0][i];
      }
      if (binned) reinterpret_cast<uint16_t *>(rec + 10 * M)[my_pos] = (uint16_t)base_idx[0];
    }
    if constexpr (U_GLOBAL) {
      // the window velocity is published early (the bulk tiles wait for it):
      // its loads are issued now and stay in registers over the entry build
      wait_window_u(ib, gstep, n_wtiles);
      u_pref = 2 * SS <= kUPrefetch * nt;
      if (u_pref) {
#pragma unroll
        for (int j = 0; j < kUPrefetch; ++j) {
          const int w = tid + j * nt;   // interleaved grid word: cell w / 2, component w % 2
          upf[j] = w < 2 * SS ? ld_agent(&ib.g_win[(w & 1) * SS + (w >> 1)]) : 0.f;
        }
      }
    }
    if constexpr (SEG)
      use_seg = seg_build(rec, M, bend, S, reinterpret_cast<uint16_t *>(lds), fcell, lcell, s_wsum, sg, binned, my_pos,
                          base_idx[0]);
    if (!use_seg) n_act = active_cells(bend, act, S, &s_nact);   // ends with a barrier
  }

  // window velocity
  if constexpr (U_GLOBAL) {
    if (u_pref) {
#pragma unroll
      for (int j = 0; j < kUPrefetch; ++j) {
        const int w = tid + j * nt;
        if (w < 2 * SS) gflat[w] = upf[j];
      }
    } else {
      if (!BINS) wait_window_u(ib, gstep, n_wtiles);
      for (int k = tid; k < SS; k += nt) grid[k] = make_float2(ld_agent(&ib.g_win[k]), ld_agent(&ib.g_win[SS + k]));
    }
  } else {
    for (int base = 0; base < SS; base += UW * nt) {
      float fu[UW][9];
      bool fu_on[UW];
      window_pull<SRC, PULL, UW>(c, src, base, S, xc, yc, fu, fu_on);
      window_store<UW>(grid, base, fu, fu_on);
    }
    // cells on a domain wall (rare: only when the window touches a wall), one
    // at a time, with nothing else in flight across the call
    if (PULL && c.p.n_walls > 0) {
      for (int k = tid; k < SS; k += nt) {
        const int co[3] = {xc + k / S, yc + k % S, 0};
        if (wall_class<D2Q9>(c, co) != -1) grid[k] = window_wall_u<SRC, TAG>(c, src, co[0], co[1]);
      }
    }
  }
  __syncthreads();

  // weights of marker q (from LDS when binned)
  auto weights = [&](int q, float(&wxq)[4], float(&wyq)[4]) {
    if constexpr (BINS) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        wxq[i] = rec[i * M + my_pos];
        wyq[i] = rec[(4 + i) * M + my_pos];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        wxq[i] = wx[q][i];
        wyq[i] = wy[q][i];
      }
    }
  };

  // initial interpolation of the window velocity
  float ub0x[MPT], ub0y[MPT], ubx[MPT], uby[MPT], Fx[MPT], Fy[MPT];
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    ub0x[q] = ub0y[q] = Fx[q] = Fy[q] = 0.f;
    if (mk[q] < M) {
      float wxq[4], wyq[4];
      weights(q, wxq, wyq);
      const float2 v = interp_marker(grid, base_idx[q], S, wxq, wyq);
      ub0x[q] = v.x;
      ub0y[q] = v.y;
    }
    ubx[q] = ub0x[q];
    uby[q] = ub0y[q];
  }
  const float Ux = (float)s_body[2], Uy = (float)s_body[3];
  __syncthreads();   // the window velocity is dead from here; the grid accumulates the spread force
  for (int k = tid; k < 2 * SS; k += nt) gflat[k] = 0.f;

  // multi-direct forcing (ib/mdf.py:45-64)
  for (int it = 0; it < ib.n_iter; ++it) {
    float sfx[MPT], sfy[MPT];
#pragma unroll
    for (int q = 0; q < MPT; ++q) {
      sfx[q] = (Ux - ubx[q]) * ds2[q];
      sfy[q] = (Uy - uby[q]) * ds2[q];
      Fx[q] += sfx[q];
      Fy[q] += sfy[q];
    }
    if constexpr (BINS) {
      if (mk[0] < M && regular[0]) rforce[my_pos] = make_float2(sfx[0], sfy[0]);
    }
    __syncthreads();
    if constexpr (BINS) {
      if (SEG && use_seg) seg_spread(grid, rec, M, S, sg, fcell, lcell, headp);
      else bin_spread<SPREAD_RUNS>(grid, rec, M, bend, act, n_act, S);
      if (s_irregular) __syncthreads();
    }
    if (!BINS || s_irregular) {
#pragma unroll
      for (int q = 0; q < MPT; ++q) {
        if (mk[q] < M && !regular[q]) {
          float wxq[4], wyq[4];
          weights(q, wxq, wyq);
          scatter_marker(gflat, base_idx[q], S, wxq, wyq, sfx[q], sfy[q]);
        }
      }
    }
    __syncthreads();
    if (it + 1 == ib.n_iter) break;
#pragma unroll
    for (int q = 0; q < MPT; ++q) {
      if (mk[q] < M) {
        float wxq[4], wyq[4];
        weights(q, wxq, wyq);
        const float2 v = interp_marker(grid, base_idx[q], S, wxq, wyq);
        ubx[q] = ub0x[q] + 0.5f * v.x;
        uby[q] = ub0y[q] + 0.5f * v.y;
      }
    }
  }
  if (ib.n_iter == 0) __syncthreads();

  // force window -> global
  for (int k = tid; k < SS; k += nt) {
    const float2 g = grid[k];
    st_agent(&ib.g_win[k], g.x);
    st_agent(&ib.g_win[SS + k], g.y);
  }

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

// ------------------------------------------------------------------ IB workgroup with an atomic spread
// The spread accumulates into LDS with integer atomics instead of a gather
// over sorted entries, so a step needs no binning or entry build.  Every
// product F w is rounded once to int64 fixed point with 2^-36 resolution; the
// sums are then exact, hence independent of the order of the atomics
// (deterministic), and rounded once to FP32 when read.  LDS float atomics
// would be order-dependent and, measured on gfx950, cost about 15x more per
// spread (20.7 us against 1.2-1.4 us for 512 markers on a 60 x 60 window).
constexpr double kFixScale = 68719476736.0;        // 2^36
constexpr double kFixMagic = 6755399441055744.0;   // 1.5 * 2^52: fma(v, 2^36, magic) keeps the integer in the mantissa
constexpr float kFixClamp = 16384.f;               // |contribution| beyond this saturates (|v| 2^36 < 2^51)

__device__ __forceinline__ unsigned long long to_fix(float v) {
  v = fminf(fmaxf(v, -kFixClamp), kFixClamp);
  const double t = fma((double)v, kFixScale, kFixMagic);
  return (unsigned long long)(__double_as_longlong(t) - __double_as_longlong(kFixMagic));
}
__device__ __forceinline__ float from_fix(unsigned long long q) { return (float)((double)(long long)q * (1.0 / kFixScale)); }

// accumulation grid: two planes of SS int64 words, x components then y
// components (cell k at words k and SS + k).  Eight-byte elements per plane
// spread the lanes of one atomic over twice as many LDS banks as interleaved
// (x, y) pairs would.
__device__ __forceinline__ void acc_add(float *lds, int SS, int k, float vx, float vy) {
  unsigned long long *g = reinterpret_cast<unsigned long long *>(lds);
  atomicAdd(&g[k], to_fix(vx));
  atomicAdd(&g[SS + k], to_fix(vy));
}
__device__ __forceinline__ float2 acc_get(const float *lds, int SS, int k) {
  const unsigned long long *g = reinterpret_cast<const unsigned long long *>(lds);
  return make_float2(from_fix(g[k]), from_fix(g[SS + k]));
}
__host__ __device__ inline int ib_atomic_lds_words(int S) { return 4 * S * S; }
// SPLIT layout (pipelined IB kernel): velocity, grid, mark array, cell list
__host__ __device__ inline int ib_split_lds_words(int S) { return 6 * S * S + (S * S + 3) / 4 + (S * S + 1) / 2; }

// Thread t of the IB workgroup works on marker mk and stencil rows
// [r0, r0 + rows): with M <= nt / 4 four threads share a marker (one row
// each, a group of four lanes), with M <= nt / 2 two threads (rows 0-1 and
// 2-3, adjacent lanes), else one thread takes all four rows.  Consecutive
// lane groups hold markers nwave apart along the body, so the lanes of one
// atomic mostly address different cells.
struct IbLane {
  int mk, r0, rows, tpm;
  // r0 recomputed from the thread index where it is used: a value kept live
  // across the IB workgroup's phases would take a register (or a spill slot)
  __device__ __forceinline__ int first_row() const {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return tpm == 4 ? (t & 3) : tpm == 2 ? 2 * (t & 1) : 0;
  }
};
// Sum of a per-thread partial over the threads of the marker; every thread of
// the marker gets the same value ((r0 + r1) + (r2 + r3) with four threads)
__device__ __forceinline__ float marker_sum(float v, const IbLane &l) {
  if (l.tpm >= 2) v += quad_get<kQuadXor1>(v);
  if (l.tpm == 4) v += quad_get<kQuadXor2>(v);
  return v;
}
// Threads per marker for M markers and nt threads (host and device).
__host__ __device__ inline int ib_tpm(int M, int nt) { return 4 * M <= nt ? 4 : (2 * M <= nt ? 2 : 1); }
// TPM: the mapping fixed at compile time (ib_tpm(M, blockDim.x), chosen by the
// launcher), so that a kernel holds the code of one mapping only; 0: chosen here.
template <int TPM = 0>
__device__ __forceinline__ IbLane ib_lane(int M) {
  const int tid = threadIdx.x, nt = blockDim.x, nwave = nt >> 6, lane = tid & 63, wave = tid >> 6;
  IbLane l;
  // uniform: kept in a scalar register
  l.tpm = TPM ? TPM : __builtin_amdgcn_readfirstlane(ib_tpm(M, nt));
  if (l.tpm == 4) {
    l.mk = (lane >> 2) * nwave + wave;
    l.r0 = lane & 3;
    l.rows = 1;
  } else if (l.tpm == 2) {
    l.mk = (lane >> 1) * nwave + wave;
    l.r0 = 2 * (lane & 1);
    l.rows = 2;
  } else {
    l.mk = lane * nwave + wave;
    l.r0 = 0;
    l.rows = 4;
  }
  return l;
}

// Partial interpolation over this thread's rows (gather rule of ib/stencil.py:54-78
// for flat indices outside the window), summed over the threads of the marker.
// INSIDE (uniform over the wave): every stencil index of every lane lies in the
// window, so a row is a base address plus constant offsets.
template <bool ACC, bool INSIDE>
__device__ __forceinline__ float2 interp_rows(const float *grid, int base, int S, const IbLane &l, const float (&wr)[4],
                                              const float (&wy)[4]) {
  const int SS = S * S;
  asm volatile("" : "+v"(base));   // indices recomputed at every call (register budget)
  auto fetch = [&](int p) {
    const int k = INSIDE ? p : gather_index32(p, SS);
    return ACC ? acc_get(grid, SS, k) : reinterpret_cast<const float2 *>(grid)[k];
  };
  if (l.tpm == 4) {
    // the summation order of two threads per marker, (row 0 then row 1) +
    // (row 2 then row 3), so that both mappings give the same bits: rows 1
    // and 3 continue from the sums of rows 0 and 2, taken from the lane before
    float2 v[4];
    float w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = fetch(base + j);
      w[j] = wr[0] * wy[j];
    }
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sx += w[j] * v[j].x;
      sy += w[j] * v[j].y;
    }
    float tx = quad_get<0xA0>(sx), ty = quad_get<0xA0>(sy);   // [0, 0, 2, 2]
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      tx += w[j] * v[j].x;
      ty += w[j] * v[j].y;
    }
    // lanes 1 and 3 of the group hold the two pair sums
    return make_float2(quad_get<0x55>(tx) + quad_get<0xFF>(tx), quad_get<0x55>(ty) + quad_get<0xFF>(ty));
  }
  float sx = 0.f, sy = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i < l.rows) {
      const int rb = base + i * S;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float2 v = fetch(rb + j);
        const float w = wr[i] * wy[j];
        sx += w * v.x;
        sy += w * v.y;
      }
      asm volatile("" : "+v"(sx), "+v"(sy));
    }
  }
  return make_float2(marker_sum(sx, l), marker_sum(sy, l));
}
template <bool ACC>
__device__ __forceinline__ float2 interp_rows(const float *grid, int base, int S, const IbLane &l, const float (&wr)[4],
                                              const float (&wy)[4], bool inside) {
  return inside ? interp_rows<ACC, true>(grid, base, S, l, wr, wy) : interp_rows<ACC, false>(grid, base, S, l, wr, wy);
}

// Spread of (fx, fy) over this thread's rows (scatter rule of ib/stencil.py:106-110:
// flat indices outside the window are dropped).
template <bool INSIDE>
__device__ __forceinline__ void spread_rows(float *grid, int base, int S, const IbLane &l, const float (&wr)[4],
                                            const float (&wy)[4], float fx, float fy) {
  const int SS = S * S;
  asm volatile("" : "+v"(base));
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i < l.rows) {
      const int rb = base + i * S;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = INSIDE ? rb + j : scatter_index32(rb + j, SS);
        if (k >= 0) {
          const float w = wr[i] * wy[j];
          acc_add(grid, SS, k, fx * w, fy * w);
        }
      }
    }
  }
}
__device__ __forceinline__ void spread_rows(float *grid, int base, int S, const IbLane &l, const float (&wr)[4],
                                            const float (&wy)[4], float fx, float fy, bool inside) {
  if (inside) spread_rows<true>(grid, base, S, l, wr, wy, fx, fy);
  else spread_rows<false>(grid, base, S, l, wr, wy, fx, fy);
}

// Peskin 4-point kernel (ib/kernels.py) without branches: the same float32
// expressions as delta_kernel, both sides evaluated and one selected.
__device__ __forceinline__ float peskin4(float r) {
  const float a = fabsf(r);
  const bool in = a < 1.f;
  const float q = sqrtf(in ? 1.f + 4.f * a - 4.f * (a * a) : -7.f + 12.f * a - 4.f * (a * a));
  const float w = (in ? 3.f - 2.f * a + q : 5.f - 2.f * a - q) * 0.125f;
  return a > 2.f ? 0.f : w;
}

// Stencil weights of this thread's rows (wr: x nodes r0 .. r0 + rows - 1,
// zero beyond) and of the marker's four y nodes; with two or four threads per
// marker each computes its share of the y weights and takes the others from
// the marker's other lanes.  Nodes ref - 1 .. ref + 2, ref = floor(x), weights
// phi(node - x) in float32 (ib/stencil.py:7-51).
template <class PHI>
__device__ __forceinline__ void ib_weights(PHI phi, const IbLane &l, bool live, float rx, float ry, float (&wr)[4],
                                           float (&wy)[4]) {
  const int r0 = l.first_row();
  const int xref = (int)floorf(rx), yref = (int)floorf(ry);
#pragma unroll
  for (int i = 0; i < 4; ++i) wr[i] = (live && i < l.rows) ? phi((float)(xref - 1 + r0 + i) - rx) : 0.f;
  if (l.tpm == 4) {
    // y weight r0 here; lane 4q + j of the marker holds weight j
    const float a = phi((float)(yref - 1 + r0) - ry);
    wy[0] = quad_get<0x00>(a);
    wy[1] = quad_get<0x55>(a);
    wy[2] = quad_get<0xAA>(a);
    wy[3] = quad_get<0xFF>(a);
  } else if (l.tpm == 2) {
    const float a = phi((float)(yref - 1 + r0) - ry), b = phi((float)(yref + r0) - ry);
    const float pa = quad_get<kQuadXor1>(a), pb = quad_get<kQuadXor1>(b);
    wy[0] = r0 == 0 ? a : pa;
    wy[1] = r0 == 0 ? b : pb;
    wy[2] = r0 == 0 ? pa : a;
    wy[3] = r0 == 0 ? pb : b;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) wy[j] = phi((float)(yref - 1 + j) - ry);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) wy[j] = live ? wy[j] : 0.f;
}

// Compaction of the marked window cells: thread t counts the marks of cells
// [t C, t C + C), C = ceil(n / blockDim.x); a block-wide exclusive scan of the
// counts gives where its cells go in `list` (ascending cell order).  Returns
// the number of marked cells (every thread); ends with a barrier.
__device__ __forceinline__ int mark_scan(const uint8_t *mark, uint16_t *list, int n, int *s_scan) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nwave = (nt + 63) >> 6;
  const int C = (n + nt - 1) / nt, k0 = tid * C, k1 = min(k0 + C, n);
  int cnt = 0;
  for (int k = k0; k < k1; ++k) cnt += mark[k];
  int inc = cnt;   // inclusive scan over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = lane_get(inc, lane >= o ? lane - o : lane);
    if (lane >= o) inc += v;
  }
  if (lane == 63) s_scan[wave] = inc;
  __syncthreads();
  // wave totals: lane w holds wave w's; sums over the waves below this one and over all
  const int wt = lane < nwave ? s_scan[lane] : 0;
  int below = lane < wave ? wt : 0, total = wt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    below += lane_get(below, lane ^ o);
    total += lane_get(total, lane ^ o);
  }
  int off = inc - cnt + below;
  for (int k = k0; k < k1; ++k)
    if (mark[k]) list[off++] = (uint16_t)k;
  __syncthreads();
  return total;
}

// One workgroup per body (examples/2d/vortex_induced_vibration.py:99-137),
// with the spread as LDS atomics:
//   window u     macroscopic velocity of the post-streaming state of the window
//                cells (pull + walls; PULL = false: the state is read in place),
//                pulled by this workgroup; then *go_flag = gstep + 1 (the bulk
//                tiles of the fused step start streaming after it)
//   stencil      ib/stencil.py:7-51 at marker coordinates (x_m + d) - origin (4x4)
//   MDF          ib/mdf.py:35-64 with the spreads accumulated in one grid G:
//                ub = ub0 + 0.5 interp(G), force window = final G
//   body force   h = sum_m (-F_m) + a * area in FP64, dyn.py:126-136
//   structure    Newmark-beta in FP64 on thread 0, dyn.py:5-51
// Writes the force window to ib.g_win (st_agent) and the new body state to body_out.
// Where the IB workgroup takes the window velocity from.
constexpr int kWinPull = 0;    // pull from the post-collision state src (storage SRC) with walls
constexpr int kWinLocal = 1;   // read src in place (a post-streaming state)
constexpr int kWinHalo = 2;    // pull from the FP32 hand-off buffer `halo` [9][S + 4][S + 4] whose origin
                               // is the previous step's window origin - 2 (ib.win_origin[0..1])

template <class SRC, int WSRC, int UW, int TAG, bool SPLIT = false, int TPM = 0>
__device__ __forceinline__ void ib_atomic(const StepCtx &c, const VsxIB2D &ib, const typename SRC::T *__restrict__ src,
                                          const float *halo, int hidx, int gstep, int32_t *go_flag,
                                          const double *body_in, double *body_out, float *lds,
                                          int32_t *wait_flag = nullptr, int wait_target = 0) {
  const int S = ib.win, SS = S * S, M = ib.n_markers;
  __shared__ double s_body[6];
  __shared__ int s_org[6];
  __shared__ double s_red[2][kIbThreads / 64];
  __shared__ int s_scan[kIbThreads / 64 + 1];
  const int tid = threadIdx.x, nt = blockDim.x;
  const IbLane l = ib_lane<TPM>(M);
  const bool live = l.mk < M;

  // the first IB kernel of a pipelined run (START) waits for the run's first
  // bulk pass to begin, i.e. for the caller's earlier work on its stream, and
  // then reads what that work may have written (body state, the field) with
  // agent-scope loads
  constexpr bool START = SPLIT && WSRC == kWinLocal;
  if constexpr (START) {
    if (wait_flag) {
      if (tid == 0) spin_until(ib, wait_flag, wait_target);
      __syncthreads();
    }
  }
  // round trip 1: body state and marker inputs
  if (tid < 6)
    s_body[tid] = START ? __hip_atomic_load(const_cast<double *>(body_in + tid), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT)
                        : body_in[tid];
  double mx = 0.0, my = 0.0;
  float ds2 = 0.f;
  if (live) {
    mx = ib.marker_x[l.mk];
    my = ib.marker_y[l.mk];
    ds2 = (ib.ds_arr ? ib.ds_arr[l.mk] : ib.ds) * 2.0f;
  }
  __syncthreads();
  if (tid == 0) {
    int o[4];
    window_origin(c, ib, s_body, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) s_org[k] = o[k];
    if constexpr (WSRC == kWinHalo) {
      // the hand-off buffer covers the previous window +- 2 cells: this window
      // must lie within 1 cell of it and 2 cells inside the grid
      const int ex = ib.win_origin[0] - 2, ey = ib.win_origin[1] - 2;
      const bool ok = o[2] - ex >= 1 && o[2] - ex <= 3 && o[3] - ey >= 1 && o[3] - ey <= 3 && o[2] >= 2 &&
                      o[3] >= 2 && o[2] + S + 2 <= c.dims[0] && o[3] + S + 2 <= c.dims[1];
      if (!ok) set_ib_error(ib, 2);
      s_org[4] = ex;
      s_org[5] = ey;
    }
    // agent-scope stores: read by the next IB kernel and by the caller's stream,
    // which in the pipelined form is not ordered after this kernel's end
    __hip_atomic_store(&ib.win_origin[0], o[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ib.win_origin[1], o[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // LDS: window velocity (2 SS words) and accumulation grid (4 SS words), the
  // grid over the velocity (zeroed once the velocity is dead) or, SPLIT,
  // behind it (zeroed before the wait); the mark array (one byte per window
  // cell) and the cell list behind the velocity (over the grid's start) or,
  // SPLIT, behind the grid
  float *acc = SPLIT ? lds + 2 * SS : lds;
  uint8_t *mark = reinterpret_cast<uint8_t *>(lds + (SPLIT ? 6 : 2) * SS);
  uint16_t *list = reinterpret_cast<uint16_t *>(lds + (SPLIT ? 6 : 2) * SS + (SS + 3) / 4);
  for (int k = tid; k < (SS + 3) / 4; k += nt) reinterpret_cast<uint32_t *>(mark)[k] = 0u;
  __syncthreads();
  const int x0 = s_org[0], y0 = s_org[1], xc = s_org[2], yc = s_org[3];

  // marker position relative to the window origin; mark the window cells the
  // interpolation of this thread's rows reads (zero weights included)
  float rx = 0.f, ry = 0.f;
  if (live) {
    rx = (float)(mx + s_body[4] - (double)x0);
    ry = (float)(my + s_body[5] - (double)y0);
    const int b0 = ((int)floorf(rx) - 1 + l.first_row()) * S + ((int)floorf(ry) - 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < l.rows) {
#pragma unroll
        for (int j = 0; j < 4; ++j) mark[gather_index32(b0 + i * S + j, SS)] = 1;
      }
    }
  }
  __syncthreads();
  // compaction: thread t owns cells [t C, t C + C); block-wide exclusive scan of the counts
  const int n_list = mark_scan(mark, list, SS, s_scan);

  // stencil of this thread's marker and rows: SPLIT before the wait, else after
  // the pull (whose registers it would take)
  float wr[4], wy[4];
  int base = 0;
  bool inside = true;
  auto stencil = [&]() {
    if (ib.kernel == VSX_KERNEL_PESKIN_4PT) ib_weights(peskin4, l, live, rx, ry, wr, wy);
    else ib_weights([&](float r) { return delta_kernel(ib.kernel, r); }, l, live, rx, ry, wr, wy);
    // flat window index of this thread's first stencil point, and whether every
    // stencil index of the wave lies inside the window
    base = live ? ((int)floorf(rx) - 1 + l.first_row()) * S + ((int)floorf(ry) - 1) : 0;
    inside = __all(base >= 0 && base + (l.rows - 1) * S + 3 < SS);
  };
  if constexpr (SPLIT) {
    stencil();
    for (int k = tid; k < SS; k += nt) reinterpret_cast<float4 *>(acc)[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  // the hand-off of the window populations (SPLIT: the zeroing above is
  // ordered before the spread by the barriers below)
  if (!START && wait_flag) {
    if (tid == 0) spin_until(ib, wait_flag, wait_target);
    __syncthreads();
  }

  // round trip 2: window velocity of the marked cells (float2 per cell over the
  // start of the accumulation grid; the other cells are never read)
  float2 *ugrid = reinterpret_cast<float2 *>(lds);
  for (int idx = tid; idx < n_list; idx += nt) {
    const int k = list[idx];
    float f[9], rho, u[3];
    if constexpr (WSRC == kWinHalo) {
      // a window outside the hand-off buffer (an error, flagged above) reads
      // the buffer's corner instead: every load stays in bounds and all nine
      // are issued back to back
      const int E = S + 4, EE = E * E;
      const int dx = min(max(xc - s_org[4], 1), 3), dy = min(max(yc - s_org[5], 1), 3);
      const int i = k / S + dx, j = k % S + dy;
#pragma unroll
      for (int q = 0; q < 9; ++q) f[q] = ld_agent(&halo[q * EE + (i - D2Q9::cx[q]) * E + (j - D2Q9::cy[q])]);
    } else {
      const int co[3] = {xc + k / S, yc + k % S, 0};
      // cells on a domain wall are left to the loop below
      if (WSRC == kWinPull && c.p.n_walls > 0 && wall_class<D2Q9>(c, co) >= 0) continue;
      if constexpr (WSRC == kWinPull) {
        pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f);
      } else if constexpr (START) {
        const float *p = reinterpret_cast<const float *>(src) + cell_of<D2Q9>(c, co[0], co[1], 0);
#pragma unroll
        for (int q = 0; q < 9; ++q) f[q] = ld_agent(p + (size_t)q * c.n);
      } else {
        local<D2Q9, SRC>(c, src, cell_of<D2Q9>(c, co[0], co[1], 0), f);
      }
    }
    macroscopic<D2Q9>(f, rho, u);
    ugrid[k] = make_float2(u[0], u[1]);
  }
  if (WSRC == kWinPull && c.p.n_walls > 0) {
    // marked cells on a domain wall (only when the window touches a wall;
    // fused steps have no wall corners, so every wall cell is of a
    // single-wall class), in a loop of their own: inline in the loop above
    // the wall code's registers would add to the pull's
    for (int idx = tid; idx < n_list; idx += nt) {
      const int k = list[idx];
      const int co[3] = {xc + k / S, yc + k % S, 0};
      const int cls = wall_class<D2Q9>(c, co);
      if (cls >= 0) {
        float f[9], rho, u[3];
        single_wall_state<D2Q9, SRC>(c, src, co, cls, f);
        macroscopic<D2Q9>(f, rho, u);
        ugrid[k] = make_float2(u[0], u[1]);
      }
    }
  }
  __syncthreads();
  if (tid == 0 && go_flag) __hip_atomic_store(go_flag, gstep + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

  if constexpr (!SPLIT) stencil();

  // initial interpolation of the window velocity
  float ub0x = 0.f, ub0y = 0.f;
  if (live) {
    const float2 v = interp_rows<false>(lds, base, S, l, wr, wy, inside);
    ub0x = v.x;
    ub0y = v.y;
  }
  float ubx = ub0x, uby = ub0y, Fx = 0.f, Fy = 0.f;
  const float Ux = (float)s_body[2], Uy = (float)s_body[3];
  if constexpr (!SPLIT) {
    __syncthreads();   // the window velocity is dead from here; the grid accumulates the spread force
    for (int k = tid; k < SS; k += nt) reinterpret_cast<float4 *>(acc)[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
  }

  // multi-direct forcing (ib/mdf.py:45-64)
  for (int it = 0; it < ib.n_iter; ++it) {
    const float sfx = (Ux - ubx) * ds2, sfy = (Uy - uby) * ds2;
    Fx += sfx;
    Fy += sfy;
    if (live) spread_rows(acc, base, S, l, wr, wy, sfx, sfy, inside);
    __syncthreads();
    if (it + 1 == ib.n_iter) break;
    if (live) {
      const float2 v = interp_rows<true>(acc, base, S, l, wr, wy, inside);
      ubx = ub0x + 0.5f * v.x;
      uby = ub0y + 0.5f * v.y;
    }
    __syncthreads();   // every interpolation has read G before the next spread adds to it
  }

  // force window -> global
  for (int k = tid; k < SS; k += nt) {
    const float2 g = acc_get(acc, SS, k);
    st_agent(&ib.g_win[k], g.x);
    st_agent(&ib.g_win[SS + k], g.y);
  }

  // body force h = sum(-F) (FP64), one thread per marker
  double hx = 0.0, hy = 0.0;
  if (live && l.first_row() == 0) {
    hx = -(double)Fx;
    hy = -(double)Fy;
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
    for (int k = 0; k < 6; ++k) __hip_atomic_store(&body_out[k], out[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ib.hist_d && hidx >= 0) {
      __hip_atomic_store(&ib.hist_d[2 * hidx], out[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ib.hist_d[2 * hidx + 1], out[5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (ib.hist_h && hidx >= 0) {
      __hip_atomic_store(&ib.hist_h[2 * hidx], h[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ib.hist_h[2 * hidx + 1], h[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Publish "the force window of this step is written" (its st_agent stores
// have completed in every thread) by storing `value` (the global step index
// + 1) to the ready flag.
__device__ __forceinline__ void publish_ready(const VsxIB2D &ib, int value) {
  wait_mem();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&ib.win_origin[2], value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait of a window tile until the ready flag reaches `target` (bounded, see
// spin_until); the force window is then read with ld_agent.
__device__ __forceinline__ void wait_ready(const VsxIB2D &ib, int target) {
  if (threadIdx.x == 0) spin_until(ib, &ib.win_origin[2], target);
  __syncthreads();
}

// Window tile: store the post-streaming velocity of a cell that lies in the
// force window into ib.g_win (read by the IB workgroup) ...
__device__ __forceinline__ void store_window_u(const VsxIB2D &ib, bool valid, const int (&co)[3], const int (&org)[3],
                                               const float (&f)[9]) {
  const int S = ib.win;
  const int r0 = co[0] - org[0], r1 = co[1] - org[1];
  if (valid && r0 >= 0 && r0 < S && r1 >= 0 && r1 < S) {
    float rho, u[3];
    macroscopic<D2Q9>(f, rho, u);
    st_agent(&ib.g_win[r0 * S + r1], u[0]);
    st_agent(&ib.g_win[S * S + r0 * S + r1], u[1]);
  }
}

// ... then, once every thread's stores have completed, count the tile in
// ib.win_origin[4 + (gstep & 1)]; the last of the n_wtiles window tiles also
// raises ib.win_origin[6] to gstep + 1 (the bulk tiles start their own
// traffic after it, so the window pulls are served first).
__device__ __forceinline__ void count_window_tile(const VsxIB2D &ib, int gstep, int n_wtiles) {
  wait_mem();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int before =
        __hip_atomic_fetch_add(&ib.win_origin[4 + (gstep & 1)], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (before == n_wtiles - 1)
      __hip_atomic_store(&ib.win_origin[6], gstep + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Pipelined form: once every thread's hand-off stores have completed, count
// the tile in ib.win_origin[4]; the last of the n hand-off tiles of step gstep
// resets the counter (the next step's tiles run in a later launch of the same
// stream) and raises ib.win_origin[6] to gstep + 1.
__device__ __forceinline__ void count_halo_tile(const VsxIB2D &ib, int gstep, int n) {
  wait_mem();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int before = __hip_atomic_fetch_add(&ib.win_origin[4], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (before == n - 1) {
      __hip_atomic_store(&ib.win_origin[4], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ib.win_origin[6], gstep + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
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
  VSX_REQUIRE(ib->fused >= 0 && ib->fused <= 2, "fused must be 0, 1 or 2, got %d", ib->fused);
  VSX_REQUIRE(ib->marker_x && ib->marker_y && ib->body && ib->g_win && ib->win_origin, "null IB buffer");
  VSX_REQUIRE(ib->mass + ib->gamma * ib->dt * ib->damping + ib->beta * ib->dt * ib->dt * ib->stiffness != 0.0,
              "singular Newmark system");
  int max_lds = 0;
  if (int rc_ = max_lds_per_block(&max_lds)) return rc_;
  const size_t need = 4 * (size_t)ib_lds_words(ib->win, ib->n_markers, ib->n_markers <= kIbThreads) + 2048;
  VSX_REQUIRE(need <= (size_t)max_lds, "IB window %d needs %zu B of LDS, device allows %d", ib->win, need, max_lds);
  if (ib->fused == 2) {
    const size_t need_pipe = 4 * (size_t)ib_split_lds_words(ib->win) + 2048;
    VSX_REQUIRE(need_pipe <= (size_t)max_lds, "pipelined IB kernel: window %d needs %zu B of LDS, device allows %d",
                ib->win, need_pipe, max_lds);
  }
  return VSX_OK;
}

}  // namespace vsx
