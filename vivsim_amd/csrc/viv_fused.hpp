// The fused 2D VIV step: one launch per reference time step.
//
// Blocks: block 0 is the IB workgroup (ib_atomic); blocks 1 .. n_win are the
// tiles that intersect the force window; the remaining blocks are bulk
// tiles.  Every block derives the step's window origin from the body state
// before the step (slot gstep & 1), so the step needs no host sync.
//   IB workgroup  pulls the window's post-streaming velocity itself, raises
//                 ib.win_origin[6] (the bulk tiles start streaming after it,
//                 so the window pulls are not queued behind the whole grid's
//                 traffic), runs the multi-direct forcing, writes the force
//                 window with st_agent and raises the ready flag
//                 ib.win_origin[2]
//   window tiles  pull their cells, wait (bounded) for the ready flag, read the
//                 force window with ld_agent, collide
//   bulk tiles    wait (bounded) for ib.win_origin[6], then stream and collide
// Pipelined form (ib.fused == 2): there is no IB block; the IB of step t + 1
// runs as its own kernel (k_ib_pipe) on a second stream while the bulk pass of
// step t streams.  The tiles of step t that cover the window +- 2 cells come
// first; they write their new populations (as read back from storage) into
// the FP32 hand-off buffer ib.halo (two halves, by step parity) with
// st_agent, count themselves in
// ib.win_origin[4], and the last one raises ib.win_origin[6]; the IB kernel of
// step t + 1 waits for that flag and pulls the window from the buffer.  The
// force-window tiles wait for the IB kernel's ready flag as in the one-launch
// form; no other tile waits.
// Modes (the post-collision state is what the work buffers hold):
//   kModeFull   src = post-collision state of the previous step (pull + walls)
//   kModeFirst  src = the caller's post-streaming FP32 field, read in place (the
//               reference state a run starts from: no pull, walls already applied)
//   kModeLast   the new post-collision populations are pushed to their
//               streaming destinations in an FP32 field (the caller's output);
//               the walls are applied afterwards by k_wall_line, wall by wall
#pragma once
#include "viv2d.hpp"

namespace vsx {

constexpr int kModeFull = 0, kModeFirst = 1, kModeLast = 2;
// poll interval (x 64 cycles) of the bulk tiles waiting for the IB workgroup's window pull
constexpr int kBulkSleep = 16;

// Push-order store of a collided cell: f_i goes to x + c_i (periodic), which is
// the reference's streamed field before the walls.  Destination cells are
// (row offset of x + cx) + (y + cy): six 32-bit values for the nine stores.
__device__ __forceinline__ void store_push2(const StepCtx &c, float *__restrict__ dst, const int (&co)[3],
                                            const float (&f)[9]) {
  const uint32_t ny = (uint32_t)c.dims[1];
  const uint32_t xr[3] = {(uint32_t)wrap(co[0] - 1, c.dims[0]) * ny, (uint32_t)co[0] * ny,
                          (uint32_t)wrap(co[0] + 1, c.dims[0]) * ny};
  const uint32_t yr[3] = {(uint32_t)wrap(co[1] - 1, c.dims[1]), (uint32_t)co[1], (uint32_t)wrap(co[1] + 1, c.dims[1])};
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    float *p = dst + (size_t)i * c.n;   // uniform base + 32-bit cell offset
    p[xr[D2Q9::cx[i] + 1] + yr[D2Q9::cy[i] + 1]] = f[i];
  }
}

// Post-streaming, post-walls populations of a tile cell (kModeFirst: in place).
template <class SRC, int MODE>
__device__ __forceinline__ void tile_cell_state(const StepCtx &c, const typename SRC::T *__restrict__ src,
                                                const int (&co)[3], float (&f)[9]) {
  if constexpr (MODE == kModeFirst) {
    local<D2Q9, SRC>(c, src, cell_of<D2Q9>(c, co[0], co[1], 0), f);
  } else {
    const int cls = c.p.n_walls > 0 ? wall_class<D2Q9>(c, co) : -1;
    if (cls >= 0) single_wall_state<D2Q9, SRC>(c, src, co, cls, f);
    else pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f);
  }
}

template <class DST, int MODE>
__device__ __forceinline__ void tile_store(const StepCtx &c, typename DST::T *__restrict__ dst, const int (&co)[3],
                                           const float (&f)[9]) {
  if constexpr (MODE == kModeLast) store_push2(c, dst, co, f);
  else store_cell<D2Q9, DST>(c, dst, cell_of<D2Q9>(c, co[0], co[1], 0), f);
}

// Two bulk cells per thread (16-bit storages, kModeFull): both cells'
// populations are loaded before either is computed, so a wave keeps as many
// bytes in flight as with one FP32 cell per thread.  a and b are interior
// cells (no walls) outside the IB window, in a step without a g field or
// uniform g.
template <class T>
struct Pack2;
template <>
struct Pack2<_Float16> {
  typedef _Float16 type __attribute__((ext_vector_type(2)));
};
template <>
struct Pack2<uint16_t> {
  typedef uint16_t type __attribute__((ext_vector_type(2)));
};

template <class SRC, class DST, int COLL>
__device__ __forceinline__ void bulk_pair(const StepCtx &c, const typename SRC::T *__restrict__ src,
                                          typename DST::T *__restrict__ dst, const int (&a)[3], const int (&b)[3]) {
  // the two cells' values of population i share one VGPR (low / high half)
  typename Pack2<typename SRC::T>::type r[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const typename SRC::T *p = src + (size_t)i * c.n;
    r[i] = {p[cell_of<D2Q9>(c, wrap(a[0] - D2Q9::cx[i], c.dims[0]), wrap(a[1] - D2Q9::cy[i], c.dims[1]), 0)],
            p[cell_of<D2Q9>(c, wrap(b[0] - D2Q9::cx[i], c.dims[0]), wrap(b[1] - D2Q9::cy[i], c.dims[1]), 0)]};
  }
  float f[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = SRC::cvt(r[i].x, c.shift[i]);
  const uint32_t ca = cell_of<D2Q9>(c, a[0], a[1], 0);
  finish_unforced<D2Q9, COLL>(c, f);
  store_cell<D2Q9, DST>(c, dst, ca, f);
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = SRC::cvt(r[i].y, c.shift[i]);
  const uint32_t cb = cell_of<D2Q9>(c, b[0], b[1], 0);
  finish_unforced<D2Q9, COLL>(c, f);
  store_cell<D2Q9, DST>(c, dst, cb, f);
}

// One bulk cell (any storage, walls included).
template <class SRC, class DST, int COLL, int MODE>
__device__ __forceinline__ void bulk_cell(const StepCtx &c, const typename SRC::T *__restrict__ src,
                                          typename DST::T *__restrict__ dst, const int (&co)[3], const int (&org)[3]) {
  float f[9];
  tile_cell_state<SRC, MODE>(c, src, co, f);
  finish<D2Q9, COLL>(c, cell_of<D2Q9>(c, co[0], co[1], 0), co, org, false, f);
  tile_store<DST, MODE>(c, dst, co, f);
}

// With PAIR (kModeFull only): block 0 the IB workgroup, blocks 1..n_win the
// window tiles, then one block per two bulk tiles; the host launches
// 1 + ceil((tiles + max window tiles) / 2) blocks and the surplus exits.
// PIPE: the pipelined form (ib.fused == 2) as a template argument, so that its
// kernels carry no IB workgroup code (which changes their register allocation).
template <class SRC, class DST, int COLL, bool PAIR, int MODE, bool PIPE>
__global__ __attribute__((amdgpu_flat_work_group_size(kTileCells, kTileCells), amdgpu_waves_per_eu(8, 8))) void
k_viv_fused(StepCtx c_arg, VsxIB2D ib, const typename SRC::T *__restrict__ src, typename DST::T *__restrict__ dst,
            int step, Tiles T) {
  static_assert(!PAIR || MODE == kModeFull, "paired bulk blocks only in full steps");
  const StepCtx &c = kernel_ctx();
  extern __shared__ float lds[];
  const int gstep = ib.step0 + step, slot = gstep & 1;
  const double *body_in = ib.body + 6 * slot;
  constexpr bool pipe = PIPE;
  if constexpr (MODE == kModeFirst) {
    // pipelined run: the caller's earlier work on this stream is complete; let
    // the run's first IB kernel start
    if (pipe && blockIdx.x == 0 && threadIdx.x == 0)
      __hip_atomic_store(&ib.win_origin[7], gstep + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if constexpr (!pipe) {
    if (blockIdx.x == 0) {
      ib_atomic<SRC, MODE == kModeFirst ? kWinLocal : kWinPull, 2, 1>(c, ib, src, nullptr, step, gstep,
                                                                       &ib.win_origin[6], body_in,
                                                                       ib.body + 6 * (slot ^ 1), lds);
      publish_ready(ib, gstep + 1);
      return;
    }
  }
  __shared__ int s_org[2];
  if (threadIdx.x == 0) {
    // one-launch form: the body state before this step was written by the
    // previous launch (plain loads); pipelined form: by the IB kernel of the
    // previous step on the other stream (agent-scope loads)
    double b[6];
    if (pipe) {
      b[4] = __hip_atomic_load(const_cast<double *>(body_in + 4), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      b[5] = __hip_atomic_load(const_cast<double *>(body_in + 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      b[4] = body_in[4];
      b[5] = body_in[5];
    }
    int o[4];
    window_origin(c, ib, b, o);
    s_org[0] = o[2];
    s_org[1] = o[3];
  }
  __syncthreads();
  const int org[3] = {s_org[0], s_org[1], 0};
  // first tiles: the force window's (one-launch form) or the window +- 2 cells (pipelined form)
  int wx0, wx1, wy0, wy1;
  if (pipe) {
    // clamped to the grid (a window closer than 2 cells to the edge is an error the IB kernel reports)
    wx0 = max(org[0] - 2, 0) / T.TX;
    wx1 = min(org[0] + ib.win + 1, c.dims[0] - 1) / T.TX;
    wy0 = max(org[1] - 2, 0) / T.TY;
    wy1 = min(org[1] + ib.win + 1, c.dims[1] - 1) / T.TY;
  } else {
    window_tiles(T, org[0], org[1], ib.win, wx0, wx1, wy0, wy1);
  }
  const int bid = pipe ? (int)blockIdx.x : (int)blockIdx.x - 1;   // tile-block index
  const int n_win = (wx1 - wx0 + 1) * (wy1 - wy0 + 1);
  if constexpr (PAIR) {
    if (bid >= n_win) {
      const int q = n_win + 2 * (bid - n_win), n_tiles = T.ntx * T.nty;
      if (q >= n_tiles) return;
      int ta[2], tb[2];
      bool w_;
      tile_of_block(q, T, wx0, wx1, wy0, wy1, ta[0], ta[1], w_);
      const bool has_b = q + 1 < n_tiles;
      if (has_b) tile_of_block(q + 1, T, wx0, wx1, wy0, wy1, tb[0], tb[1], w_);
      const int a[3] = {ta[0] * T.TX + (int)threadIdx.x / T.TY, ta[1] * T.TY + (int)threadIdx.x % T.TY, 0};
      const int b[3] = {has_b ? tb[0] * T.TX + (int)threadIdx.x / T.TY : c.dims[0],
                        has_b ? tb[1] * T.TY + (int)threadIdx.x % T.TY : c.dims[1], 0};
      if (!pipe) {
        if (threadIdx.x == 0) spin_until<kBulkSleep>(ib, &ib.win_origin[6], gstep + 1);
        __syncthreads();
      }
      const bool va = a[0] < c.dims[0] && a[1] < c.dims[1], vb = b[0] < c.dims[0] && b[1] < c.dims[1];
      const bool wa = va && c.p.n_walls > 0 && wall_class<D2Q9>(c, a) >= 0;
      const bool wb = vb && c.p.n_walls > 0 && wall_class<D2Q9>(c, b) >= 0;
      const bool unforced = !c.p.g && c.p.g_uniform[0] == 0.f && c.p.g_uniform[1] == 0.f;
      if (unforced && va && vb && !wa && !wb) {
        bulk_pair<SRC, DST, COLL>(c, src, dst, a, b);
      } else {
        if (va) bulk_cell<SRC, DST, COLL, MODE>(c, src, dst, a, org);
        if (vb) bulk_cell<SRC, DST, COLL, MODE>(c, src, dst, b, org);
      }
      return;
    }
  }
  int tx, ty;
  bool in_win;
  tile_of_block(bid, T, wx0, wx1, wy0, wy1, tx, ty, in_win);
  const int co[3] = {tx * T.TX + (int)threadIdx.x / T.TY, ty * T.TY + (int)threadIdx.x % T.TY, 0};
  const bool valid = co[0] < c.dims[0] && co[1] < c.dims[1];
  float f[9];
  if (pipe && in_win) {
    // tile around the window (pipelined form): collide (after the force window
    // when the tile intersects it), store, hand the cells within 2 of the window
    // to the next step's IB kernel
    const int S = ib.win;
    const bool force = tx * T.TX < org[0] + S && (tx + 1) * T.TX > org[0] && ty * T.TY < org[1] + S &&
                       (ty + 1) * T.TY > org[1];
    if (valid) tile_cell_state<SRC, MODE>(c, src, co, f);
    if (force) wait_ready(ib, gstep + 1);
    if (valid) {
      finish<D2Q9, COLL>(c, cell_of<D2Q9>(c, co[0], co[1], 0), co, org, force, f);
      tile_store<DST, MODE>(c, dst, co, f);
    }
    if constexpr (MODE != kModeLast) {
      // hand-off buffer of this step's parity (the next step's IB reads it while the
      // next bulk pass writes the other one)
      const int E = S + 4, EE = E * E, a = co[0] - (org[0] - 2), b = co[1] - (org[1] - 2);
      float *h = ib.halo + (size_t)(gstep & 1) * 9 * EE;
      if (valid && a >= 0 && a < E && b >= 0 && b < E) {
#pragma unroll
        for (int q = 0; q < 9; ++q) st_agent(&h[q * EE + a * E + b], DST::rt(f[q], c.shift[q]));
      }
      count_halo_tile(ib, gstep, n_win);
    }
    return;
  }
  if (in_win) {
    // window tile: post-streaming state -> wait for the IB workgroup's force window -> collide
    if (valid) tile_cell_state<SRC, MODE>(c, src, co, f);
    wait_ready(ib, gstep + 1);
    if (!valid) return;
    finish<D2Q9, COLL>(c, cell_of<D2Q9>(c, co[0], co[1], 0), co, org, true, f);
    tile_store<DST, MODE>(c, dst, co, f);
    return;
  }
  // bulk tile (one-launch form): wait (bounded) until the IB workgroup has pulled the window
  if (!pipe) {
    if (threadIdx.x == 0) spin_until<kBulkSleep>(ib, &ib.win_origin[6], gstep + 1);
    __syncthreads();
  }
  if (!valid) return;
  if constexpr (MODE != kModeFirst) {
    const int cls = c.p.n_walls > 0 ? wall_class<D2Q9>(c, co) : -1;
    if (cls >= 0) {
      // wall cell: own path, keeping the wall code's registers out of the interior path
      single_wall_state<D2Q9, SRC>(c, src, co, cls, f);
      finish<D2Q9, COLL>(c, cell_of<D2Q9>(c, co[0], co[1], 0), co, org, false, f);
      tile_store<DST, MODE>(c, dst, co, f);
      return;
    }
    pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f);
  } else {
    local<D2Q9, SRC>(c, src, cell_of<D2Q9>(c, co[0], co[1], 0), f);
  }
  finish<D2Q9, COLL>(c, cell_of<D2Q9>(c, co[0], co[1], 0), co, org, false, f);
  tile_store<DST, MODE>(c, dst, co, f);
}

// IB kernel of the pipelined form: step gstep's IB, pulling the window from the
// hand-off buffer (written by the previous step's bulk pass, WSRC = kWinHalo;
// it waits for that pass's flag) or, for the first step of a run, reading the
// caller's post-streaming FP32 field in place (kWinLocal).  One block of 1024
// threads; it runs on a compute unit of its own (the IB stream's CU mask), so
// its LDS and registers do not constrain the bulk kernel's occupancy.
template <int WSRC, int TPM>
__global__ __launch_bounds__(kIbThreads) void k_ib_pipe(StepCtx c_arg, VsxIB2D ib, const float *__restrict__ src,
                                                       int step) {
  const StepCtx &c = kernel_ctx();
  extern __shared__ float lds[];
  const int gstep = ib.step0 + step, slot = gstep & 1;
  const float *halo = ib.halo + (size_t)((gstep - 1) & 1) * 9 * (ib.win + 4) * (ib.win + 4);
  // everything up to the window pull (body state, marks, cell list, stencil,
  // zeroed grid) is done while the previous step's bulk pass still streams;
  // the pull waits for that pass's hand-off flag
  // (the first IB kernel of a run waits for the run's first bulk pass to start:
  // ib.win_origin[7] = step0 + 1)
  ib_atomic<Store<VSX_STORE_F32>, WSRC, 2, 2, true, TPM>(c, ib, src, halo, step, gstep, nullptr, ib.body + 6 * slot,
                                                     ib.body + 6 * (slot ^ 1), lds,
                                                     WSRC == kWinHalo ? &ib.win_origin[6] : &ib.win_origin[7],
                                                     WSRC == kWinHalo ? gstep : gstep + 1);
  publish_ready(ib, gstep + 1);
}

// One wall applied in place to a post-streaming FP32 field (the reference's
// boundary_* call on the whole array): thread j takes wall node j, reads the
// wall cell and, where the scheme needs it, the inward neighbour (both already
// carry the walls applied before this one), and rewrites the wall cell.  The
// wall must not need the pre-streaming populations.
template <class L>
__global__ __launch_bounds__(256) void k_wall_line(StepCtx c_arg, float *__restrict__ f, int k0) {
  const StepCtx &c = kernel_ctx();
  const int k = k0 + (int)blockIdx.y;   // one wall per grid row when the walls are independent
  const VsxWall &w = c.p.walls[k];
  const int a = L::w_axis[w.loc];
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int co[3] = {0, 0, 0};
  if constexpr (L::D == 2) {
    if (j >= c.dims[1 - a]) return;
    co[1 - a] = (int)j;
  } else {
    const int b0 = a == 0 ? 1 : 0, b1 = a == 2 ? 1 : 2;
    if (j >= (int64_t)c.dims[b0] * c.dims[b1]) return;
    co[b0] = (int)(j / c.dims[b1]);
    co[b1] = (int)(j % c.dims[b1]);
  }
  co[a] = L::w_sign[w.loc] > 0 ? 0 : c.dims[a] - 1;
  float fw[L::Q], fn[L::Q], fp[L::Q];
  const uint32_t cell = cell_of<L>(c, co[0], co[1], co[2]);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    fw[i] = f[(size_t)i * c.n + cell];
    fn[i] = fp[i] = 0.f;
  }
  if (wall_needs_neighbor(w)) {
    int cn[3];
    inward<L>(w.loc, co, cn);
    const uint32_t nb = cell_of<L>(c, cn[0], cn[1], cn[2]);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) fn[i] = f[(size_t)i * c.n + nb];
  }
  apply_wall<L>(w, wall_node<L>(c, w.loc, co), fw, fn, fp);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) f[(size_t)i * c.n + cell] = fw[i];
}

// Walls of a pushed field, in list order (one launch per wall).  Walls on
// distinct faces of one axis, with at least 4 cells between them, neither
// share a cell nor read one another's cells: then all go in one launch.
template <class L>
int walls_in_place(const StepCtx &c, float *f, hipStream_t s) {
  const int n = c.p.n_walls;
  if (n == 0) return VSX_OK;
  const int a0 = L::w_axis[c.p.walls[0].loc];
  bool together = c.dims[a0] >= 4;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < k; ++j)
      together = together && L::w_axis[c.p.walls[k].loc] == a0 && c.p.walls[k].loc != c.p.walls[j].loc;
  int64_t nodes = 1;
  for (int b = 0; b < L::D; ++b)
    if (b != a0) nodes *= c.dims[b];
  const unsigned gx = (unsigned)((nodes + 255) / 256);
  if (together) {
    hipLaunchKernelGGL((k_wall_line<L>), dim3(gx, (unsigned)n), dim3(256), 0, s, c, f, 0);
  } else {
    for (int k = 0; k < n; ++k) {
      const int a = L::w_axis[c.p.walls[k].loc];
      int64_t nk = 1;
      for (int b = 0; b < L::D; ++b)
        if (b != a) nk *= c.dims[b];
      hipLaunchKernelGGL((k_wall_line<L>), dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, c, f, k);
    }
  }
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

// Whether a run can end with a pushed step: no wall reads pre-streaming populations.
inline bool walls_allow_push(const VsxLbm &p) {
  for (int k = 0; k < p.n_walls; ++k)
    if (wall_needs_pre(p.walls[k])) return false;
  return true;
}

// ------------------------------------------------------------------ launches
template <class B>
struct IsStreamed : std::false_type {};
template <class B>
struct IsStreamed<Streamed<B>> : std::true_type {};

Tiles make_tiles(const StepCtx &c);

// Bulk blocks take two tiles (PAIR) for the 16-bit storages on grids of at
// least kPairMinTiles tiles: there the streaming bounds the step and the pairs
// double the bytes in flight.  On smaller grids single tiles keep the step
// shorter; FP32 tiles stay single (no gain at 4096^2, a loss at 1024^2).
constexpr int kPairMinTiles = 4096;

template <class SRC, class DST, int COLL, bool PAIR, int MODE>
int launch_fused_p(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s,
                   const Tiles &T) {
  const bool pipe = ib.fused == 2;
  const size_t lds = pipe ? 0 : 4 * (size_t)ib_atomic_lds_words(ib.win);
  const int n_tiles = T.ntx * T.nty;
  int grid = n_tiles;
  if (PAIR) {
    const int side = pipe ? ib.win + 4 : ib.win;
    const int nw_max = ((side + T.TX - 1) / T.TX + 1) * ((side + T.TY - 1) / T.TY + 1);
    grid = (n_tiles + nw_max + 1) / 2;
    if (grid > n_tiles) grid = n_tiles;
  }
  timer_begin(s);
  if (pipe)
    hipLaunchKernelGGL((k_viv_fused<SRC, DST, COLL, PAIR, MODE, true>), dim3(grid), dim3(kTileCells), lds, s, c, ib,
                       static_cast<const typename SRC::T *>(src), static_cast<typename DST::T *>(dst), step, T);
  else if constexpr (!IsStreamed<DST>::value)
    hipLaunchKernelGGL((k_viv_fused<SRC, DST, COLL, PAIR, MODE, false>), dim3(1 + grid), dim3(kTileCells), lds, s, c,
                       ib, static_cast<const typename SRC::T *>(src), static_cast<typename DST::T *>(dst), step, T);
  else
    return fail(VSX_ERR_VALUE, "non-temporal field stores are built for the pipelined form only");
  timer_end(s);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

template <class SRC, class DST, int COLL, int MODE>
int launch_fused_c(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  const Tiles T = make_tiles(c);
  if constexpr (MODE == kModeFull) {
    // full steps of the pipelined form whose work buffer alone exceeds the
    // last-level cache: the field stores are non-temporal (such a 16-bit field
    // has more than kPairMinTiles tiles, so its bulk blocks take tile pairs).
    // Buffers that exceed the cache only together (2048^2 FP32, 2 x 151 MB)
    // still gain from plain stores.
    if (ib.fused == 2 && (size_t)c.n * 9 * sizeof(typename DST::T) > kLastLevelCacheBytes)
      return launch_fused_p<SRC, Streamed<DST>, COLL, sizeof(typename SRC::T) == 2, MODE>(c, ib, src, dst, step, s, T);
  }
  if constexpr (MODE == kModeFull && sizeof(typename SRC::T) == 2) {
    if (T.ntx * T.nty >= kPairMinTiles) return launch_fused_p<SRC, DST, COLL, true, MODE>(c, ib, src, dst, step, s, T);
  }
  return launch_fused_p<SRC, DST, COLL, false, MODE>(c, ib, src, dst, step, s, T);
}

// storage ST of the work buffers; kModeFirst reads FP32, kModeLast writes FP32
template <class ST, int MODE>
int launch_fused_mode(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  using F32 = Store<VSX_STORE_F32>;
  using SRC = std::conditional_t<MODE == kModeFirst, F32, ST>;
  using DST = std::conditional_t<MODE == kModeLast, F32, ST>;
  switch (c.p.collision) {
    case VSX_COLL_BGK: return launch_fused_c<SRC, DST, VSX_COLL_BGK, MODE>(c, ib, src, dst, step, s);
    case VSX_COLL_MRT: return launch_fused_c<SRC, DST, VSX_COLL_MRT, MODE>(c, ib, src, dst, step, s);
    case VSX_COLL_KBC: return launch_fused_c<SRC, DST, VSX_COLL_KBC, MODE>(c, ib, src, dst, step, s);
    default: return launch_fused_c<SRC, DST, VSX_COLL_REG, MODE>(c, ib, src, dst, step, s);
  }
}

template <int MODE>
int fused_mode_any(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  switch (c.p.storage) {
    case VSX_STORE_F32: return launch_fused_mode<Store<VSX_STORE_F32>, MODE>(c, ib, src, dst, step, s);
    case VSX_STORE_F16S: return launch_fused_mode<Store<VSX_STORE_F16S>, MODE>(c, ib, src, dst, step, s);
    case VSX_STORE_BF16S: return launch_fused_mode<Store<VSX_STORE_BF16S>, MODE>(c, ib, src, dst, step, s);
    default: return launch_fused_mode<Store<VSX_STORE_E3M12S>, MODE>(c, ib, src, dst, step, s);
  }
}

// IB kernel of step `step` of the pipelined form (first: read the caller's field src in place)
inline int launch_ib_pipe(const StepCtx &c, const VsxIB2D &ib, const float *src, bool first, int step,
                          hipStream_t s) {
  const size_t lds = 4 * (size_t)ib_split_lds_words(ib.win);
  const int nt = kIbThreads;
  auto go = [&](auto tpm) {
    constexpr int T = decltype(tpm)::value;
    if (first) hipLaunchKernelGGL((k_ib_pipe<kWinLocal, T>), dim3(1), dim3(nt), lds, s, c, ib, src, step);
    else hipLaunchKernelGGL((k_ib_pipe<kWinHalo, T>), dim3(1), dim3(nt), lds, s, c, ib, src, step);
  };
  switch (ib_tpm(ib.n_markers, nt)) {
    case 4: go(std::integral_constant<int, 4>()); break;
    case 2: go(std::integral_constant<int, 2>()); break;
    default: go(std::integral_constant<int, 1>()); break;
  }
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

// defined in the translation unit that instantiates each mode
int fused_full(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s);
int fused_first(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s);
int fused_last(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s);

}  // namespace vsx
