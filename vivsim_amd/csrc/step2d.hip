// D2Q9 fused steps and the 2D IB-LBM vortex-induced-vibration driver.
#include "step_impl.hpp"
#include "viv2d.hpp"

namespace vsx {

// Stand-alone IB kernel (the first step of a run, and the non-fused path).
template <class SRC, bool PULL, int MPT>
__global__ __launch_bounds__(kIbThreads) void k_viv_ib(StepCtx c_arg, VsxIB2D ib,
                                                        const typename SRC::T *__restrict__ src, int step) {
  const StepCtx &c = kernel_ctx();
  extern __shared__ float lds[];
  const int gstep = ib.step0 + step, slot = gstep & 1;
  ib_block<SRC, PULL, MPT, 4, 0, false, 4>(c, ib, src, step, gstep, 0, ib.body + 6 * slot, ib.body + 6 * (slot ^ 1),
                                        lds);
  publish_ready(ib, gstep + 1);
}

// Window velocity of the fused step: the window tiles publish the
// post-streaming velocity of their cells (store_window_u, count_window_tile)
// and the IB workgroup waits for all of them.

// Two bulk cells per thread (16-bit storages): both cells' populations are
// loaded before either is computed, so a wave keeps as many bytes in flight
// as with one FP32 cell per thread.  a and b are interior cells (no walls)
// outside the IB window, in a step without a g field or uniform g.
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
template <class SRC, class DST, int COLL>
__device__ __forceinline__ void bulk_cell(const StepCtx &c, const typename SRC::T *__restrict__ src,
                                          typename DST::T *__restrict__ dst, const int (&co)[3], const int (&org)[3]) {
  float f[9];
  const uint32_t cell = cell_of<D2Q9>(c, co[0], co[1], 0);
  const int cls = c.p.n_walls > 0 ? wall_class<D2Q9>(c, co) : -1;
  if (cls >= 0) {
    single_wall_state<D2Q9, SRC>(c, src, co, cls, f);
    finish<D2Q9, COLL>(c, cell, co, org, false, f);
    store_cell<D2Q9, DST>(c, dst, cell, f);
    return;
  }
  pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f);
  finish<D2Q9, COLL>(c, cell, co, org, false, f);
  store_cell<D2Q9, DST>(c, dst, cell, f);
}

// Blocks of the fused step with PAIR: block 0 the IB workgroup, blocks
// 1..n_win the window tiles, then one block per two bulk tiles; the host
// launches 1 + ceil((tiles + max window tiles) / 2) blocks and the surplus exits.
template <class SRC, class DST, int COLL, bool PAIR>
__global__ __attribute__((amdgpu_flat_work_group_size(kTileCells, kTileCells), amdgpu_waves_per_eu(8, 8))) void
k_viv_fused(StepCtx c_arg, VsxIB2D ib, const typename SRC::T *__restrict__ src, typename DST::T *__restrict__ dst,
            int step, Tiles T) {
  const StepCtx &c = kernel_ctx();
  extern __shared__ float lds[];
  const int gstep = ib.step0 + step, slot = gstep & 1;
  const double *body_in = ib.body + 6 * slot;
  __shared__ int s_org[2];
  if (threadIdx.x == 0) {
    int o[4];
    window_origin(c, ib, body_in, o);
    s_org[0] = o[2];
    s_org[1] = o[3];
  }
  __syncthreads();
  const int org[3] = {s_org[0], s_org[1], 0};
  int wx0, wx1, wy0, wy1;
  window_tiles(T, org[0], org[1], ib.win, wx0, wx1, wy0, wy1);
  if (blockIdx.x == 0) {
    ib_block<SRC, true, 1, 1, 1, true, 1>(c, ib, src, step, gstep, (wx1 - wx0 + 1) * (wy1 - wy0 + 1), body_in,
                                          ib.body + 6 * (slot ^ 1), lds);
    publish_ready(ib, gstep + 1);
    return;
  }
  const int n_win = (wx1 - wx0 + 1) * (wy1 - wy0 + 1);
  if constexpr (PAIR) {
    if ((int)blockIdx.x - 1 >= n_win) {
    const int q = n_win + 2 * ((int)blockIdx.x - 1 - n_win), n_tiles = T.ntx * T.nty;
    if (q >= n_tiles) return;
    int ta[2], tb[2];
    bool w_;
    tile_of_block(q, T, wx0, wx1, wy0, wy1, ta[0], ta[1], w_);
    const bool has_b = q + 1 < n_tiles;
    if (has_b) tile_of_block(q + 1, T, wx0, wx1, wy0, wy1, tb[0], tb[1], w_);
    const int a[3] = {ta[0] * T.TX + (int)threadIdx.x / T.TY, ta[1] * T.TY + (int)threadIdx.x % T.TY, 0};
    const int b[3] = {has_b ? tb[0] * T.TX + (int)threadIdx.x / T.TY : c.dims[0],
                      has_b ? tb[1] * T.TY + (int)threadIdx.x % T.TY : c.dims[1], 0};
    if (threadIdx.x == 0) spin_until(ib, &ib.win_origin[6], gstep + 1);
    __syncthreads();
    const bool va = a[0] < c.dims[0] && a[1] < c.dims[1], vb = b[0] < c.dims[0] && b[1] < c.dims[1];
    const bool wa = va && c.p.n_walls > 0 && wall_class<D2Q9>(c, a) >= 0;
    const bool wb = vb && c.p.n_walls > 0 && wall_class<D2Q9>(c, b) >= 0;
    const bool unforced = !c.p.g && c.p.g_uniform[0] == 0.f && c.p.g_uniform[1] == 0.f;
    if (unforced && va && vb && !wa && !wb) {
      bulk_pair<SRC, DST, COLL>(c, src, dst, a, b);
    } else {
      if (va) bulk_cell<SRC, DST, COLL>(c, src, dst, a, org);
      if (vb) bulk_cell<SRC, DST, COLL>(c, src, dst, b, org);
    }
    return;
    }
  }
  int tx, ty;
  bool in_win;
  tile_of_block(blockIdx.x - 1, T, wx0, wx1, wy0, wy1, tx, ty, in_win);
  const int co[3] = {tx * T.TX + (int)threadIdx.x / T.TY, ty * T.TY + (int)threadIdx.x % T.TY, 0};
  const bool valid = co[0] < c.dims[0] && co[1] < c.dims[1];
  const uint32_t cell = valid ? cell_of<D2Q9>(c, co[0], co[1], 0) : 0;
  float f[9];
  if (in_win) {
    // window tile: post-streaming state -> window velocity for the IB
    // workgroup -> wait for its force window -> collide
    if (valid) {
      const int cls = c.p.n_walls > 0 ? wall_class<D2Q9>(c, co) : -1;
      if (cls >= 0) single_wall_state<D2Q9, SRC>(c, src, co, cls, f);
      else pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f);
    }
    store_window_u(ib, valid, co, org, f);
    count_window_tile(ib, gstep, (wx1 - wx0 + 1) * (wy1 - wy0 + 1));
    wait_ready(ib, gstep + 1);
    if (!valid) return;
    finish<D2Q9, COLL>(c, cell, co, org, true, f);
    store_cell<D2Q9, DST>(c, dst, cell, f);
    return;
  }
  // bulk tile: wait (bounded) until the window tiles have published the window
  // velocity, so that their pulls are not queued behind the bulk traffic
  if (threadIdx.x == 0) spin_until(ib, &ib.win_origin[6], gstep + 1);
  __syncthreads();
  if (!valid) return;
  const int cls = c.p.n_walls > 0 ? wall_class<D2Q9>(c, co) : -1;
  if (cls >= 0) {
    single_wall_state<D2Q9, SRC>(c, src, co, cls, f);
    finish<D2Q9, COLL>(c, cell, co, org, false, f);
    store_cell<D2Q9, DST>(c, dst, cell, f);
    return;
  }
  pull<D2Q9, SRC>(c, src, co[0], co[1], 0, f);
  finish<D2Q9, COLL>(c, cell, co, org, false, f);
  store_cell<D2Q9, DST>(c, dst, cell, f);
}

// ------------------------------------------------------------------ host side

template <class SRC, bool PULL>
int launch_viv_ib(const StepCtx &c, const VsxIB2D &ib, const void *src, int step, hipStream_t s) {
  const size_t lds = 4 * (size_t)ib_lds_words(ib.win, ib.n_markers, ib.n_markers <= kIbThreads);
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

// Bulk blocks take two tiles (PAIR) for the 16-bit storages on grids of at
// least kPairMinTiles tiles: there the streaming bounds the step and the pairs
// double the bytes in flight (+12 % at 4096^2).  On smaller grids the IB
// workgroup bounds the step and single tiles keep it shorter; FP32 tiles stay
// single (no gain at 4096^2, a loss at 1024^2).
constexpr int kPairMinTiles = 4096;

template <class ST, int COLL, bool PAIR>
int launch_fused_p(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s,
                   const Tiles &T) {
  const size_t lds = 4 * (size_t)ib_lds_words(ib.win, ib.n_markers, true);
  const int n_tiles = T.ntx * T.nty;
  int grid = n_tiles;
  if (PAIR) {
    const int nw_max = ((ib.win + T.TX - 1) / T.TX + 1) * ((ib.win + T.TY - 1) / T.TY + 1);
    grid = (n_tiles + nw_max + 1) / 2;
    if (grid > n_tiles) grid = n_tiles;
  }
  timer_begin(s);
  hipLaunchKernelGGL((k_viv_fused<ST, ST, COLL, PAIR>), dim3(1 + grid), dim3(kTileCells), lds, s, c, ib,
                     static_cast<const typename ST::T *>(src), static_cast<typename ST::T *>(dst), step, T);
  timer_end(s);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

template <class ST, int COLL>
int launch_fused_c(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  const Tiles T = make_tiles(c);
  if constexpr (sizeof(typename ST::T) == 2) {
    if (T.ntx * T.nty >= kPairMinTiles) return launch_fused_p<ST, COLL, true>(c, ib, src, dst, step, s, T);
  }
  return launch_fused_p<ST, COLL, false>(c, ib, src, dst, step, s, T);
}

template <class ST>
int launch_fused(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  switch (c.p.collision) {
    case VSX_COLL_BGK: return launch_fused_c<ST, VSX_COLL_BGK>(c, ib, src, dst, step, s);
    case VSX_COLL_MRT: return launch_fused_c<ST, VSX_COLL_MRT>(c, ib, src, dst, step, s);
    case VSX_COLL_KBC: return launch_fused_c<ST, VSX_COLL_KBC>(c, ib, src, dst, step, s);
    default: return launch_fused_c<ST, VSX_COLL_REG>(c, ib, src, dst, step, s);
  }
}

int fused_any(const StepCtx &c, const VsxIB2D &ib, const void *src, void *dst, int step, hipStream_t s) {
  switch (c.p.storage) {
    case VSX_STORE_F32: return launch_fused<Store<VSX_STORE_F32>>(c, ib, src, dst, step, s);
    case VSX_STORE_F16S: return launch_fused<Store<VSX_STORE_F16S>>(c, ib, src, dst, step, s);
    case VSX_STORE_BF16S: return launch_fused<Store<VSX_STORE_BF16S>>(c, ib, src, dst, step, s);
    default: return launch_fused<Store<VSX_STORE_E3M12S>>(c, ib, src, dst, step, s);
  }
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
  // fused launches need one marker per thread of the IB workgroup, and walls
  // without corner cells (every window cell is then finished by its tile)
  const bool fused = ib->fused && ib->n_markers <= kIbThreads && !(p->n_walls > 0 && walls_need_edge<D2Q9>(*p));
  const void *cur = f_in;
  void *bufs[2] = {wa, wb};
  for (int t = 0; t < n; ++t) {
    void *dst = bufs[t & 1];
    if (t == 0 || !fused) {
      rc = viv_ib_any(c, *ib, cur, t == 0, t, s);
      if (rc) return rc;
      rc = step_any<D2Q9>(c, cur, t == 0, dst, false, t == 0 ? VSX_PHASE_COLLIDE : VSX_PHASE_FULL, s);
    } else {
      rc = fused_any(c, *ib, cur, dst, t, s);
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
