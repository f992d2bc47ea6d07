// Batched 2D VIV sweep: every case of a batch advances in one launch per step.
//
// Launch layout (grid = n_cases + total tiles, 1024 threads per workgroup):
//   blocks [0, n_cases)           the IB workgroup of case b (window velocity,
//                                 multi-direct forcing, FP64 body force, Newmark)
//   blocks [n_cases, ...)         bulk tiles, through a block map: first the
//                                 tiles of every case outside the x-band of its
//                                 IB window, then the tiles inside the bands.
// A tile that intersects its case's force window waits (bounded) for the
// case's ready flag; the band tiles come last so that by the time they start
// the IB workgroups, which start first and need nothing from the tiles, are
// normally done: with many cases the IB work of all of them overlaps the
// streaming of the whole batch.
// Every kernel reads the batch clock (global step index of the run's first
// step) from device memory, so a captured HIP graph replays correctly: the
// ready flags, tile counters, body slots and history rows all follow the
// clock, which each run advances.
// Split mode (IB windows whose LDS would leave room for only one workgroup
// per CU, which halves the bulk tiles' occupancy): each step is two
// launches, the IB workgroups of all cases (k_viv_batch_ib), then the bulk
// tiles without LDS, which apply the force windows without waiting.
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "step_impl.hpp"
#include "viv2d.hpp"

namespace vsx {

struct BatchCase {
  StepCtx c;          // per-case step context; c.win = {ib.win_origin, ib.g_win, {S, S, 1}}
  VsxIB2D ib;
  int64_t f_off;      // population offset of the case in the batch buffers
  Tiles T;
  int32_t hist_cap, band_lo, band_hi, reserved;
};

constexpr int kCaseShift = 20;   // block map entry = case << 20 | tile index

// Populations of one tile cell of case `bc`.  PHASE: COLLIDE reads the
// caller's post-streaming state in place, FULL / STREAM pull and apply walls.
template <class SRC, class DST, int COLL, int PHASE, bool SPLIT>
__global__ __attribute__((amdgpu_flat_work_group_size(kTileCells, kTileCells), amdgpu_waves_per_eu(8, 8))) void
k_viv_batch(const BatchCase *__restrict__ tab, const uint32_t *__restrict__ bmap, int n_cases,
            const typename SRC::T *__restrict__ src, typename DST::T *__restrict__ dst, int step,
            const int32_t *__restrict__ clock) {
  extern __shared__ float lds[];
  constexpr bool IB = PHASE != VSX_PHASE_STREAM;
  const int b = SPLIT ? (int)blockIdx.x + n_cases : (int)blockIdx.x;   // split: tiles only
  const int gstep = IB ? *clock + step : 0;
  const int slot = gstep & 1;
  const BatchCase &bc = b < n_cases ? tab[b] : tab[bmap[b - n_cases] >> kCaseShift];
  const StepCtx &c = bc.c;
  int org[3] = {0, 0, 0};
  int wx0 = 0, wx1 = -1, wy0 = 0, wy1 = -1;
  if constexpr (IB) {
    __shared__ int s_org[2];
    if (threadIdx.x == 0) {
      int o[4];
      window_origin(c, bc.ib, bc.ib.body + 6 * slot, o);
      s_org[0] = o[2];
      s_org[1] = o[3];
    }
    __syncthreads();
    org[0] = s_org[0];
    org[1] = s_org[1];
    window_tiles(bc.T, org[0], org[1], bc.ib.win, wx0, wx1, wy0, wy1);
  }
  if (!SPLIT && b < n_cases) {
    if constexpr (IB && !SPLIT) {
      ib_atomic<SRC, PHASE == VSX_PHASE_FULL ? kWinPull : kWinLocal, 2, 2, false>(
          c, bc.ib, src + bc.f_off, nullptr, gstep < bc.hist_cap ? gstep : -1, gstep, nullptr, bc.ib.body + 6 * slot,
          bc.ib.body + 6 * (slot ^ 1), lds);
      publish_ready(bc.ib, gstep + 1);
    }
    return;
  }
  const int tile = (int)(bmap[b - n_cases] & ((1u << kCaseShift) - 1));
  const int TX = bc.T.TX, TY = bc.T.TY;
  const int tx = tile / bc.T.nty, ty = tile % bc.T.nty;
  const bool win = IB && tx >= wx0 && tx <= wx1 && ty >= wy0 && ty <= wy1;
  const int co[3] = {tx * TX + (int)threadIdx.x / TY, ty * TY + (int)threadIdx.x % TY, 0};
  const bool valid = co[0] < c.dims[0] && co[1] < c.dims[1];
  const uint32_t cell = valid ? cell_of<D2Q9>(c, co[0], co[1], 0) : 0;
  const typename SRC::T *__restrict__ s_case = src + bc.f_off;
  typename DST::T *__restrict__ d_case = dst + bc.f_off;
  float f[9];
  int cls = -1;
  if (valid) {
    if constexpr (PHASE == VSX_PHASE_COLLIDE) {
      local<D2Q9, SRC>(c, s_case, cell, f);
    } else {
      cls = c.p.n_walls > 0 ? wall_class<D2Q9>(c, co) : -1;
      if (cls >= 0) single_wall_state<D2Q9, SRC>(c, s_case, co, cls, f);
      else pull<D2Q9, SRC>(c, s_case, co[0], co[1], 0, f);
    }
  }
  if constexpr (IB && !SPLIT) {
    if (win) wait_ready(bc.ib, gstep + 1);
  }
  if (!valid) return;
  if constexpr (PHASE != VSX_PHASE_STREAM) finish<D2Q9, COLL>(c, cell, co, org, win, f);
  store_cell<D2Q9, DST>(c, d_case, cell, f);
}

// Split mode: the IB workgroup of every case (one workgroup per case), the
// single-case IB workgroup (marked-cell window pull, fixed-point LDS spreads,
// several threads per marker); dynamic LDS >= 16 S^2 bytes for every case.
template <class SRC, bool PULL>
__global__ __launch_bounds__(kIbThreads) void k_viv_batch_ib(const BatchCase *__restrict__ tab,
                                                             const typename SRC::T *__restrict__ src, int step,
                                                             const int32_t *__restrict__ clock) {
  extern __shared__ float lds[];
  const BatchCase &bc = tab[blockIdx.x];
  const int gstep = *clock + step, slot = gstep & 1;
  ib_atomic<SRC, PULL ? kWinPull : kWinLocal, 2, 3, false>(bc.c, bc.ib, src + bc.f_off, nullptr,
                                                          gstep < bc.hist_cap ? gstep : -1, gstep, nullptr,
                                                          bc.ib.body + 6 * slot, bc.ib.body + 6 * (slot ^ 1), lds);
}

// Dynamic LDS above which the IB workgroups run in their own launch (two
// 1024-thread bulk workgroups per CU need each at most half of the 160 KB).
constexpr int kBatchSplitLds = 76 * 1024;

__global__ void k_clock_advance(int32_t *clock, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *clock += n;
}

// ------------------------------------------------------------------ host side

int64_t tiles_of(const VsxVivCase &k, Tiles &T) {
  T.TY = k.tile_y;
  T.TX = kTileCells / k.tile_y;
  T.ntx = (k.lbm.nx + T.TX - 1) / T.TX;
  T.nty = (k.lbm.ny + T.TY - 1) / T.TY;
  return (int64_t)T.ntx * T.nty;
}

int64_t table_bytes(const VsxVivCase *cases, int n) {
  int64_t tiles = 0;
  for (int k = 0; k < n; ++k) {
    Tiles T;
    if (cases[k].tile_y < 64 || cases[k].tile_y > kTileCells) return -1;
    tiles += tiles_of(cases[k], T);
  }
  return (int64_t)sizeof(BatchCase) * n + 4 * tiles;
}

template <class SRC, class DST, int COLL, int PHASE>
void launch_batch_c(const VsxVivBatch &b, const void *src, void *dst, int step, hipStream_t s) {
  const auto *tab = static_cast<const BatchCase *>(b.table);
  const auto *bmap = reinterpret_cast<const uint32_t *>(static_cast<const char *>(b.table) +
                                                        sizeof(BatchCase) * b.n_cases);
  auto S = static_cast<const typename SRC::T *>(src);
  auto D = static_cast<typename DST::T *>(dst);
  const int32_t *clk = b.clock;
  if (b.lds_bytes > kBatchSplitLds) {
    if constexpr (PHASE != VSX_PHASE_STREAM)
      hipLaunchKernelGGL((k_viv_batch_ib<SRC, PHASE == VSX_PHASE_FULL>), dim3((unsigned)b.n_cases), dim3(kIbThreads),
                         (size_t)b.lds_bytes, s, tab, S, step, clk);
    hipLaunchKernelGGL((k_viv_batch<SRC, DST, COLL, PHASE, true>), dim3((unsigned)(b.n_blocks - b.n_cases)),
                       dim3(kTileCells), 0, s, tab, bmap, b.n_cases, S, D, step, clk);
  } else {
    hipLaunchKernelGGL((k_viv_batch<SRC, DST, COLL, PHASE, false>), dim3((unsigned)b.n_blocks), dim3(kTileCells),
                       (size_t)b.lds_bytes, s, tab, bmap, b.n_cases, S, D, step, clk);
  }
}

template <class SRC, class DST, int PHASE>
void launch_batch_coll(const VsxVivBatch &b, const void *src, void *dst, int step, hipStream_t s) {
  switch (b.collision) {
    case VSX_COLL_BGK: return launch_batch_c<SRC, DST, VSX_COLL_BGK, PHASE>(b, src, dst, step, s);
    case VSX_COLL_MRT: return launch_batch_c<SRC, DST, VSX_COLL_MRT, PHASE>(b, src, dst, step, s);
    case VSX_COLL_KBC: return launch_batch_c<SRC, DST, VSX_COLL_KBC, PHASE>(b, src, dst, step, s);
    default: return launch_batch_c<SRC, DST, VSX_COLL_REG, PHASE>(b, src, dst, step, s);
  }
}

template <class ST, int PHASE>
void launch_batch_p(const VsxVivBatch &b, const void *src, void *dst, int step, hipStream_t s) {
  using F32 = Store<VSX_STORE_F32>;
  using SRC = std::conditional_t<PHASE == VSX_PHASE_COLLIDE, F32, ST>;
  using DST = std::conditional_t<PHASE == VSX_PHASE_STREAM, F32, ST>;
  if constexpr (PHASE == VSX_PHASE_STREAM) {
    launch_batch_c<SRC, DST, VSX_COLL_BGK, PHASE>(b, src, dst, step, s);
  } else {
    // full steps whose work buffer exceeds the last-level cache: non-temporal field stores
    if constexpr (PHASE == VSX_PHASE_FULL) {
      if ((size_t)b.total_pops * sizeof(typename ST::T) > kLastLevelCacheBytes)
        return launch_batch_coll<SRC, Streamed<DST>, PHASE>(b, src, dst, step, s);
    }
    launch_batch_coll<SRC, DST, PHASE>(b, src, dst, step, s);
  }
}

template <int PHASE>
void launch_batch(const VsxVivBatch &b, const void *src, void *dst, int step, hipStream_t s) {
  switch (b.storage) {
    case VSX_STORE_F32: return launch_batch_p<Store<VSX_STORE_F32>, PHASE>(b, src, dst, step, s);
    case VSX_STORE_F16S: return launch_batch_p<Store<VSX_STORE_F16S>, PHASE>(b, src, dst, step, s);
    case VSX_STORE_BF16S: return launch_batch_p<Store<VSX_STORE_BF16S>, PHASE>(b, src, dst, step, s);
    default: return launch_batch_p<Store<VSX_STORE_E3M12S>, PHASE>(b, src, dst, step, s);
  }
}

int batch_prepare(const VsxVivCase *cases, int n, void *table, int32_t *clock, VsxVivBatch *out) {
  VSX_REQUIRE(cases && table && clock && out, "null argument");
  VSX_REQUIRE(n > 0 && n < (1 << 11), "case count %d out of range (1..2047)", n);
  const VsxLbm &p0 = cases[0].lbm;
  int max_lds = 0;
  if (int rc_ = max_lds_per_block(&max_lds)) return rc_;
  std::vector<BatchCase> tab(n);
  std::vector<uint32_t> outer, band;
  int64_t pops = 0;
  int lds = 0;
  for (int k = 0; k < n; ++k) {
    const VsxVivCase &vc = cases[k];
    int rc = validate_lbm<D2Q9>(&vc.lbm);
    if (rc) return rc;
    rc = validate_ib2d(&vc.lbm, &vc.ib);
    if (rc) return rc;
    VSX_REQUIRE(vc.lbm.collision == p0.collision && vc.lbm.forcing == p0.forcing && vc.lbm.storage == p0.storage,
                "case %d: collision, forcing and storage must match the first case", k);
    VSX_REQUIRE(vc.ib.n_markers <= kIbThreads, "case %d: at most %d markers per case", k, kIbThreads);
    VSX_REQUIRE(vc.tile_y == 64 || vc.tile_y == 128 || vc.tile_y == 256 || vc.tile_y == 512 || vc.tile_y == 1024,
                "case %d: tile_y %d not in {64, 128, 256, 512, 1024}", k, vc.tile_y);
    VSX_REQUIRE(!(vc.lbm.n_walls > 0 && walls_need_edge<D2Q9>(vc.lbm)),
                "case %d: batched cases take walls without shared corners (no two walls on different axes)", k);
    VSX_REQUIRE(vc.f_offset >= 0 && vc.f_offset % 64 == 0, "case %d: f_offset must be a non-negative multiple of 64", k);
    VSX_REQUIRE(vc.hist_cap >= 0, "case %d: negative history capacity", k);
    BatchCase &bc = tab[k];
    bc.ib = vc.ib;
    bc.c = make_ctx<D2Q9>(vc.lbm, Window{vc.ib.win_origin, vc.ib.g_win, {vc.ib.win, vc.ib.win, 1}});
    bc.f_off = vc.f_offset;
    const int64_t ntiles = tiles_of(vc, bc.T);
    VSX_REQUIRE(ntiles < (1 << kCaseShift), "case %d: too many tiles", k);
    bc.hist_cap = vc.hist_cap;
    // x-band of tile rows that hold the window with the body at rest
    const int x0 = std::max(0, std::min((int)vc.ib.win_x0, vc.lbm.nx - vc.ib.win));
    bc.band_lo = x0 / bc.T.TX;
    bc.band_hi = (x0 + vc.ib.win - 1) / bc.T.TX;
    bc.reserved = 0;
    for (int tx = 0; tx < bc.T.ntx; ++tx)
      for (int ty = 0; ty < bc.T.nty; ++ty) {
        const uint32_t entry = ((uint32_t)k << kCaseShift) | (uint32_t)(tx * bc.T.nty + ty);
        (tx >= bc.band_lo && tx <= bc.band_hi ? band : outer).push_back(entry);
      }
    pops = std::max(pops, vc.f_offset + (int64_t)9 * vc.lbm.nx * vc.lbm.ny);
    lds = std::max(lds, 4 * ib_atomic_lds_words(vc.ib.win));   // the IB workgroup's (ib_atomic)
  }
  VSX_REQUIRE(lds + 1024 <= max_lds, "IB windows need %d B of LDS, device allows %d", lds, max_lds);
  std::vector<char> host(sizeof(BatchCase) * n + 4 * (outer.size() + band.size()));
  std::memcpy(host.data(), tab.data(), sizeof(BatchCase) * n);
  char *bm = host.data() + sizeof(BatchCase) * n;
  std::memcpy(bm, outer.data(), 4 * outer.size());
  std::memcpy(bm + 4 * outer.size(), band.data(), 4 * band.size());
  VSX_HIP(hipMemcpy(table, host.data(), host.size(), hipMemcpyHostToDevice));
  out->n_cases = n;
  out->collision = p0.collision;
  out->storage = p0.storage;
  out->lds_bytes = lds;
  out->n_blocks = n + (int64_t)(outer.size() + band.size());
  out->total_pops = pops;
  out->table = table;
  out->clock = clock;
  return VSX_OK;
}

int batch_run(const VsxVivBatch *b, const float *f_in, float *f_out, void *wa, void *wb, int n, hipStream_t s) {
  VSX_REQUIRE(b && b->table && b->clock && b->n_cases > 0, "batch not prepared");
  VSX_REQUIRE(f_in && f_out && wa, "null buffer");
  VSX_REQUIRE(n >= 0, "negative step count");
  VSX_REQUIRE(n <= 1 || wb, "need two work buffers for more than one step");
  if (n == 0) {
    if (f_out != f_in) VSX_HIP(hipMemcpyAsync(f_out, f_in, sizeof(float) * b->total_pops, hipMemcpyDeviceToDevice, s));
    return VSX_OK;
  }
  const void *cur = f_in;
  void *bufs[2] = {wa, wb};
  for (int t = 0; t < n; ++t) {
    void *dst = bufs[t & 1];
    if (t == 0) {
      launch_batch<VSX_PHASE_COLLIDE>(*b, cur, dst, t, s);
    } else {
      timer_begin(s);
      launch_batch<VSX_PHASE_FULL>(*b, cur, dst, t, s);
      timer_end(s);
    }
    VSX_CHECK_LAUNCH();
    cur = dst;
  }
  launch_batch<VSX_PHASE_STREAM>(*b, cur, f_out, 0, s);
  hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, s, b->clock, n);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

}  // namespace vsx

using namespace vsx;

extern "C" int64_t vsx_viv2d_batch_table_bytes(const VsxVivCase *cases, int32_t n_cases) {
  if (!cases || n_cases <= 0) return -1;
  return table_bytes(cases, n_cases);
}

extern "C" int vsx_viv2d_batch_prepare(const VsxVivCase *cases, int32_t n_cases, void *table, int32_t *clock,
                                       VsxVivBatch *out) {
  return batch_prepare(cases, n_cases, table, clock, out);
}

extern "C" int vsx_viv2d_batch_run(const VsxVivBatch *b, const float *f_in, float *f_out, void *work_a, void *work_b,
                                   int32_t n_steps, void *stream) {
  return batch_run(b, f_in, f_out, work_a, work_b, n_steps, as_stream(stream));
}

extern "C" int vsx_viv2d_batch_graph(const VsxVivBatch *b, const float *f_in, float *f_out, void *work_a,
                                     void *work_b, int32_t n_steps, void **graph_exec) {
  VSX_REQUIRE(graph_exec, "null graph output");
  hipStream_t s;
  VSX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  int rc = VSX_OK;
  if (e == hipSuccess) {
    rc = batch_run(b, f_in, f_out, work_a, work_b, n_steps, s);
    e = hipStreamEndCapture(s, &g);
  }
  hipGraphExec_t ex = nullptr;
  if (e == hipSuccess && rc == VSX_OK) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(s);
  if (rc != VSX_OK) return rc;
  if (e != hipSuccess) return fail(VSX_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  *graph_exec = ex;
  return VSX_OK;
}

extern "C" int vsx_graph_launch(void *graph_exec, void *stream) {
  VSX_REQUIRE(graph_exec, "null graph");
  VSX_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), as_stream(stream)));
  return VSX_OK;
}

extern "C" int vsx_graph_destroy(void *graph_exec) {
  if (graph_exec) VSX_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
  return VSX_OK;
}
