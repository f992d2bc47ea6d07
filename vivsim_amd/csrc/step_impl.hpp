// Host-side launch logic of the fused steps (validation, template dispatch,
// the N-step driver).  Included by step2d.hip and step3d.hip.
#pragma once
#include "step.hpp"

namespace vsx {

template <class L>
int validate_lbm(const VsxLbm *p) {
  if (!p) return fail(VSX_ERR_VALUE, "null parameters");
  if (p->nx < 3 || p->ny < 3 || (L::D == 3 ? p->nz < 3 : p->nz != 1))
    return fail(VSX_ERR_VALUE, "bad grid %d x %d x %d", p->nx, p->ny, p->nz);
  if ((int64_t)p->nx * p->ny * p->nz >= ((int64_t)1 << 31)) return fail(VSX_ERR_VALUE, "grid too large (>= 2^31 cells)");
  if (p->collision < VSX_COLL_BGK || p->collision > VSX_COLL_REG) return fail(VSX_ERR_VALUE, "unknown collision %d", p->collision);
  if (p->forcing < VSX_FORCE_NONE || p->forcing > VSX_FORCE_EDM) return fail(VSX_ERR_VALUE, "unknown forcing %d", p->forcing);
  if (p->storage < VSX_STORE_F32 || p->storage > VSX_STORE_E3M12S) return fail(VSX_ERR_VALUE, "unknown storage %d", p->storage);
  if (p->n_walls < 0 || p->n_walls > L::n_walls) return fail(VSX_ERR_VALUE, "bad wall count %d", p->n_walls);
  for (int k = 0; k < p->n_walls; ++k) {
    const VsxWall &w = p->walls[k];
    if (w.loc < 0 || w.loc >= L::n_walls) return fail(VSX_ERR_KEY, "unknown wall location %d", w.loc);
    if (w.kind < VSX_WALL_NEE || w.kind > VSX_WALL_SPECULAR) return fail(VSX_ERR_VALUE, "unknown wall kind %d", w.kind);
    if (w.mode < VSX_MODE_PLAIN || w.mode > VSX_MODE_FORCE_CORRECTED) return fail(VSX_ERR_VALUE, "unknown wall mode %d", w.mode);
  }
  return VSX_OK;
}

template <class L>
StepCtx make_ctx(const VsxLbm &p, Window win) {
  StepCtx c;
  c.p = p;
  c.win = win;
  c.dims[0] = p.nx;
  c.dims[1] = p.ny;
  c.dims[2] = L::D == 3 ? p.nz : 1;
  c.n = (int64_t)c.dims[0] * c.dims[1] * c.dims[2];
  for (int l = 0; l < 6; ++l) {
    c.order[l] = -1;
    c.at[l] = VsxWall{};
  }
  for (int k = 0; k < p.n_walls; ++k) {
    const int l = p.walls[k].loc;
    c.order[l] = c.order[l] == -1 ? k : -2;
    c.at[l] = p.walls[k];
  }
  // shift of the 16-bit storages: feq_i(1, store_u) (lbm/basic.py:114-135 at rho = 1)
  const float u0 = p.store_u[0], u1 = p.store_u[1], u2 = L::D == 3 ? p.store_u[2] : 0.f;
  const float usq = u0 * u0 + u1 * u1 + u2 * u2;
  for (int i = 0; i < 19; ++i) c.shift[i] = 0.f;
  for (int i = 0; i < L::Q; ++i) {
    const float cu = L::cx[i] * u0 + L::cy[i] * u1 + L::cz[i] * u2;
    c.shift[i] = L::w[i] * (1.f + 3.f * cu + 4.5f * cu * cu - 1.5f * usq);
  }
  return c;
}

template <class L>
void bulk_grid(const StepCtx &c, dim3 &grid, dim3 &block) {
  const int inner = L::D == 2 ? c.dims[1] : c.dims[2];
  int b = ((inner + 63) / 64) * 64;
  if (b > 256) b = 256;
  block = dim3(b);
  if (L::D == 2) grid = dim3((inner + b - 1) / b, c.dims[0]);
  else grid = dim3((inner + b - 1) / b, c.dims[1], c.dims[0]);
}

template <class L>
int64_t edge_per_wall(const StepCtx &c) {
  int64_t m = 0;
  for (int k = 0; k < c.p.n_walls; ++k) {
    const int a = L::w_axis[c.p.walls[k].loc];
    int64_t nodes = 1;
    for (int b = 0; b < L::D; ++b)
      if (b != a) nodes *= c.dims[b];
    if (nodes > m) m = nodes;
  }
  return m;
}

template <class L, int PHASE, class SRC, class DST>
int launch_phase(const StepCtx &c, const void *src, void *dst, hipStream_t s) {
  dim3 grid, block;
  bulk_grid<L>(c, grid, block);
  auto S = static_cast<const typename SRC::T *>(src);
  auto D = static_cast<typename DST::T *>(dst);
  if constexpr (PHASE == VSX_PHASE_FULL) timer_begin(s);
  if constexpr (PHASE == VSX_PHASE_STREAM) {
    hipLaunchKernelGGL((k_bulk<L, VSX_COLL_BGK, PHASE, SRC, DST>), grid, block, 0, s, c, S, D);
  } else {
    switch (c.p.collision) {
      case VSX_COLL_BGK: hipLaunchKernelGGL((k_bulk<L, VSX_COLL_BGK, PHASE, SRC, DST>), grid, block, 0, s, c, S, D); break;
      case VSX_COLL_MRT: hipLaunchKernelGGL((k_bulk<L, VSX_COLL_MRT, PHASE, SRC, DST>), grid, block, 0, s, c, S, D); break;
      case VSX_COLL_KBC: hipLaunchKernelGGL((k_bulk<L, VSX_COLL_KBC, PHASE, SRC, DST>), grid, block, 0, s, c, S, D); break;
      default: hipLaunchKernelGGL((k_bulk<L, VSX_COLL_REG, PHASE, SRC, DST>), grid, block, 0, s, c, S, D); break;
    }
  }
  if constexpr (PHASE == VSX_PHASE_FULL) timer_end(s);
  if constexpr (PHASE != VSX_PHASE_COLLIDE) {
    if (c.p.n_walls > 0 && walls_need_edge<L>(c.p)) {
      const int64_t pw = edge_per_wall<L>(c);
      const int64_t tot = pw * c.p.n_walls;
      hipLaunchKernelGGL((k_edge<L, PHASE, SRC, DST>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, c, S, D, pw);
    }
  }
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

// runtime (phase, src storage, dst storage) -> template instantiation
template <class L, class ST>
int phase_dispatch(const StepCtx &c, const void *src, bool src_f32, void *dst, bool dst_f32, int phase, hipStream_t s) {
  using F32 = Store<VSX_STORE_F32>;
  switch (phase) {
    case VSX_PHASE_COLLIDE:
      return src_f32 ? launch_phase<L, VSX_PHASE_COLLIDE, F32, ST>(c, src, dst, s)
                     : launch_phase<L, VSX_PHASE_COLLIDE, ST, ST>(c, src, dst, s);
    case VSX_PHASE_FULL:
      // an FP32 work buffer larger than the last-level cache: non-temporal
      // field stores (FP32 only, which keeps the kernel count of the generic
      // step down; a 16-bit field needs twice the cells to reach that size)
      if constexpr (std::is_same_v<ST, Store<VSX_STORE_F32>>) {
        if ((size_t)c.n * L::Q * sizeof(float) > kLastLevelCacheBytes)
          return launch_phase<L, VSX_PHASE_FULL, ST, Streamed<ST>>(c, src, dst, s);
      }
      return launch_phase<L, VSX_PHASE_FULL, ST, ST>(c, src, dst, s);
    case VSX_PHASE_STREAM:
      return dst_f32 ? launch_phase<L, VSX_PHASE_STREAM, ST, F32>(c, src, dst, s)
                     : launch_phase<L, VSX_PHASE_STREAM, ST, ST>(c, src, dst, s);
    default:
      return fail(VSX_ERR_VALUE, "unknown phase %d", phase);
  }
}

template <class L>
int step_any(const StepCtx &c, const void *src, bool src_f32, void *dst, bool dst_f32, int phase, hipStream_t s) {
  switch (c.p.storage) {
    case VSX_STORE_F32: return phase_dispatch<L, Store<VSX_STORE_F32>>(c, src, true, dst, true, phase, s);
    case VSX_STORE_F16S: return phase_dispatch<L, Store<VSX_STORE_F16S>>(c, src, src_f32, dst, dst_f32, phase, s);
    case VSX_STORE_BF16S: return phase_dispatch<L, Store<VSX_STORE_BF16S>>(c, src, src_f32, dst, dst_f32, phase, s);
    default: return phase_dispatch<L, Store<VSX_STORE_E3M12S>>(c, src, src_f32, dst, dst_f32, phase, s);
  }
}

template <class L>
int step_entry(const VsxLbm *p, const void *src, int src_f32, void *dst, int dst_f32, int phase, void *stream) {
  int rc = validate_lbm<L>(p);
  if (rc) return rc;
  VSX_REQUIRE(src && dst && src != dst, "step needs distinct source and destination buffers");
  const StepCtx c = make_ctx<L>(*p, Window{nullptr, nullptr, {0, 0, 0}});
  return step_any<L>(c, src, src_f32 != 0, dst, dst_f32 != 0, phase, as_stream(stream));
}

// n reference steps: COLLIDE(f_in -> a), FULL x (n-1) (a <-> b), STREAM(-> f_out)
template <class L>
int run_entry(const VsxLbm *p, const float *f_in, float *f_out, void *wa, void *wb, int n, void *stream) {
  int rc = validate_lbm<L>(p);
  if (rc) return rc;
  VSX_REQUIRE(f_in && f_out && wa, "null buffer");
  VSX_REQUIRE(n >= 0, "negative step count");
  VSX_REQUIRE(n <= 1 || wb, "need two work buffers for more than one step");
  hipStream_t s = as_stream(stream);
  const StepCtx c = make_ctx<L>(*p, Window{nullptr, nullptr, {0, 0, 0}});
  if (n == 0) {
    if (f_out != f_in) VSX_HIP(hipMemcpyAsync(f_out, f_in, sizeof(float) * L::Q * c.n, hipMemcpyDeviceToDevice, s));
    return VSX_OK;
  }
  rc = step_any<L>(c, f_in, true, wa, false, VSX_PHASE_COLLIDE, s);
  if (rc) return rc;
  void *cur = wa, *nxt = wb;
  for (int t = 1; t < n; ++t) {
    rc = step_any<L>(c, cur, false, nxt, false, VSX_PHASE_FULL, s);
    if (rc) return rc;
    void *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  return step_any<L>(c, cur, false, f_out, true, VSX_PHASE_STREAM, s);
}

}  // namespace vsx
