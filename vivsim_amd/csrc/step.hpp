// Fused pull-scheme LBM step, generic over the lattice.
//
// The reference composes a time step from pure functions
//   macroscopic -> equilibrium -> collision -> forcing -> streaming -> walls
// (README.md:105-121, examples/2d/lid_driven_cavity.py:55-64,
// examples/2d/poiseuille_channel.py:80-128).  Here the stored state is the
// POST-COLLISION population field; one pass per cell
//   1. pulls f_i(x) = f_post_i(x - c_i) (periodic),
//   2. applies the ordered wall list to wall cells,
//   3. computes rho, u, adds the body force, collides and forces in registers,
//   4. writes the new post-collision populations of the cell.
// So every population is read once and written once per step (Q loads + Q
// stores = 72 B/cell D2Q9, 152 B/cell D3Q19 in FP32).  A run of N reference
// steps is one COLLIDE pass (the input is a post-streaming state), N-1 FULL
// passes and one STREAM pass that recovers the reference's post-streaming state.
//
// Wall cells: the reference applies each wall to the whole streamed array, in
// order, so a later wall can read cells an earlier wall rewrote (NEE reads the
// neighbour line) and the last writer wins at corners.  A cell on exactly one
// wall whose update reads no cell rewritten by an earlier wall is finished in
// the bulk kernel (one extra pull for the neighbour line).  The remaining cells
// (corners and corner-adjacent cells of perpendicular wall pairs) go to a
// separate edge kernel that replays the wall list recursively (depth <= number
// of walls); it is launched only when such cells can exist.
#pragma once
#include <type_traits>
#include <utility>

#include "cell.hpp"
#include "common.hpp"
#include "lattice.hpp"
#include "wall.hpp"

namespace vsx {

// Body-force window written by the immersed-boundary kernels (device pointers).
struct Window {
  const int32_t *origin;   // [3] clamped window origin, or null (no window)
  const float *g;          // [D][S0][S1(][S2)]
  int32_t size[3];
};

struct StepCtx {
  VsxLbm p;
  Window win;
  int32_t dims[3];         // nx, ny, nz (nz = 1 in 2D)
  int64_t n;               // cells
  // Walls by location, so device code indexes walls only by compile-time
  // location: order[LOC] = position of the wall at LOC in p.walls (-1: none,
  // -2: several walls at LOC), at[LOC] = that wall.
  int32_t order[6];
  VsxWall at[6];
  float shift[19];         // 16-bit storage: stored value = f_i - shift[i] = f_i - feq_i(1, p.store_u)
};

// The kernel's StepCtx argument, addressed in the kernarg segment (it is always
// the first kernel parameter, at offset 0).  Taking the context by reference
// from here, instead of from the by-value parameter, keeps the compiler from
// making a private per-lane copy of the ~1.4 KB block whenever a reference to
// it reaches a function that is not inlined.
__device__ __forceinline__ const StepCtx &kernel_ctx() {
  return *reinterpret_cast<const StepCtx *>((const char *)__builtin_amdgcn_kernarg_segment_ptr());
}

// Cell index (< 2^31 cells, checked on the host).  Populations are addressed as
// (uniform base of population i) + (32-bit cell offset), so every load is a
// scalar-base + 32-bit vector-offset access.
template <class L>
__device__ __forceinline__ uint32_t cell_of(const StepCtx &c, int x, int y, int z) {
  return ((uint32_t)x * (uint32_t)c.dims[1] + (uint32_t)y) * (uint32_t)c.dims[2] + (uint32_t)z;
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

template <class L, class SRC>
__device__ __forceinline__ void pull(const StepCtx &c, const typename SRC::T *__restrict__ src, int x, int y, int z,
                                     float (&f)[L::Q]) {
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const int xs = wrap(x - L::cx[i], c.dims[0]);
    const int ys = wrap(y - L::cy[i], c.dims[1]);
    const int zs = L::D == 3 ? wrap(z - L::cz[i], c.dims[2]) : 0;
    f[i] = SRC::ld(src + (size_t)i * c.n, cell_of<L>(c, xs, ys, zs), c.shift[i]);
  }
}

template <class L, class SRC>
__device__ __forceinline__ void local(const StepCtx &c, const typename SRC::T *__restrict__ src, uint32_t cell,
                                      float (&f)[L::Q]) {
#pragma unroll
  for (int i = 0; i < L::Q; ++i) f[i] = SRC::ld(src + (size_t)i * c.n, cell, c.shift[i]);
}

// component a of a coordinate triple, by selects (no dynamic register indexing)
__device__ __forceinline__ int comp3(const int (&v)[3], int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

// coordinates of the inward neighbour across wall `loc`
template <class L>
__device__ __forceinline__ void inward(int loc, const int (&co)[3], int (&cn)[3]) {
  const int a = L::w_axis[loc], sg = L::w_sign[loc];
  cn[0] = co[0] + (a == 0 ? sg : 0);
  cn[1] = co[1] + (a == 1 ? sg : 0);
  cn[2] = co[2] + (a == 2 ? sg : 0);
}

template <class L>
__device__ __forceinline__ bool on_wall(const StepCtx &c, int loc, const int (&co)[3]) {
  const int a = L::w_axis[loc];
  return comp3(co, a) == (L::w_sign[loc] > 0 ? 0 : comp3(c.dims, a) - 1);
}

// index of a cell along the wall (matches the per-node wall arrays)
template <class L>
__device__ __forceinline__ int64_t wall_node(const StepCtx &c, int loc, const int (&co)[3]) {
  const int a = L::w_axis[loc];
  if (a == 0) return (int64_t)co[1] * c.dims[2] + co[2];
  if (a == 1) return (int64_t)co[0] * c.dims[2] + co[2];
  return (int64_t)co[0] * c.dims[1] + co[1];
}

template <class L>
__device__ __forceinline__ bool on_any_wall(const StepCtx &c, const int (&co)[3], int upto) {
  bool on = false;
#pragma unroll
  for (int k = 0; k < L::n_walls; ++k)
    if (k < upto) on |= on_wall<L>(c, c.p.walls[k].loc, co);
  return on;
}

// Post-streaming, post-walls populations of cell co after walls [0, J).
// TAG only separates instantiations: an out-of-line function takes its register
// budget from the kernels that call it, so a kernel with a tight budget calls
// its own copy.
template <class L, class SRC, int J, int TAG = 0>
__device__ __noinline__ void wall_state(const StepCtx &c, const typename SRC::T *__restrict__ src, const int (&co)[3],
                                        float (&f)[L::Q]) {
  if constexpr (J == 0) {
    pull<L, SRC>(c, src, co[0], co[1], co[2], f);
  } else {
    wall_state<L, SRC, J - 1, TAG>(c, src, co, f);
    if (J - 1 < c.p.n_walls) {
      const VsxWall &w = c.p.walls[J - 1];
      if (on_wall<L>(c, w.loc, co)) {
        float fn[L::Q], fp[L::Q];
#pragma unroll
        for (int i = 0; i < L::Q; ++i) fn[i] = fp[i] = 0.f;
        if (wall_needs_neighbor(w)) {
          int cn[3];
          inward<L>(w.loc, co, cn);
          wall_state<L, SRC, J - 1, TAG>(c, src, cn, fn);
        }
        if (wall_needs_pre(w)) local<L, SRC>(c, src, cell_of<L>(c, co[0], co[1], co[2]), fp);
        apply_wall<L>(w, wall_node<L>(c, w.loc, co), f, fn, fp);
      }
    }
  }
}

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1 (compile-time loop).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <class L, int LOC>
__device__ __forceinline__ bool on_wall_at(const StepCtx &c, const int (&co)[3]) {
  constexpr int a = L::w_axis[LOC];
  return co[a] == (L::w_sign[LOC] > 0 ? 0 : c.dims[a] - 1);
}

// Wall class of a cell: -1 on no wall; LOC when the cell is on exactly one wall
// (at location LOC) and that wall reads no neighbour rewritten by an earlier
// wall, so one pull and one wall update reproduce the ordered sequence; -2
// otherwise (corners and corner-adjacent cells, replayed recursively by the
// edge kernel).
template <class L>
__device__ __forceinline__ int wall_class(const StepCtx &c, const int (&co)[3]) {
  int loc_on = -1, count = 0;
  static_for<L::n_walls>([&](auto LC) {
    constexpr int LOC = decltype(LC)::value;
    if (c.order[LOC] != -1 && on_wall_at<L, LOC>(c, co)) {
      loc_on = LOC;
      count += c.order[LOC] == -2 ? 2 : 1;
    }
  });
  if (count == 0) return -1;
  if (count > 1) return -2;
  int result = loc_on;
  static_for<L::n_walls>([&](auto LC) {
    constexpr int LOC = decltype(LC)::value;
    if (LOC == loc_on && wall_needs_neighbor(c.at[LOC])) {
      int cn[3];
      inward<L>(LOC, co, cn);
      static_for<L::n_walls>([&](auto L2C) {
        constexpr int L2 = decltype(L2C)::value;
        if (c.order[L2] >= 0 && c.order[L2] < c.order[LOC] && on_wall_at<L, L2>(c, cn)) result = -2;
      });
    }
  });
  return result;
}

// Populations of a cell of wall class cls >= 0 (the wall location).
template <class L, class SRC>
__device__ __forceinline__ void single_wall_state(const StepCtx &c, const typename SRC::T *__restrict__ src,
                                                  const int (&co)[3], int cls, float (&f)[L::Q]) {
  pull<L, SRC>(c, src, co[0], co[1], co[2], f);
  static_for<L::n_walls>([&](auto LC) {
    constexpr int LOC = decltype(LC)::value;
    if (LOC != cls) return;
    const VsxWall &w = c.at[LOC];
    float fn[L::Q], fp[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) fn[i] = fp[i] = 0.f;
    if (wall_needs_neighbor(w)) {
      int cn[3];
      inward<L>(LOC, co, cn);
      pull<L, SRC>(c, src, cn[0], cn[1], cn[2], fn);
    }
    if (wall_needs_pre(w)) local<L, SRC>(c, src, cell_of<L>(c, co[0], co[1], co[2]), fp);
    apply_wall_at<L, LOC>(w, wall_node<L>(c, LOC, co), f, fn, fp);
  });
}

// Post-streaming, post-walls populations of any cell.
template <class L, class SRC, int TAG = 0>
__device__ __forceinline__ void cell_state(const StepCtx &c, const typename SRC::T *__restrict__ src, const int (&co)[3],
                                           float (&f)[L::Q]) {
  const int cls = c.p.n_walls > 0 ? wall_class<L>(c, co) : -1;
  if (cls == -1) pull<L, SRC>(c, src, co[0], co[1], co[2], f);
  else if (cls >= 0) single_wall_state<L, SRC>(c, src, co, cls, f);
  else wall_state<L, SRC, L::n_walls, TAG>(c, src, co, f);
}

// Whether any cell can be of class -2: two walls on different axes (shared
// corners) or two walls at the same location.
template <class L>
inline bool walls_need_edge(const VsxLbm &p) {
  for (int i = 0; i < p.n_walls; ++i)
    for (int j = i + 1; j < p.n_walls; ++j)
      if (L::w_axis[p.walls[i].loc] != L::w_axis[p.walls[j].loc] || p.walls[i].loc == p.walls[j].loc) return true;
  return false;
}

// body force at a cell: uniform + optional full field + immersed-boundary window
// (window origin `org`, used when `win`)
template <class L>
__device__ __forceinline__ void body_force(const StepCtx &c, uint32_t cell, const int (&co)[3], const int (&org)[3],
                                           bool win, float (&g)[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) g[a] = a < L::D ? c.p.g_uniform[a] : 0.f;
  if (c.p.g) {
#pragma unroll
    for (int a = 0; a < L::D; ++a) g[a] += c.p.g[(size_t)a * c.n + cell];
  }
  if (win) {
    int r[3];
    bool in = true;
#pragma unroll
    for (int a = 0; a < L::D; ++a) {
      r[a] = co[a] - org[a];
      in = in && r[a] >= 0 && r[a] < c.win.size[a];
    }
    if (in) {
      int k = r[0];
#pragma unroll
      for (int a = 1; a < L::D; ++a) k = k * c.win.size[a] + r[a];
      const int wn = c.win.size[0] * c.win.size[1] * (L::D == 3 ? c.win.size[2] : 1);
#pragma unroll
      for (int a = 0; a < L::D; ++a) {
        // 2D: the force window may come from the IB workgroup of this launch;
        // 3D: it was written by an earlier launch
        if constexpr (L::D == 2) g[a] += ld_agent(&c.win.g[a * wn + k]);
        else g[a] += c.win.g[a * wn + k];
      }
    }
  }
}

// window origin of the bulk kernels: read from the window buffer (device)
template <class L>
__device__ __forceinline__ bool window_origin_of(const StepCtx &c, int (&org)[3]) {
  org[0] = org[1] = org[2] = 0;
  if (!c.win.origin) return false;
#pragma unroll
  for (int a = 0; a < L::D; ++a) org[a] = c.win.origin[a];
  return true;
}

// collision + forcing of one cell, composed as in the reference examples:
//   Guo: u* = u + g / (2 rho) in the equilibrium and the forcing term (poiseuille_channel.py:99-128)
//   EDM: uncorrected u, f += F(g, u)                               (poiseuille_channel.py:84-97)
template <class L, int COLL>
__device__ __forceinline__ void collide_force(float (&f)[L::Q], float rho, const float (&u)[3], const float (&g)[3],
                                              int forcing, float omega, const float *rates) {
  const bool has_g = forcing != VSX_FORCE_NONE && (g[0] != 0.f || g[1] != 0.f || g[2] != 0.f);
  float ue[3] = {u[0], u[1], u[2]};
  if (has_g && forcing == VSX_FORCE_GUO) {
#pragma unroll
    for (int a = 0; a < L::D; ++a) ue[a] = u[a] + g[a] * 0.5f / rho;
  }
  float feq[L::Q];
  equilibrium<L>(rho, ue, feq);
  float F[L::Q];
  if (has_g) {
    const float (&uf)[3] = forcing == VSX_FORCE_GUO ? ue : u;
    const float ug = dot3<L>(g, uf);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) F[i] = guo_i<L>(i, uf, g, ug);
  }
  bool forced = false;
  if constexpr (COLL == VSX_COLL_BGK) {
    const float a = 1.f - omega;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) f[i] = a * f[i] + omega * feq[i];
  } else if constexpr (COLL == VSX_COLL_KBC) {
    if constexpr (L::Q == 9) kbc_d2q9(f, feq, omega);
    else kbc_d3q19<L>(f, feq, omega);
  } else if constexpr (COLL == VSX_COLL_REG) {
    if constexpr (L::Q == 9) reg_d2q9(f, feq, omega);
    else reg_d3q19<L>(f, feq, omega);
  } else {  // MRT
    if constexpr (L::Q == 9) {
      float r[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) r[k] = rates[k];
      if (has_g && forcing == VSX_FORCE_GUO) {
        mrt_d2q9<true>(f, feq, F, r);
        forced = true;
      } else {
        mrt_d2q9<false>(f, feq, F, r);
      }
    } else {
      forced = has_g && forcing == VSX_FORCE_GUO;
      mrt_d3q19(f, feq, F, forced, rates);
    }
  }
  if (has_g && !forced) {
    const float s = forcing == VSX_FORCE_GUO ? (1.f - 0.5f * omega) : 1.f;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) f[i] += F[i] * s;
  }
}

template <class L, class DST>
__device__ __forceinline__ void store_cell(const StepCtx &c, typename DST::T *__restrict__ dst, uint32_t cell,
                                           const float (&f)[L::Q]) {
#pragma unroll
  for (int i = 0; i < L::Q; ++i) DST::st(dst + (size_t)i * c.n, cell, f[i], c.shift[i]);
}

// finish() of a cell that carries no body force (no g field, no uniform g, not
// in the IB window): the forced branches are compiled out.
template <class L, int COLL>
__device__ __forceinline__ void finish_unforced(const StepCtx &c, float (&f)[L::Q]) {
  float rho, u[3];
  macroscopic<L>(f, rho, u);
  if constexpr (COLL == VSX_COLL_MRT && L::Q == 9) {
    float r[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) r[k] = c.p.mrt_rates[k];
    mrt_d2q9_unforced(f, rho, u, r);
  } else {
    const float g[3] = {0.f, 0.f, 0.f};
    collide_force<L, COLL>(f, rho, u, g, VSX_FORCE_NONE, c.p.omega, c.p.mrt_rates);
  }
}

template <class L, int COLL>
__device__ __forceinline__ void finish(const StepCtx &c, uint32_t cell, const int (&co)[3], const int (&org)[3],
                                       bool win, float (&f)[L::Q]) {
  bool unforced = !win && !c.p.g;
#pragma unroll
  for (int a = 0; a < L::D; ++a) unforced = unforced && c.p.g_uniform[a] == 0.f;
  if (unforced) {
    finish_unforced<L, COLL>(c, f);
    return;
  }
  float rho, u[3], g[3];
  macroscopic<L>(f, rho, u);
  body_force<L>(c, cell, co, org, win, g);
  collide_force<L, COLL>(f, rho, u, g, c.p.forcing, c.p.omega, c.p.mrt_rates);
}

// ------------------------------------------------------------------ kernels
// Bulk: one thread per cell; blockIdx.x tiles the contiguous axis, blockIdx.y
// (and z in 3D) the others, so a wavefront touches 64 consecutive cells.
template <class L, int COLL, int PHASE, class SRC, class DST>
__global__ __launch_bounds__(256) void k_bulk(StepCtx c_arg, const typename SRC::T *__restrict__ src,
                                              typename DST::T *__restrict__ dst) {
  const StepCtx &c = kernel_ctx();
  int co[3];
  if constexpr (L::D == 2) {
    co[1] = blockIdx.x * blockDim.x + threadIdx.x;
    co[0] = blockIdx.y;
    co[2] = 0;
    if (co[1] >= c.dims[1]) return;
  } else {
    co[2] = blockIdx.x * blockDim.x + threadIdx.x;
    co[1] = blockIdx.y;
    co[0] = blockIdx.z;
    if (co[2] >= c.dims[2]) return;
  }
  const uint32_t cell = cell_of<L>(c, co[0], co[1], co[2]);
  int org[3];
  const bool win = window_origin_of<L>(c, org);
  float f[L::Q];
  if constexpr (PHASE == VSX_PHASE_COLLIDE) {
    local<L, SRC>(c, src, cell, f);
  } else {
    const int cls = c.p.n_walls > 0 ? wall_class<L>(c, co) : -1;
    if (cls != -1) {
      // wall cell: a separate path (own finish and store) keeps the wall code's
      // register demand out of the interior path below
      if (cls < 0) return;   // corner cells: edge kernel
      single_wall_state<L, SRC>(c, src, co, cls, f);
      if constexpr (PHASE != VSX_PHASE_STREAM) finish<L, COLL>(c, cell, co, org, win, f);
      store_cell<L, DST>(c, dst, cell, f);
      return;
    }
    pull<L, SRC>(c, src, co[0], co[1], co[2], f);
  }
  if constexpr (PHASE != VSX_PHASE_STREAM) finish<L, COLL>(c, cell, co, org, win, f);
  store_cell<L, DST>(c, dst, cell, f);
}

// Edge: the cells of wall class -2, enumerated as (wall, node) pairs; a cell on
// several walls is handled under the first wall (in list order) that contains it.
template <class L, int PHASE, class SRC, class DST>
__global__ __launch_bounds__(256) void k_edge(StepCtx c_arg, const typename SRC::T *__restrict__ src,
                                              typename DST::T *__restrict__ dst, int64_t per_wall) {
  const StepCtx &c = kernel_ctx();
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int k = (int)(t / per_wall);
  const int64_t j = t % per_wall;
  if (k >= c.p.n_walls) return;
  const int loc = c.p.walls[k].loc, a = L::w_axis[loc];
  int co[3];
  // node j -> coordinates on the wall (same order as the per-node arrays)
  int b0 = a == 0 ? 1 : 0, b1 = a == 2 ? 1 : 2;
  const int n1 = L::D == 2 ? 1 : c.dims[b1];
  const int64_t nodes = (int64_t)c.dims[b0] * (L::D == 2 ? 1 : c.dims[b1]);
  if (L::D == 2) { b0 = 1 - a; b1 = 2; }
  if (j >= (L::D == 2 ? (int64_t)c.dims[1 - a] : nodes)) return;
  co[a] = L::w_sign[loc] > 0 ? 0 : c.dims[a] - 1;
  if constexpr (L::D == 2) {
    co[1 - a] = (int)j;
    co[2] = 0;
  } else {
    co[b0] = (int)(j / n1);
    co[b1] = (int)(j % n1);
  }
  if (on_any_wall<L>(c, co, k)) return;     // handled under the first wall that contains it
  if (wall_class<L>(c, co) != -2) return;   // bulk kernel
  const uint32_t cell = cell_of<L>(c, co[0], co[1], co[2]);
  float f[L::Q];
  wall_state<L, SRC, L::n_walls>(c, src, co, f);
  if constexpr (PHASE == VSX_PHASE_FULL) {
    int org[3];
    const bool win = window_origin_of<L>(c, org);
    switch (c.p.collision) {
      case VSX_COLL_BGK: finish<L, VSX_COLL_BGK>(c, cell, co, org, win, f); break;
      case VSX_COLL_MRT: finish<L, VSX_COLL_MRT>(c, cell, co, org, win, f); break;
      case VSX_COLL_KBC: finish<L, VSX_COLL_KBC>(c, cell, co, org, win, f); break;
      default: finish<L, VSX_COLL_REG>(c, cell, co, org, win, f); break;
    }
  }
  store_cell<L, DST>(c, dst, cell, f);
}

}  // namespace vsx
