// Domain-wall boundary conditions for one wall cell, generic over the lattice.
//
// A wall rewrites the populations of its cells from (a) the current populations
// of the cell, (b) for NEE / pressure modes, the current populations of the
// inward neighbour cell, (c) for bounce-back / specular, the pre-streaming
// populations of the cell, and (d) the wall data (scalar or per-node arrays).
//   NEE          lbm/boundary/nee.py:19-60          lbm3d/boundary/nee.py:13-33
//   NEBB         lbm/boundary/nebb.py:20-58 (Zou/He) lbm3d/boundary/nebb.py:13-30 (generic)
//   equilibrium  lbm/boundary/eq.py:25-56           lbm3d/boundary/eq.py:13-26
//   bounce-back  lbm/boundary/bb.py:15-95           lbm3d/boundary/bb.py:9-53
//   wrappers     lbm/boundary/_helpers.py:80-239    lbm3d/boundary/_helpers.py:35-133
#pragma once
#include "../../include/vivsim_amd.h"
#include "cell.hpp"
#include "lattice.hpp"

namespace vsx {

template <class L>
__host__ __device__ inline int wall_axis(int loc) { return L::w_axis[loc]; }

__device__ __forceinline__ float wv(const float *arr, float s, int64_t j) { return arr ? arr[j] : s; }

__host__ __device__ inline bool wall_needs_neighbor(const VsxWall &w) {
  return w.kind == VSX_WALL_NEE || (w.mode == VSX_MODE_PRESSURE &&
                                    (w.kind == VSX_WALL_NEBB || w.kind == VSX_WALL_EQ));
}
__host__ __device__ inline bool wall_needs_pre(const VsxWall &w) {
  return w.kind == VSX_WALL_BB || w.kind == VSX_WALL_SPECULAR;
}

// sum of wall-parallel populations + 2 * sum of outgoing ones
// (lbm/boundary/_helpers.py:135-145, lbm3d/boundary/_helpers.py:35-42)
template <class L, int LOC>
__device__ __forceinline__ float rho_wall_numerator(const float (&f)[L::Q]) {
  constexpr int a = L::w_axis[LOC], sg = L::w_sign[LOC];
  float z = 0.f, o = 0.f;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const int cn = cax<L>(i, a) * sg;
    if (cn == 0) z += f[i];
    if (cn < 0) o += f[i];
  }
  return z + 2.f * o;
}

// Resolve (rho_w, u_w) at wall node j for the wrapper mode of the wall.
template <class L, int LOC>
__device__ __forceinline__ void wall_state(const VsxWall &w, int64_t j, const float (&f)[L::Q], const float (&fn)[L::Q],
                                           float &rho, float (&u)[3]) {
  constexpr int a = L::w_axis[LOC], sg = L::w_sign[LOC];
  rho = wv(w.rho_arr, w.rho, j);
  u[0] = wv(w.ux_arr, w.ux, j);
  u[1] = wv(w.uy_arr, w.uy, j);
  u[2] = wv(w.uz_arr, w.uz, j);
  if (w.mode == VSX_MODE_FORCE_CORRECTED) {
    // u - 0.5 g / rho   (_helpers.py:156-177)
    u[0] = u[0] - wv(w.gx_arr, w.gx, j) * 0.5f / rho;
    u[1] = u[1] - wv(w.gy_arr, w.gy, j) * 0.5f / rho;
    u[2] = u[2] - wv(w.gz_arr, w.gz, j) * 0.5f / rho;
  } else if (w.mode == VSX_MODE_VELOCITY) {
    // rho = numerator / (1 - u_n)   (_helpers.py:80-95)
    const float un = sg * u[a];
    rho = rho_wall_numerator<L, LOC>(f) / (1.f - un);
  } else if (w.mode == VSX_MODE_PRESSURE) {
    // u_n from the prescribed density, tangential velocity of the neighbour (_helpers.py:98-132)
    float rn, un_b[3];
    macroscopic<L>(fn, rn, un_b);
    u[0] = un_b[0];
    u[1] = un_b[1];
    u[2] = un_b[2];
    u[a] = sg * (1.f - rho_wall_numerator<L, LOC>(f) / rho);
  }
}

template <class L, int LOC>
__device__ __forceinline__ void apply_wall_at(const VsxWall &w, int64_t j, float (&f)[L::Q], const float (&fn)[L::Q],
                                              const float (&fpre)[L::Q]) {
  constexpr int a = L::w_axis[LOC], sg = L::w_sign[LOC];
  if (w.kind == VSX_WALL_BB || w.kind == VSX_WALL_SPECULAR) {
    float uw[3] = {wv(w.ux_arr, w.ux, j), wv(w.uy_arr, w.uy, j), wv(w.uz_arr, w.uz, j)};
    if constexpr (L::D == 2) {
      // in_k := pre[out_k] + {2/3 u_n, (u_n + u_t)/6, (u_n - u_t)/6}; specular swaps the diagonal targets
      const float un = sg * uw[a], ut = sg * uw[1 - a];
      const float v0 = fpre[L::w_out[LOC][0]] + 2.f / 3.f * un;
      const float v1 = fpre[L::w_out[LOC][1]] + 1.f / 6.f * (un + ut);
      const float v2 = fpre[L::w_out[LOC][2]] + 1.f / 6.f * (un - ut);
      f[L::w_in[LOC][0]] = v0;
      const bool bb = w.kind == VSX_WALL_BB;
      f[L::w_in[LOC][1]] = bb ? v1 : v2;
      f[L::w_in[LOC][2]] = bb ? v2 : v1;
    } else {
      // in_k := pre[out_k] + 2 w_in rho_pre (c_in . u_w) / cs2 (bounce-back and specular share the body)
      float rho = 0.f;
#pragma unroll
      for (int i = 0; i < L::Q; ++i) rho += fpre[i];
#pragma unroll
      for (int k = 0; k < L::n_in; ++k) {
        const int i = L::w_in[LOC][k];
        f[i] = fpre[L::w_out[LOC][k]] + 2.0f * L::w[i] * rho * cdot<L>(i, uw) * 3.0f;
      }
    }
    return;
  }

  float rho, u[3];
  wall_state<L, LOC>(w, j, f, fn, rho, u);

  if (w.kind == VSX_WALL_EQ) {
    equilibrium<L>(rho, u, f);
  } else if (w.kind == VSX_WALL_NEE) {
    // wall := feq(rho_w, u_w) + (f_nb - feq(macro(f_nb)))
    float rn, un_b[3];
    macroscopic<L>(fn, rn, un_b);
    const float u2w = usq<L>(u), u2n = usq<L>(un_b);
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
      f[i] = feq_i<L>(i, rho, u, u2w) + (fn[i] - feq_i<L>(i, rn, un_b, u2n));
  } else {  // NEBB
    if constexpr (L::D == 2) {
      const float un = sg * u[a], ut = sg * u[1 - a];
      const float shear = 0.5f * (f[L::w_tan[LOC][0]] - f[L::w_tan[LOC][1]]) * (float)sg;
      const float nt = 1.f / 6.f * un * rho, tt = 0.5f * ut * rho;
      const float v0 = f[L::w_out[LOC][0]] + 2.f / 3.f * un * rho;
      const float v1 = f[L::w_out[LOC][1]] - shear + nt + tt;
      const float v2 = f[L::w_out[LOC][2]] + shear + nt - tt;
      f[L::w_in[LOC][0]] = v0;
      f[L::w_in[LOC][1]] = v1;
      f[L::w_in[LOC][2]] = v2;
    } else {
      // f_in := f_opp(in) + feq_in - feq_opp(in)
      const float u2 = usq<L>(u);
#pragma unroll
      for (int k = 0; k < L::n_in; ++k) {
        const int i = L::w_in[LOC][k], o = L::opp[i];
        f[i] = f[o] + feq_i<L>(i, rho, u, u2) - feq_i<L>(o, rho, u, u2);
      }
    }
  }
}

// Apply wall w at its node j: dispatch on the location so that every lattice
// table lookup is a compile-time constant (no dynamic register indexing).
template <class L>
__device__ __forceinline__ void apply_wall(const VsxWall &w, int64_t j, float (&f)[L::Q], const float (&fn)[L::Q],
                                           const float (&fpre)[L::Q]) {
  switch (w.loc) {
    case 0: apply_wall_at<L, 0>(w, j, f, fn, fpre); break;
    case 1: apply_wall_at<L, 1>(w, j, f, fn, fpre); break;
    case 2: apply_wall_at<L, 2>(w, j, f, fn, fpre); break;
    case 3: apply_wall_at<L, 3>(w, j, f, fn, fpre); break;
    default:
      if constexpr (L::n_walls > 4) {
        if (w.loc == 4) apply_wall_at<L, 4>(w, j, f, fn, fpre);
        else apply_wall_at<L, 5>(w, j, f, fn, fpre);
      }
      break;
  }
}

}  // namespace vsx
