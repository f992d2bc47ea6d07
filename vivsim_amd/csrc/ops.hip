// Standalone kernels behind the pure-function API of vivsim.lbm / vivsim.lbm3d.
// Each entry point is one reference function on arrays in the reference layout
// ([Q][n] populations, [D][n] vectors, [n] scalars; n = product of the spatial
// shape).  Cell-parallel, one thread per cell, grid-stride; the Q loads of a
// thread are strided by n so every wavefront reads 64 consecutive floats per
// population (coalesced 256-byte segments).
#include <cstdarg>
#include <cstring>

#include "cell.hpp"
#include "common.hpp"
#include "wall.hpp"

namespace vsx {

template <int N>
struct Mat {
  float m[N * N];
};

constexpr int kBlock = 256;

#define VSX_GRID_STRIDE(idx, n) \
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < (n); idx += (int64_t)gridDim.x * blockDim.x)

template <class L>
__device__ __forceinline__ void load_q(const float *__restrict__ f, int64_t n, int64_t c, float (&v)[L::Q]) {
#pragma unroll
  for (int i = 0; i < L::Q; ++i) v[i] = f[i * n + c];
}
template <class L>
__device__ __forceinline__ void store_q(float *__restrict__ f, int64_t n, int64_t c, const float (&v)[L::Q]) {
#pragma unroll
  for (int i = 0; i < L::Q; ++i) f[i * n + c] = v[i];
}
template <class L>
__device__ __forceinline__ void load_vec(const float *__restrict__ u, int64_t n, int64_t c, float (&v)[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) v[a] = a < L::D ? u[a * n + c] : 0.f;
}

// ------------------------------------------------------------------ kernels

template <class L>
__global__ void k_macroscopic(const float *__restrict__ f, float *__restrict__ rho, float *__restrict__ u, int64_t n) {
  VSX_GRID_STRIDE(c, n) {
    float v[L::Q], r, uu[3];
    load_q<L>(f, n, c, v);
    macroscopic<L>(v, r, uu);
    rho[c] = r;
#pragma unroll
    for (int a = 0; a < L::D; ++a) u[a * n + c] = uu[a];
  }
}

template <class L>
__global__ void k_equilibrium(const float *__restrict__ rho, const float *__restrict__ u, float *__restrict__ feq, int64_t n) {
  VSX_GRID_STRIDE(c, n) {
    float uu[3], v[L::Q];
    load_vec<L>(u, n, c, uu);
    equilibrium<L>(rho[c], uu, v);
    store_q<L>(feq, n, c, v);
  }
}

template <class L>
__global__ void k_bgk(const float *__restrict__ f, const float *__restrict__ feq, float omega, float *__restrict__ out, int64_t n) {
  const float a = 1.f - omega;
  VSX_GRID_STRIDE(c, n) {
#pragma unroll
    for (int i = 0; i < L::Q; ++i) out[i * n + c] = a * f[i * n + c] + omega * feq[i * n + c];
  }
}

// out = f + op (feq - f)  with a dense Q x Q operator (the reference's own matvec)
template <class L>
__global__ void k_mrt(const float *__restrict__ f, const float *__restrict__ feq, Mat<L::Q> op, float *__restrict__ out, int64_t n) {
  VSX_GRID_STRIDE(c, n) {
    float d[L::Q], v[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
      v[i] = f[i * n + c];
      d[i] = feq[i * n + c] - v[i];
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
      float s = 0.f;
#pragma unroll
      for (int j = 0; j < L::Q; ++j) s += op.m[i * L::Q + j] * d[j];
      out[i * n + c] = v[i] + s;
    }
  }
}

template <class L>
__global__ void k_kbc(const float *__restrict__ f, const float *__restrict__ feq, float omega, float *__restrict__ out, int64_t n) {
  VSX_GRID_STRIDE(c, n) {
    float v[L::Q], e[L::Q];
    load_q<L>(f, n, c, v);
    load_q<L>(feq, n, c, e);
    if constexpr (L::Q == 9) kbc_d2q9(v, e, omega);
    else kbc_d3q19<L>(v, e, omega);
    store_q<L>(out, n, c, v);
  }
}

template <class L>
__global__ void k_reg(const float *__restrict__ f, const float *__restrict__ feq, float omega, float *__restrict__ out, int64_t n) {
  VSX_GRID_STRIDE(c, n) {
    float v[L::Q], e[L::Q];
    load_q<L>(f, n, c, v);
    load_q<L>(feq, n, c, e);
    if constexpr (L::Q == 9) reg_d2q9(v, e, omega);
    else reg_d3q19<L>(v, e, omega);
    store_q<L>(out, n, c, v);
  }
}

// mode 0: out = F ; mode 1: out = f + F * scale ; mode 2: out = f + op F
template <class L, int MODE>
__global__ void k_guo(const float *__restrict__ f, const float *__restrict__ g, const float *__restrict__ u, float scale,
                      Mat<L::Q> op, float *__restrict__ out, int64_t n) {
  VSX_GRID_STRIDE(c, n) {
    float uu[3], gg[3], F[L::Q];
    load_vec<L>(u, n, c, uu);
    load_vec<L>(g, n, c, gg);
    const float ug = dot3<L>(gg, uu);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) F[i] = guo_i<L>(i, uu, gg, ug);
    if constexpr (MODE == 0) {
      store_q<L>(out, n, c, F);
    } else if constexpr (MODE == 1) {
#pragma unroll
      for (int i = 0; i < L::Q; ++i) out[i * n + c] = f[i * n + c] + F[i] * scale;
    } else {
#pragma unroll
      for (int i = 0; i < L::Q; ++i) {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < L::Q; ++j) s += op.m[i * L::Q + j] * F[j];
        out[i * n + c] = f[i * n + c] + s;
      }
    }
  }
}

template <class L>
__global__ void k_obstacle(const float *__restrict__ f, const uint8_t *__restrict__ mask, float *__restrict__ out, int64_t n) {
  VSX_GRID_STRIDE(c, n) {
    float v[L::Q];
    load_q<L>(f, n, c, v);
    const bool m = mask[c] != 0;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) out[i * n + c] = m ? v[L::opp[i]] : v[i];
  }
}

// periodic pull streaming f'_i(x) = f_i(x - c_i)
template <class L>
__global__ void k_stream(const float *__restrict__ f, float *__restrict__ out, int nx, int ny, int nz) {
  const int64_t n = (int64_t)nx * ny * nz;
  VSX_GRID_STRIDE(c, n) {
    const int z = (int)(c % nz);
    const int y = (int)((c / nz) % ny);
    const int x = (int)(c / ((int64_t)nz * ny));
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
      int xs = x - L::cx[i], ys = y - L::cy[i], zs = z - L::cz[i];
      xs += xs < 0 ? nx : (xs >= nx ? -nx : 0);
      ys += ys < 0 ? ny : (ys >= ny ? -ny : 0);
      zs += zs < 0 ? nz : (zs >= nz ? -nz : 0);
      out[i * n + c] = f[i * n + ((int64_t)xs * ny + ys) * nz + zs];
    }
  }
}

__global__ void k_vcorr(const float *__restrict__ g, const float *__restrict__ rho, float rs, float *__restrict__ out,
                        int comps, int64_t n) {
  VSX_GRID_STRIDE(i, (int64_t)comps * n) {
    const float r = rho ? rho[i % n] : rs;
    out[i] = g[i] * 0.5f / r;
  }
}

// ------------------------------------------------------------------ walls on a full field

struct Dims {
  int n[3];
};

// cell index of wall node j at layer `layer` (0 = wall line, 1 = first inward neighbour)
template <class L>
__device__ __forceinline__ int64_t wall_cell(int loc, int64_t j, int layer, Dims d) {
  const int a = L::w_axis[loc], sg = L::w_sign[loc];
  const int idx = sg > 0 ? layer : d.n[a] - 1 - layer;
  int co[3];
  if (a == 0) { co[0] = idx; co[1] = (int)(j / d.n[2]); co[2] = (int)(j % d.n[2]); }
  else if (a == 1) { co[1] = idx; co[0] = (int)(j / d.n[2]); co[2] = (int)(j % d.n[2]); }
  else { co[2] = idx; co[0] = (int)(j / d.n[1]); co[1] = (int)(j % d.n[1]); }
  return ((int64_t)co[0] * d.n[1] + co[1]) * d.n[2] + co[2];
}

template <class L>
__global__ void k_wall(VsxWall w, const float *__restrict__ fpre, const float *f, float *out, Dims d, int64_t nw) {
  const int64_t n = (int64_t)d.n[0] * d.n[1] * d.n[2];
  VSX_GRID_STRIDE(j, nw) {
    float v[L::Q], fn[L::Q], fp[L::Q];
    const int64_t cw = wall_cell<L>(w.loc, j, 0, d);
    load_q<L>(f, n, cw, v);
    if (wall_needs_neighbor(w)) load_q<L>(f, n, wall_cell<L>(w.loc, j, 1, d), fn);
    else {
#pragma unroll
      for (int i = 0; i < L::Q; ++i) fn[i] = 0.f;
    }
    if (wall_needs_pre(w)) load_q<L>(fpre, n, cw, fp);
    else {
#pragma unroll
      for (int i = 0; i < L::Q; ++i) fp[i] = 0.f;
    }
    apply_wall<L>(w, j, v, fn, fp);
    store_q<L>(out, n, cw, v);
  }
}

// characteristic BC (lbm/boundary/cbc.py:14-53, lbm3d/boundary/cbc.py:16-51)
template <class L>
__global__ void k_cbc(const float *__restrict__ rho, const float *__restrict__ u, int loc, float *__restrict__ rho_next,
                      float *__restrict__ u_next, Dims d, int64_t nw) {
  const int64_t n = (int64_t)d.n[0] * d.n[1] * d.n[2];
  const int a = L::w_axis[loc], sg = L::w_sign[loc];
  const float cs = 0.57735026918962576f;  // 1 / sqrt(3)
  VSX_GRID_STRIDE(j, nw) {
    const int64_t c1 = wall_cell<L>(loc, j, 0, d), c2 = wall_cell<L>(loc, j, 1, d), c3 = wall_cell<L>(loc, j, 2, d);
    const float r1 = rho[c1], r2 = rho[c2], r3 = rho[c3];
    const float n1 = u[a * n + c1], n2 = u[a * n + c2], n3 = u[a * n + c3];
    const float k = -0.5f * sg;
    const float drho = k * (3.f * r1 - 4.f * r2 + r3);
    const float dun = k * (3.f * n1 - 4.f * n2 + n3);
    const float ch = dun - sg * cs / r1 * drho;
    const float lo = (n1 - sg * cs) * ch;
    rho_next[j] = r1 - 0.5f * r1 / cs * lo;
#pragma unroll
    for (int b = 0; b < L::D; ++b) u_next[b * nw + j] = (b == a) ? n1 - 0.5f * lo : u[b * n + c1];
  }
}

// ------------------------------------------------------------------ host helpers

template <class L>
Mat<L::Q> host_mat(const float *op) {
  Mat<L::Q> m;
  std::memcpy(m.m, op, sizeof(m.m));
  return m;
}

template <class L>
int64_t wall_nodes(int loc, Dims d) {
  const int a = L::w_axis[loc];
  int64_t p = 1;
  for (int b = 0; b < 3; ++b)
    if (b != a) p *= d.n[b];
  return p;
}

template <class L>
int check_wall(const VsxWall *w, Dims d) {
  if (!w) return fail(VSX_ERR_VALUE, "wall is NULL");
  if (w->loc < 0 || w->loc >= L::n_walls) return fail(VSX_ERR_KEY, "unknown wall location %d", w->loc);
  if (w->kind < VSX_WALL_NEE || w->kind > VSX_WALL_SPECULAR) return fail(VSX_ERR_VALUE, "unknown wall kind %d", w->kind);
  if (w->mode < VSX_MODE_PLAIN || w->mode > VSX_MODE_FORCE_CORRECTED) return fail(VSX_ERR_VALUE, "unknown wall mode %d", w->mode);
  if (d.n[L::w_axis[w->loc]] < 3) return fail(VSX_ERR_VALUE, "grid too small along the wall normal");
  return VSX_OK;
}

template <class L>
int boundary(const VsxWall *w, const float *fpre, const float *f, float *out, Dims d, void *stream) {
  int rc = check_wall<L>(w, d);
  if (rc) return rc;
  VSX_REQUIRE(f && out, "null field pointer");
  VSX_REQUIRE(!wall_needs_pre(*w) || fpre, "bounce-back needs the pre-streaming field");
  const int64_t n = (int64_t)d.n[0] * d.n[1] * d.n[2];
  hipStream_t s = as_stream(stream);
  if (out != f) VSX_HIP(hipMemcpyAsync(out, f, sizeof(float) * L::Q * n, hipMemcpyDeviceToDevice, s));
  const int64_t nw = wall_nodes<L>(w->loc, d);
  hipLaunchKernelGGL(k_wall<L>, dim3(grid_for(nw, kBlock)), dim3(kBlock), 0, s, *w, fpre, f, out, d, nw);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

template <class L>
int cbc(const float *rho, const float *u, int loc, float *rn, float *un, Dims d, void *stream) {
  if (loc < 0 || loc >= L::n_walls) return fail(VSX_ERR_VALUE, "unknown wall location %d", loc);
  VSX_REQUIRE(d.n[L::w_axis[loc]] >= 3, "need 3 nodes along the wall normal");
  const int64_t nw = wall_nodes<L>(loc, d);
  hipLaunchKernelGGL(k_cbc<L>, dim3(grid_for(nw, kBlock)), dim3(kBlock), 0, as_stream(stream), rho, u, loc, rn, un, d, nw);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

}  // namespace vsx

using namespace vsx;

#define LAUNCH(kern, n, ...)                                                                          \
  do {                                                                                               \
    if ((n) <= 0) return VSX_OK;                                                                     \
    hipLaunchKernelGGL(kern, dim3(grid_for((n), kBlock)), dim3(kBlock), 0, as_stream(stream), __VA_ARGS__); \
    VSX_CHECK_LAUNCH();                                                                              \
    return VSX_OK;                                                                                   \
  } while (0)

#define VSX_POINTWISE_API(pfx, L)                                                                                     \
  extern "C" int vsx_##pfx##_get_macroscopic(const float *f, float *rho, float *u, int64_t n, void *stream) {          \
    LAUNCH(k_macroscopic<L>, n, f, rho, u, n);                                                                        \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_get_equilibrium(const float *rho, const float *u, float *feq, int64_t n, void *stream) {  \
    LAUNCH(k_equilibrium<L>, n, rho, u, feq, n);                                                                      \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_collision_bgk(const float *f, const float *feq, float omega, float *out, int64_t n,      \
                                           void *stream) {                                                            \
    LAUNCH(k_bgk<L>, n, f, feq, omega, out, n);                                                                       \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_collision_mrt(const float *f, const float *feq, const float *op, float *out, int64_t n,  \
                                           void *stream) {                                                            \
    if (!op) return fail(VSX_ERR_VALUE, "null operator");                                                            \
    LAUNCH(k_mrt<L>, n, f, feq, host_mat<L>(op), out, n);                                                             \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_collision_kbc(const float *f, const float *feq, float omega, float *out, int64_t n,      \
                                           void *stream) {                                                            \
    LAUNCH(k_kbc<L>, n, f, feq, omega, out, n);                                                                       \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_collision_reg(const float *f, const float *feq, float omega, float *out, int64_t n,      \
                                           void *stream) {                                                            \
    LAUNCH(k_reg<L>, n, f, feq, omega, out, n);                                                                       \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_get_guo_forcing_term(const float *g, const float *u, float *out, int64_t n,              \
                                                  void *stream) {                                                     \
    LAUNCH((k_guo<L, 0>), n, nullptr, g, u, 0.f, Mat<L::Q>{}, out, n);                                                \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_forcing_guo_bgk(const float *f, const float *g, const float *u, float omega, float *out, \
                                             int64_t n, void *stream) {                                               \
    LAUNCH((k_guo<L, 1>), n, f, g, u, (float)(1 - 0.5 * (double)omega), Mat<L::Q>{}, out, n);                         \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_forcing_guo_mrt(const float *f, const float *g, const float *u, const float *op,         \
                                             float *out, int64_t n, void *stream) {                                   \
    if (!op) return fail(VSX_ERR_VALUE, "null operator");                                                            \
    LAUNCH((k_guo<L, 2>), n, f, g, u, 0.f, host_mat<L>(op), out, n);                                                  \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_forcing_edm(const float *f, const float *g, const float *u, float *out, int64_t n,       \
                                         void *stream) {                                                              \
    LAUNCH((k_guo<L, 1>), n, f, g, u, 1.f, Mat<L::Q>{}, out, n);                                                      \
  }                                                                                                                   \
  extern "C" int vsx_##pfx##_obstacle_bounce_back(const float *f, const uint8_t *mask, float *out, int64_t n,         \
                                                  void *stream) {                                                     \
    LAUNCH(k_obstacle<L>, n, f, mask, out, n);                                                                        \
  }

VSX_POINTWISE_API(d2q9, D2Q9)
VSX_POINTWISE_API(d3q19, D3Q19)

extern "C" int vsx_velocity_correction(const float *g, const float *rho, float rho_scalar, float *out, int32_t comps,
                                       int64_t n, void *stream) {
  VSX_REQUIRE(comps > 0 && n >= 0, "bad sizes");
  LAUNCH(k_vcorr, (int64_t)comps * n, g, rho, rho_scalar, out, comps, n);
}

extern "C" int vsx_d2q9_streaming(const float *f, float *out, int32_t nx, int32_t ny, void *stream) {
  VSX_REQUIRE(nx > 0 && ny > 0, "bad grid %d x %d", nx, ny);
  VSX_REQUIRE(f != out, "streaming cannot run in place");
  const int64_t n = (int64_t)nx * ny;
  LAUNCH(k_stream<D2Q9>, n, f, out, nx, ny, 1);
}

extern "C" int vsx_d3q19_streaming(const float *f, float *out, int32_t nx, int32_t ny, int32_t nz, void *stream) {
  VSX_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad grid %d x %d x %d", nx, ny, nz);
  VSX_REQUIRE(f != out, "streaming cannot run in place");
  const int64_t n = (int64_t)nx * ny * nz;
  LAUNCH(k_stream<D3Q19>, n, f, out, nx, ny, nz);
}

extern "C" int vsx_d2q9_boundary(const VsxWall *wall, const float *f_pre, const float *f, float *out, int32_t nx,
                                 int32_t ny, void *stream) {
  return boundary<D2Q9>(wall, f_pre, f, out, Dims{{nx, ny, 1}}, stream);
}

extern "C" int vsx_d3q19_boundary(const VsxWall *wall, const float *f_pre, const float *f, float *out, int32_t nx,
                                  int32_t ny, int32_t nz, void *stream) {
  return boundary<D3Q19>(wall, f_pre, f, out, Dims{{nx, ny, nz}}, stream);
}

extern "C" int vsx_d2q9_boundary_characteristic(const float *rho, const float *u, int32_t loc, float *rho_next,
                                                float *u_next, int32_t nx, int32_t ny, void *stream) {
  return cbc<D2Q9>(rho, u, loc, rho_next, u_next, Dims{{nx, ny, 1}}, stream);
}

extern "C" int vsx_d3q19_boundary_characteristic(const float *rho, const float *u, int32_t loc, float *rho_next,
                                                 float *u_next, int32_t nx, int32_t ny, int32_t nz, void *stream) {
  return cbc<D3Q19>(rho, u, loc, rho_next, u_next, Dims{{nx, ny, nz}}, stream);
}
