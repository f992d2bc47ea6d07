// Standalone immersed-boundary, geometry and structural kernels behind the
// pure-function API of vivsim.ib, vivsim.ib3d and vivsim.dyn.
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "ib_common.hpp"

namespace vsx {

constexpr int kB = 256;

#define GS_LOOP(idx, n) \
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < (n); idx += (int64_t)gridDim.x * blockDim.x)

__global__ void k_fill(float *p, float v, int64_t n) {
  GS_LOOP(i, n) p[i] = v;
}

__global__ void k_delta(int kind, const float *__restrict__ r, float *__restrict__ out, int64_t n) {
  GS_LOOP(i, n) out[i] = delta_kernel(kind, r[i]);
}

// ib/stencil.py:7-51 — one thread per marker, NS x NS points, x-major / y-minor
template <int NS>
__global__ void k_stencil2d(const float *__restrict__ mx, const float *__restrict__ my, int m, int ny, int kind,
                            float *__restrict__ W, int32_t *__restrict__ I) {
  GS_LOOP(k, (int64_t)m) {
    const Axis<NS> ax = axis_stencil<NS>(kind, mx[k]);
    const Axis<NS> ay = axis_stencil<NS>(kind, my[k]);
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const int s = i * NS + j;
        W[k * NS * NS + s] = ax.w[i] * ay.w[j];
        I[k * NS * NS + s] = (ax.ref - NS / 2 + 1 + i) * ny + (ay.ref - NS / 2 + 1 + j);
      }
  }
}

// ib3d/stencil.py:7-59 — idx = sx*(ny*nz) + sy*nz + sz
template <int NS>
__global__ void k_stencil3d(const float *__restrict__ xyz, int m, int ny, int nz, int kind, float *__restrict__ W,
                            int32_t *__restrict__ I) {
  GS_LOOP(k, (int64_t)m) {
    const Axis<NS> ax = axis_stencil<NS>(kind, xyz[3 * k]);
    const Axis<NS> ay = axis_stencil<NS>(kind, xyz[3 * k + 1]);
    const Axis<NS> az = axis_stencil<NS>(kind, xyz[3 * k + 2]);
    constexpr int P = NS * NS * NS;
    for (int i = 0; i < NS; ++i)
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int s = (i * NS + j) * NS + l;
          W[k * P + s] = ax.w[i] * ay.w[j] * az.w[l];
          I[k * P + s] = (ax.ref - NS / 2 + 1 + i) * (ny * nz) + (ay.ref - NS / 2 + 1 + j) * nz + (az.ref - NS / 2 + 1 + l);
        }
  }
}

// out[m][c] (=|+=) scale * sum_s W[m][s] grid[c][I[m][s]]   (ib/stencil.py:54-78)
template <bool ACC>
__global__ void k_interp(const float *__restrict__ grid, int c, int64_t ng, const float *__restrict__ W,
                         const int32_t *__restrict__ I, int m, int s, float scale, float *__restrict__ out) {
  GS_LOOP(t, (int64_t)m * c) {
    const int k = (int)(t / c), comp = (int)(t % c);
    float acc = 0.f;
    for (int q = 0; q < s; ++q) {
      const int64_t g = gather_index(I[(int64_t)k * s + q], ng);
      acc += W[(int64_t)k * s + q] * (grid[comp * ng + g] * scale);
    }
    if (ACC) out[t] += acc;
    else out[t] = acc;
  }
}

// grid[c][I[m][s]] += marker[m][c] * W[m][s]   (ib/stencil.py:81-110; float atomics)
__global__ void k_spread(const float *__restrict__ mv, float *grid, int c, int64_t ng, const float *__restrict__ W,
                         const int32_t *__restrict__ I, int m, int s) {
  GS_LOOP(t, (int64_t)m * s) {
    const int k = (int)(t / s);
    const int64_t g = scatter_index(I[t], ng);
    if (g < 0) continue;
    const float w = W[t];
    for (int comp = 0; comp < c; ++comp) atomicAdd(&grid[comp * ng + g], mv[(int64_t)k * c + comp] * w);
  }
}

// zero the grid entries a stencil touches (cheaper than clearing the whole grid)
__global__ void k_touch_zero(float *grid, int c, int64_t ng, const int32_t *__restrict__ I, int64_t ms) {
  GS_LOOP(t, ms) {
    const int64_t g = scatter_index(I[t], ng);
    if (g < 0) continue;
    for (int comp = 0; comp < c; ++comp) grid[comp * ng + g] = 0.f;
  }
}

// MDF marker update: dF = (U - ub) * (2 ds); F += dF   (ib/mdf.py:51-53)
__global__ void k_mdf_update(const float *__restrict__ ub, const float *__restrict__ U, const float *__restrict__ ds_arr,
                             float ds, int m, int c, float *__restrict__ F, float *__restrict__ dF) {
  GS_LOOP(t, (int64_t)m * c) {
    const int k = (int)(t / c);
    const float d2 = (ds_arr ? ds_arr[k] : ds) * 2.0f;
    const float v = (U[t] - ub[t]) * d2;
    dF[t] = v;
    F[t] += v;
  }
}

__global__ void k_neg(const float *__restrict__ a, float *__restrict__ out, int64_t n) {
  GS_LOOP(i, n) out[i] = -a[i];
}

// ------------------------------------------------------------------ reductions (FP64)
template <int BS>
__device__ double block_sum(double v) {
  __shared__ double red[BS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BS / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// shoelace area 0.5 |sum x_i y_{i-1} - y_i x_{i-1}|   (ib/geometry.py:6-19)
__global__ void k_area(const float *__restrict__ p, int m, float *out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const int j = (i + m - 1) % m;
    acc += (double)p[2 * i] * p[2 * j + 1] - (double)p[2 * i + 1] * p[2 * j];
  }
  const double s = block_sum<kB>(acc);
  if (threadIdx.x == 0) *out = (float)(0.5 * fabs(s));
}

// ib/geometry.py:22-43
__global__ void k_ds(const float *__restrict__ p, int m, int closed, float *__restrict__ out) {
  GS_LOOP(i, (int64_t)m) {
    auto seg = [&](int a, int b) { return hypotf(p[2 * a] - p[2 * b], p[2 * a + 1] - p[2 * b + 1]); };
    if (closed) {
      const int nx = (int)((i + 1) % m), pv = (int)((i + m - 1) % m);
      out[i] = (seg((int)i, nx) + seg(pv, (int)i)) / 2.0f;
    } else {
      const float a = i > 0 ? seg((int)i, (int)i - 1) / 2.f : 0.f;
      const float b = i + 1 < m ? seg((int)i + 1, (int)i) / 2.f : 0.f;
      out[i] = a + b;
    }
  }
}

// vertices of triangle k; vertex indices use gather semantics (wrap negative once, clamp)
__device__ __forceinline__ void tri(const float *v, int nv, const int32_t *f, int k, double (&a)[3], double (&b)[3],
                                    double (&c)[3]) {
  const int64_t i0 = gather_index(f[3 * k], nv), i1 = gather_index(f[3 * k + 1], nv), i2 = gather_index(f[3 * k + 2], nv);
  for (int d = 0; d < 3; ++d) {
    a[d] = v[3 * i0 + d];
    b[d] = v[3 * i1 + d];
    c[d] = v[3 * i2 + d];
  }
}

__device__ __forceinline__ double tri_area(const double (&a)[3], const double (&b)[3], const double (&c)[3]) {
  const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2], z = e1[0] * e2[1] - e1[1] * e2[0];
  return 0.5 * sqrt(x * x + y * y + z * z);
}

// ib3d/geometry.py:6-18
__global__ void k_tri_areas(const float *__restrict__ v, int nv, const int32_t *__restrict__ f, int nf,
                            float *__restrict__ out) {
  GS_LOOP(k, (int64_t)nf) {
    double a[3], b[3], c[3];
    tri(v, nv, f, (int)k, a, b, c);
    out[k] = (float)tri_area(a, b, c);
  }
}

// ib3d/geometry.py:21-45 (mode 0: total area, 1: |signed volume|)
__global__ void k_mesh_sum(const float *__restrict__ v, int nv, const int32_t *__restrict__ f, int nf, int mode,
                           float *out) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < nf; k += blockDim.x) {
    double a[3], b[3], c[3];
    tri(v, nv, f, k, a, b, c);
    if (mode == 0) acc += tri_area(a, b, c);
    else acc += (a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) + a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
  }
  const double s = block_sum<kB>(acc);
  if (threadIdx.x == 0) *out = (float)(mode == 0 ? s : fabs(s));
}

// ib3d/geometry.py:48-67 (area / 3 to each vertex of each triangle)
__global__ void k_vertex_ds(const float *__restrict__ v, int nv, const int32_t *__restrict__ f, int nf, float *out) {
  GS_LOOP(k, (int64_t)nf) {
    double a[3], b[3], c[3];
    tri(v, nv, f, (int)k, a, b, c);
    const float third = (float)(tri_area(a, b, c) / 3.0);
    for (int q = 0; q < 3; ++q) {
      const int64_t t = scatter_index(f[3 * k + q], nv);   // scatter-add drops out-of-range vertices
      if (t >= 0) atomicAdd(&out[t], third);
    }
  }
}

// ------------------------------------------------------------------ dyn.py (FP64)
// Newmark-beta (dyn.py:5-51): scalar m, k, c or dense matrices (Gaussian elimination
// with partial pivoting for the matrix system, ndof <= 8).
__global__ void k_newmark(const double *a, const double *v, const double *d, const double *h, int n, const double *m,
                          const double *kk, const double *cc, int matrix, double dt, double gamma, double beta,
                          double *ao, double *vo, double *dout) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double c1 = gamma * dt, c2 = beta * dt * dt;
  double v1[8], v2[8], an[8];
  for (int i = 0; i < n; ++i) {
    v1[i] = v[i] + dt * (1 - gamma) * a[i];
    v2[i] = d[i] + dt * v[i] + dt * dt * (0.5 - beta) * a[i];
  }
  if (!matrix) {
    for (int i = 0; i < n; ++i) an[i] = (h[i] - cc[0] * v1[i] - kk[0] * v2[i]) / (m[0] + c1 * cc[0] + c2 * kk[0]);
  } else {
    double A[8][9];
    for (int i = 0; i < n; ++i) {
      double r = h[i];
      for (int j = 0; j < n; ++j) {
        A[i][j] = m[i * n + j] + c1 * cc[i * n + j] + c2 * kk[i * n + j];
        r -= cc[i * n + j] * v1[j] + kk[i * n + j] * v2[j];
      }
      A[i][n] = r;
    }
    for (int col = 0; col < n; ++col) {
      int piv = col;
      for (int r = col + 1; r < n; ++r)
        if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
      for (int j = 0; j <= n; ++j) {
        const double t = A[col][j];
        A[col][j] = A[piv][j];
        A[piv][j] = t;
      }
      for (int r = col + 1; r < n; ++r) {
        const double fct = A[r][col] / A[col][col];
        for (int j = col; j <= n; ++j) A[r][j] -= fct * A[col][j];
      }
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = A[i][n];
      for (int j = i + 1; j < n; ++j) s -= A[i][j] * an[j];
      an[i] = s / A[i][i];
    }
  }
  for (int i = 0; i < n; ++i) {
    ao[i] = an[i];
    vo[i] = c1 * an[i] + v1[i];
    dout[i] = c2 * an[i] + v2[i];
  }
}

// dyn.py:126-136
__global__ void k_force_sum(const float *__restrict__ h, int m, int c, double *out) {
  for (int comp = 0; comp < c; ++comp) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) acc += (double)h[(int64_t)i * c + comp];
    const double s = block_sum<kB>(acc);
    if (threadIdx.x == 0) out[comp] = s;
  }
}

// dyn.py:139-154
__global__ void k_torque(const float *__restrict__ x, const float *__restrict__ y, double xc, double yc, const double *d,
                         const float *__restrict__ h, int m, double *out) {
  const double cx = xc + d[0], cy = yc + d[1];
  double acc = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) acc += (x[i] - cx) * (double)h[2 * i + 1] - (y[i] - cy) * (double)h[2 * i];
  const double s = block_sum<kB>(acc);
  if (threadIdx.x == 0) *out = s;
}

// dyn.py:68-100
__global__ void k_coords(const float *__restrict__ x0, const float *__restrict__ y0, int m, double xc, double yc,
                         const double *d, int rotate, float *__restrict__ x, float *__restrict__ y) {
  GS_LOOP(i, (int64_t)m) {
    if (!rotate) {
      x[i] = (float)(x0[i] + d[0]);
      y[i] = (float)(y0[i] + d[1]);
    } else {
      const double xr = x0[i] - xc, yr = y0[i] - yc, cs = cos(d[2]), sn = sin(d[2]);
      x[i] = (float)(xc + d[0] + xr * cs - yr * sn);
      y[i] = (float)(yc + d[1] + xr * sn + yr * cs);
    }
  }
}

// dyn.py:103-120
__global__ void k_vel3(const float *__restrict__ x, const float *__restrict__ y, int m, double xc, double yc, const double *d,
                       const double *v, float *__restrict__ out) {
  GS_LOOP(i, (int64_t)m) {
    const double xr = x[i] - xc - d[0], yr = y[i] - yc - d[1];
    out[2 * i] = (float)(v[0] - v[2] * yr);
    out[2 * i + 1] = (float)(v[1] + v[2] * xr);
  }
}

}  // namespace vsx

using namespace vsx;

#define L1(kern, n, ...)                                                                                       \
  do {                                                                                                         \
    if ((n) > 0) hipLaunchKernelGGL(kern, dim3(grid_for((n), kB)), dim3(kB), 0, as_stream(stream), __VA_ARGS__); \
    VSX_CHECK_LAUNCH();                                                                                        \
  } while (0)

extern "C" int vsx_ib_kernel(int32_t kernel, const float *r, float *out, int64_t n, void *stream) {
  VSX_REQUIRE(kernel >= VSX_KERNEL_PESKIN_2PT && kernel <= VSX_KERNEL_COSINE_4PT, "unknown kernel %d", kernel);
  L1(k_delta, n, kernel, r, out, n);
  return VSX_OK;
}

extern "C" int vsx_ib_get_ib_stencil2d(const float *mx, const float *my, int32_t m, int32_t ny, int32_t kernel,
                                       int32_t radius, float *W, int32_t *I, void *stream) {
  VSX_REQUIRE(kernel >= VSX_KERNEL_PESKIN_2PT && kernel <= VSX_KERNEL_COSINE_4PT, "unknown kernel %d", kernel);
  VSX_REQUIRE(m >= 0, "negative marker count");
  switch (radius) {
    case 1: L1(k_stencil2d<2>, (int64_t)m, mx, my, m, ny, kernel, W, I); break;
    case 2: L1(k_stencil2d<4>, (int64_t)m, mx, my, m, ny, kernel, W, I); break;
    case 3: L1(k_stencil2d<6>, (int64_t)m, mx, my, m, ny, kernel, W, I); break;
    case 4: L1(k_stencil2d<8>, (int64_t)m, mx, my, m, ny, kernel, W, I); break;
    default: return fail(VSX_ERR_VALUE, "stencil_radius must be 1..4, got %d", radius);
  }
  return VSX_OK;
}

extern "C" int vsx_ib_get_ib_stencil3d(const float *xyz, int32_t m, int32_t nx, int32_t ny, int32_t nz, int32_t kernel,
                                       int32_t radius, float *W, int32_t *I, void *stream) {
  (void)nx;
  VSX_REQUIRE(kernel >= VSX_KERNEL_PESKIN_2PT && kernel <= VSX_KERNEL_COSINE_4PT, "unknown kernel %d", kernel);
  VSX_REQUIRE(m >= 0, "negative marker count");
  switch (radius) {
    case 1: L1(k_stencil3d<2>, (int64_t)m, xyz, m, ny, nz, kernel, W, I); break;
    case 2: L1(k_stencil3d<4>, (int64_t)m, xyz, m, ny, nz, kernel, W, I); break;
    case 3: L1(k_stencil3d<6>, (int64_t)m, xyz, m, ny, nz, kernel, W, I); break;
    case 4: L1(k_stencil3d<8>, (int64_t)m, xyz, m, ny, nz, kernel, W, I); break;
    default: return fail(VSX_ERR_VALUE, "stencil_radius must be 1..4, got %d", radius);
  }
  return VSX_OK;
}

extern "C" int vsx_ib_interpolate(const float *grid, int32_t c, int64_t ng, const float *W, const int32_t *I, int32_t m,
                                  int32_t s, float *out, void *stream) {
  VSX_REQUIRE(c > 0 && ng > 0 && m >= 0 && s > 0, "bad interpolate sizes");
  L1(k_interp<false>, (int64_t)m * c, grid, c, ng, W, I, m, s, 1.0f, out);
  return VSX_OK;
}

extern "C" int vsx_ib_spread(const float *mv, const float *grid_in, float *out, int32_t c, int64_t ng, const float *W,
                             const int32_t *I, int32_t m, int32_t s, void *stream) {
  VSX_REQUIRE(c > 0 && ng > 0 && m >= 0 && s > 0, "bad spread sizes");
  if (out != grid_in)
    VSX_HIP(hipMemcpyAsync(out, grid_in, sizeof(float) * c * ng, hipMemcpyDeviceToDevice, as_stream(stream)));
  L1(k_spread, (int64_t)m * s, mv, out, c, ng, W, I, m, s);
  return VSX_OK;
}

namespace vsx {

// ib/mdf.py:35-64 on caller buffers: work holds ub, F, dF ([m][c] each).
static int mdf_run(const float *grid_u, int32_t c, int64_t ng, const float *W, const int32_t *I, int32_t m, int32_t s,
                   const float *U, const float *ds_arr, float ds, int32_t n_iter, float *grid_force,
                   float *marker_force, float *work, void *stream) {
  const int64_t mc = (int64_t)m * c, ms = (int64_t)m * s;
  float *ub = work, *F = work + mc, *dF = work + 2 * mc;
  L1(k_fill, (int64_t)c * ng, grid_force, 0.f, (int64_t)c * ng);
  L1(k_fill, mc, F, 0.f, mc);
  L1(k_interp<false>, mc, grid_u, c, ng, W, I, m, s, 1.0f, ub);
  for (int it = 0; it < n_iter; ++it) {
    L1(k_mdf_update, mc, ub, U, ds_arr, ds, m, c, F, dF);
    L1(k_touch_zero, ms, grid_force, c, ng, I, ms);
    L1(k_spread, ms, dF, grid_force, c, ng, W, I, m, s);
    L1(k_interp<true>, mc, grid_force, c, ng, W, I, m, s, 0.5f, ub);
  }
  L1(k_touch_zero, ms, grid_force, c, ng, I, ms);
  L1(k_spread, ms, F, grid_force, c, ng, W, I, m, s);
  L1(k_neg, mc, F, marker_force, mc);
  return VSX_OK;
}

// Scratch of vsx_ib_multi_direct_forcing_noalloc: ub, F and dF for up to
// kMdfScratchMC marker components (markers x components) each, 768 KiB per
// device, held until the process ends.
constexpr int64_t kMdfScratchMC = 65536;
static __device__ __align__(256) float scratch[3 * kMdfScratchMC];

// Device address of `scratch` on the current device (each device holds its own
// copy), looked up once per device with that device current.  In host code the
// name `scratch` is only a handle for this lookup, never a device address.
static int scratch_ptr(float **out) {
  static std::mutex mu;
  static std::vector<float *> by_dev;
  int dev = 0;
  VSX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)by_dev.size() <= dev) by_dev.resize(dev + 1, nullptr);
  if (!by_dev[dev]) {
    void *p = nullptr;
    VSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    by_dev[dev] = static_cast<float *>(p);
  }
  *out = by_dev[dev];
  return VSX_OK;
}

}  // namespace vsx

extern "C" int vsx_ib_multi_direct_forcing(const float *grid_u, int32_t c, int64_t ng, const float *W, const int32_t *I,
                                           int32_t m, int32_t s, const float *U, const float *ds_arr, float ds,
                                           int32_t n_iter, float *grid_force, float *marker_force, float *work,
                                           void *stream) {
  VSX_REQUIRE(c > 0 && ng > 0 && m > 0 && s > 0 && n_iter >= 0, "bad multi_direct_forcing sizes");
  VSX_REQUIRE(work && grid_force && marker_force, "null buffer");
  return mdf_run(grid_u, c, ng, W, I, m, s, U, ds_arr, ds, n_iter, grid_force, marker_force, work, stream);
}

extern "C" int vsx_ib_multi_direct_forcing_noalloc(const float *grid_u, int32_t c, int64_t ng, const float *W,
                                                   const int32_t *I, int32_t m, int32_t s, const float *U,
                                                   const float *ds_arr, float ds, int32_t n_iter, float *grid_force,
                                                   float *marker_force, void *stream) {
  VSX_REQUIRE(c > 0 && ng > 0 && m > 0 && s > 0 && n_iter >= 0, "bad multi_direct_forcing sizes");
  VSX_REQUIRE((int64_t)m * c <= kMdfScratchMC, "multi_direct_forcing_noalloc: %lld marker components exceed the "
              "scratch limit of %lld", (long long)m * c, (long long)kMdfScratchMC);
  VSX_REQUIRE(grid_force && marker_force, "null buffer");
  float *work = nullptr;
  const int rc = scratch_ptr(&work);
  if (rc) return rc;
  return mdf_run(grid_u, c, ng, W, I, m, s, U, ds_arr, ds, n_iter, grid_force, marker_force, work, stream);
}

extern "C" int vsx_ib_get_area(const float *p, int32_t m, float *out, void *stream) {
  VSX_REQUIRE(m > 0, "no markers");
  hipLaunchKernelGGL(k_area, dim3(1), dim3(kB), 0, as_stream(stream), p, m, out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_ib_get_ds(const float *p, int32_t m, int32_t closed, float *out, void *stream) {
  VSX_REQUIRE(m > 0, "no markers");
  L1(k_ds, (int64_t)m, p, m, closed, out);
  return VSX_OK;
}

extern "C" int vsx_ib3d_get_triangle_areas(const float *v, int32_t nv, const int32_t *f, int32_t nf, float *out,
                                           void *stream) {
  VSX_REQUIRE(nv > 0 && nf >= 0, "bad mesh sizes");
  L1(k_tri_areas, (int64_t)nf, v, nv, f, nf, out);
  return VSX_OK;
}

extern "C" int vsx_ib3d_get_surface_area(const float *v, int32_t nv, const int32_t *f, int32_t nf, float *out,
                                         void *stream) {
  VSX_REQUIRE(nv > 0 && nf >= 0, "bad mesh sizes");
  hipLaunchKernelGGL(k_mesh_sum, dim3(1), dim3(kB), 0, as_stream(stream), v, nv, f, nf, 0, out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_ib3d_get_volume(const float *v, int32_t nv, const int32_t *f, int32_t nf, float *out, void *stream) {
  VSX_REQUIRE(nv > 0 && nf >= 0, "bad mesh sizes");
  hipLaunchKernelGGL(k_mesh_sum, dim3(1), dim3(kB), 0, as_stream(stream), v, nv, f, nf, 1, out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_ib3d_get_ds(const float *v, int32_t nv, const int32_t *f, int32_t nf, float *out, void *stream) {
  VSX_REQUIRE(nv > 0 && nf >= 0, "bad mesh sizes");
  L1(k_fill, (int64_t)nv, out, 0.f, (int64_t)nv);
  L1(k_vertex_ds, (int64_t)nf, v, nv, f, nf, out);
  return VSX_OK;
}

extern "C" int vsx_dyn_newmark(const double *a, const double *v, const double *d, const double *h, int32_t ndof,
                               const double *m, const double *k, const double *c, int32_t matrix, double dt,
                               double gamma, double beta, double *ao, double *vo, double *dout, void *stream) {
  VSX_REQUIRE(ndof >= 1 && ndof <= 8, "ndof must be 1..8, got %d", ndof);
  hipLaunchKernelGGL(k_newmark, dim3(1), dim3(64), 0, as_stream(stream), a, v, d, h, ndof, m, k, c, matrix, dt, gamma,
                     beta, ao, vo, dout);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_dyn_get_force_to_obj(const float *h, int32_t m, int32_t c, double *out, void *stream) {
  VSX_REQUIRE(m > 0 && c > 0, "bad sizes");
  hipLaunchKernelGGL(k_force_sum, dim3(1), dim3(kB), 0, as_stream(stream), h, m, c, out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_dyn_get_torque_to_obj(const float *x, const float *y, double xc, double yc, const double *d,
                                         const float *h, int32_t m, double *out, void *stream) {
  VSX_REQUIRE(m > 0, "no markers");
  hipLaunchKernelGGL(k_torque, dim3(1), dim3(kB), 0, as_stream(stream), x, y, xc, yc, d, h, m, out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_dyn_markers_coords(const float *x0, const float *y0, int32_t m, double xc, double yc, const double *d,
                                      int32_t rotate, float *x, float *y, void *stream) {
  L1(k_coords, (int64_t)m, x0, y0, m, xc, yc, d, rotate, x, y);
  return VSX_OK;
}

extern "C" int vsx_dyn_markers_velocity_3dof(const float *x, const float *y, int32_t m, double xc, double yc,
                                             const double *d, const double *v, float *out, void *stream) {
  L1(k_vel3, (int64_t)m, x, y, m, xc, yc, d, v, out);
  return VSX_OK;
}
