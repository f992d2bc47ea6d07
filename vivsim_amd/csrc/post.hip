// Field diagnostics of vivsim.post (velocity gradient, vorticity, strain,
// Q-criterion, energies) and the force coefficients of the VIV example, on
// device so that a run can monitor its flow without copying fields to the
// host; and the grid-refinement line transfers of vivsim.multigrid.
//
// Derivatives follow jnp.gradient with unit spacing and edge_order 1 (the
// reference's post.py:18-31): (u[k+1] - u[k-1]) * 0.5 inside, u[1] - u[0] and
// u[n-1] - u[n-2] at the two ends.  One thread per cell computes the full
// local gradient from its neighbours (reads only, contiguous along the last
// axis), then the requested quantity; nothing is materialised in between.
#include <cmath>

#include "common.hpp"

namespace vsx {

constexpr int kPB = 256;

// d field / d axis at index k of an axis of n >= 2 points; `at(j)` reads the
// field at index j of that axis.
template <class F>
__device__ __forceinline__ float grad1(F at, int k, int n) {
  if (k == 0) return at(1) - at(0);
  if (k == n - 1) return at(n - 1) - at(n - 2);
  return (at(k + 1) - at(k - 1)) * 0.5f;
}

// G[i][j] = du_i / dx_j at cell (x, y, z) of a [D][nx][ny][nz] field (nz = 1 in 2D)
template <int D>
__device__ __forceinline__ void local_gradient(const float *__restrict__ u, int nx, int ny, int nz, int x, int y, int z,
                                               float (&G)[3][3]) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int dims[3] = {nx, ny, nz};
  const int co[3] = {x, y, z};
  const int64_t stride[3] = {(int64_t)ny * nz, nz, 1};
  const int64_t cell = x * stride[0] + y * stride[1] + z;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[i][j] = 0.f;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const float *ui = u + i * n;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int64_t base = cell - (int64_t)co[j] * stride[j];
      G[i][j] = grad1([&](int k) { return ui[base + (int64_t)k * stride[j]]; }, co[j], dims[j]);
    }
  }
}

// quantities of one cell (VSX_POST_*); out has `comps` planes of n cells
template <int D, int OP>
__global__ __launch_bounds__(kPB) void k_post(const float *__restrict__ u, int nx, int ny, int nz,
                                              float *__restrict__ out) {
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(c % nz), y = (int)((c / nz) % ny), x = (int)(c / ((int64_t)ny * nz));
    if constexpr (OP == VSX_POST_VELOCITY_MAGNITUDE || OP == VSX_POST_KINETIC_ENERGY) {
      float s = 0.f;
#pragma unroll
      for (int i = 0; i < D; ++i) s += u[i * n + c] * u[i * n + c];
      out[c] = OP == VSX_POST_KINETIC_ENERGY ? 0.5f * s : sqrtf(s);
    } else {
      float G[3][3];
      local_gradient<D>(u, nx, ny, nz, x, y, z, G);
      if constexpr (OP == VSX_POST_GRADIENT) {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j) out[(i * D + j) * n + c] = G[i][j];
      } else if constexpr (OP == VSX_POST_STRAIN_RATE) {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j) out[(i * D + j) * n + c] = 0.5f * (G[i][j] + G[j][i]);
      } else if constexpr (OP == VSX_POST_STRAIN_RATE_MAGNITUDE) {
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j) {
            const float e = 0.5f * (G[i][j] + G[j][i]);
            s += e * e;
          }
        out[c] = sqrtf(s);
      } else if constexpr (OP == VSX_POST_DIVERGENCE) {
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < D; ++i) s += G[i][i];
        out[c] = s;
      } else if constexpr (OP == VSX_POST_Q_CRITERION) {
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j) s += G[i][j] * G[j][i];
        out[c] = -0.5f * s;
      } else {   // vorticity, its magnitude, enstrophy
        float w[3];
        if constexpr (D == 2) {
          w[0] = G[1][0] - G[0][1];
        } else {
          w[0] = G[2][1] - G[1][2];
          w[1] = G[0][2] - G[2][0];
          w[2] = G[1][0] - G[0][1];
        }
        constexpr int NW = D == 2 ? 1 : 3;
        if constexpr (OP == VSX_POST_VORTICITY) {
#pragma unroll
          for (int k = 0; k < NW; ++k) out[k * n + c] = w[k];
        } else {
          float s = 0.f;
#pragma unroll
          for (int k = 0; k < NW; ++k) s += w[k] * w[k];
          if constexpr (OP == VSX_POST_VORTICITY_MAGNITUDE) out[c] = D == 2 ? fabsf(w[0]) : sqrtf(s);
          else out[c] = 0.5f * s;   // enstrophy
        }
      }
    }
  }
}

// Mean of the kinetic energy or enstrophy field: FP64 partial sums of the
// float32 per-cell values per block (fixed grid, fixed order), then one block
// sums the partials in order: deterministic.
constexpr int kMeanBlocks = 1024;

template <int D, int OP>
__global__ __launch_bounds__(kPB) void k_mean_partial(const float *__restrict__ u, int nx, int ny, int nz,
                                                      double *__restrict__ partial) {
  const int64_t n = (int64_t)nx * ny * nz;
  double acc = 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(c % nz), y = (int)((c / nz) % ny), x = (int)(c / ((int64_t)ny * nz));
    float v;
    if constexpr (OP == VSX_POST_KINETIC_ENERGY) {
      float s = 0.f;
#pragma unroll
      for (int i = 0; i < D; ++i) s += u[i * n + c] * u[i * n + c];
      v = 0.5f * s;
    } else {
      float G[3][3];
      local_gradient<D>(u, nx, ny, nz, x, y, z, G);
      float s;
      if constexpr (D == 2) {
        const float w = G[1][0] - G[0][1];
        s = w * w;
      } else {
        const float a = G[2][1] - G[1][2], b = G[0][2] - G[2][0], d = G[1][0] - G[0][1];
        s = a * a + b * b + d * d;
      }
      v = 0.5f * s;
    }
    acc += (double)v;
  }
  __shared__ double red[kPB];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kPB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kPB) void k_mean_final(const double *__restrict__ partial, int nb, double inv_n,
                                                    float *__restrict__ out) {
  __shared__ double red[kPB];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += kPB) acc += partial[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kPB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = (float)(red[0] * inv_n);
}

__global__ void k_pressure(const float *__restrict__ rho, int64_t n, float cs2, float *__restrict__ out) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    out[c] = rho[c] * cs2;
}

// Cd, Cl = 2 h / (D U0^2) of an FP64 [n][2] force history
__global__ void k_force_coeff(const double *__restrict__ h, int n, double scale, double *__restrict__ cd,
                              double *__restrict__ cl) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    cd[t] = h[2 * t] * scale;
    cl[t] = h[2 * t + 1] * scale;
  }
}

// ------------------------------------------------------------------ grid refinement (multigrid.py)
// Populations crossing an interface of a block pair whose grids differ by a
// factor 2 (lbm/lattice.py:58-61): right +x {1, 5, 8}, left -x {3, 7, 6},
// up +y {2, 5, 6}, down -y {4, 7, 8}.
__constant__ int kMgDirs[4][3] = {{3, 7, 6}, {1, 5, 8}, {2, 5, 6}, {4, 7, 8}};

// coarse buffer line := mean of the 2 x 2 fine cells next to the interface
// (multigrid.py:58-98); one thread per (direction, coarse node)
__global__ void k_fine_to_coarse(const float *__restrict__ ff, int nxf, int nyf, float *__restrict__ fc, int nxc,
                                 int nyc, int dir) {
  const int nodes = dir < 2 ? nyc : nxc;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nodes) return;
  const int q = kMgDirs[dir][t / nodes], j = t % nodes;
  const float *p = ff + (int64_t)q * nxf * nyf;
  auto F = [&](int x, int y) { return p[(int64_t)x * nyf + y]; };
  float v;
  int64_t dst;
  if (dir == 0) {          // left: fine columns 0, 1 -> coarse column nxc - 1
    v = F(0, 2 * j) + F(0, 2 * j + 1) + F(1, 2 * j) + F(1, 2 * j + 1);
    dst = (int64_t)(nxc - 1) * nyc + j;
  } else if (dir == 1) {   // right: fine columns nxf - 1, nxf - 2 -> coarse column 0
    v = F(nxf - 1, 2 * j) + F(nxf - 1, 2 * j + 1) + F(nxf - 2, 2 * j) + F(nxf - 2, 2 * j + 1);
    dst = j;
  } else if (dir == 2) {   // up: fine rows nyf - 1, nyf - 2 -> coarse row 0
    v = F(2 * j, nyf - 1) + F(2 * j + 1, nyf - 1) + F(2 * j, nyf - 2) + F(2 * j + 1, nyf - 2);
    dst = (int64_t)j * nyc;
  } else {                 // down: fine rows 0, 1 -> coarse row nyc - 1
    v = F(2 * j, 0) + F(2 * j + 1, 0) + F(2 * j, 1) + F(2 * j + 1, 1);
    dst = (int64_t)j * nyc + nyc - 1;
  }
  fc[(int64_t)q * nxc * nyc + dst] = 0.25f * v;
}

// fine buffer line := each coarse interface value on the two fine cells it
// covers (multigrid.py:100-130); one thread per (direction, fine node)
__global__ void k_coarse_to_fine(const float *__restrict__ fc, int nxc, int nyc, float *__restrict__ ff, int nxf,
                                 int nyf, int dir) {
  const int nodes = dir < 2 ? nyf : nxf;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nodes) return;
  const int q = kMgDirs[dir][t / nodes], j = t % nodes, jc = j / 2;
  const float *p = fc + (int64_t)q * nxc * nyc;
  float *o = ff + (int64_t)q * nxf * nyf;
  if (dir == 0) o[(int64_t)(nxf - 1) * nyf + j] = p[jc];                          // left
  else if (dir == 1) o[j] = p[(int64_t)(nxc - 1) * nyc + jc];                      // right
  else if (dir == 2) o[(int64_t)j * nyf] = p[(int64_t)jc * nyc + nyc - 1];         // up
  else o[(int64_t)j * nyf + nyf - 1] = p[(int64_t)jc * nyc];                       // down
}

}  // namespace vsx

using namespace vsx;

namespace {
unsigned blocks_for(int64_t n) { return grid_for(n, kPB); }

template <int D>
int post_dispatch(int op, const float *u, int nx, int ny, int nz, float *out, hipStream_t s) {
  const int64_t n = (int64_t)nx * ny * nz;
#define VSX_POST_CASE(OP)                                                                                    \
  case OP:                                                                                                   \
    hipLaunchKernelGGL((k_post<D, OP>), dim3(blocks_for(n)), dim3(kPB), 0, s, u, nx, ny, nz, out);         \
    break;
  switch (op) {
    VSX_POST_CASE(VSX_POST_GRADIENT)
    VSX_POST_CASE(VSX_POST_VORTICITY)
    VSX_POST_CASE(VSX_POST_VORTICITY_MAGNITUDE)
    VSX_POST_CASE(VSX_POST_DIVERGENCE)
    VSX_POST_CASE(VSX_POST_STRAIN_RATE)
    VSX_POST_CASE(VSX_POST_STRAIN_RATE_MAGNITUDE)
    VSX_POST_CASE(VSX_POST_Q_CRITERION)
    VSX_POST_CASE(VSX_POST_ENSTROPHY)
    VSX_POST_CASE(VSX_POST_VELOCITY_MAGNITUDE)
    VSX_POST_CASE(VSX_POST_KINETIC_ENERGY)
    default: return fail(VSX_ERR_VALUE, "unknown diagnostic %d", op);
  }
#undef VSX_POST_CASE
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

int check_field(int dim, int nx, int ny, int nz, bool needs_gradient) {
  VSX_REQUIRE(dim == 2 || dim == 3, "velocity field must have 2 or 3 components, got %d", dim);
  VSX_REQUIRE(nx > 0 && ny > 0 && nz > 0 && (dim == 3 || nz == 1), "bad field shape %d x %d x %d", nx, ny, nz);
  if (needs_gradient)   // np.gradient / jnp.gradient need two points along every axis
    VSX_REQUIRE(nx >= 2 && ny >= 2 && (dim == 2 || nz >= 2),
                "each spatial axis needs at least 2 points for a gradient (got %d x %d x %d)", nx, ny, nz);
  return VSX_OK;
}
}  // namespace

extern "C" int vsx_post_field(const float *u, int32_t dim, int32_t nx, int32_t ny, int32_t nz, int32_t op, float *out,
                              void *stream) {
  const bool grad = op != VSX_POST_VELOCITY_MAGNITUDE && op != VSX_POST_KINETIC_ENERGY;
  int rc = check_field(dim, nx, ny, nz, grad);
  if (rc) return rc;
  VSX_REQUIRE(u && out, "null buffer");
  return dim == 2 ? post_dispatch<2>(op, u, nx, ny, 1, out, as_stream(stream))
                  : post_dispatch<3>(op, u, nx, ny, nz, out, as_stream(stream));
}

extern "C" int vsx_post_mean(const float *u, int32_t dim, int32_t nx, int32_t ny, int32_t nz, int32_t op,
                             double *partial, float *out, void *stream) {
  VSX_REQUIRE(op == VSX_POST_KINETIC_ENERGY || op == VSX_POST_ENSTROPHY, "mean of op %d is not provided", op);
  int rc = check_field(dim, nx, ny, nz, op == VSX_POST_ENSTROPHY);
  if (rc) return rc;
  VSX_REQUIRE(u && partial && out, "null buffer");
  hipStream_t s = as_stream(stream);
  const int64_t n = (int64_t)nx * ny * nz;
  if (dim == 2) {
    if (op == VSX_POST_KINETIC_ENERGY)
      hipLaunchKernelGGL((k_mean_partial<2, VSX_POST_KINETIC_ENERGY>), dim3(kMeanBlocks), dim3(kPB), 0, s, u, nx, ny, 1, partial);
    else
      hipLaunchKernelGGL((k_mean_partial<2, VSX_POST_ENSTROPHY>), dim3(kMeanBlocks), dim3(kPB), 0, s, u, nx, ny, 1, partial);
  } else {
    if (op == VSX_POST_KINETIC_ENERGY)
      hipLaunchKernelGGL((k_mean_partial<3, VSX_POST_KINETIC_ENERGY>), dim3(kMeanBlocks), dim3(kPB), 0, s, u, nx, ny, nz, partial);
    else
      hipLaunchKernelGGL((k_mean_partial<3, VSX_POST_ENSTROPHY>), dim3(kMeanBlocks), dim3(kPB), 0, s, u, nx, ny, nz, partial);
  }
  hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kPB), 0, s, partial, kMeanBlocks, 1.0 / (double)n, out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_post_pressure(const float *rho, int64_t n, float cs2, float *out, void *stream) {
  VSX_REQUIRE(n >= 0 && (n == 0 || (rho && out)), "bad pressure arguments");
  if (n > 0) hipLaunchKernelGGL(k_pressure, dim3(blocks_for(n)), dim3(kPB), 0, as_stream(stream), rho, n, cs2, out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_force_coefficients(const double *h, int32_t n, double D, double U0, double *cd, double *cl,
                                      void *stream) {
  VSX_REQUIRE(n >= 0 && (n == 0 || (h && cd && cl)), "bad force-coefficient arguments");
  VSX_REQUIRE(D * U0 * U0 != 0.0, "D * U0^2 must be non-zero");
  if (n > 0)
    hipLaunchKernelGGL(k_force_coeff, dim3(blocks_for(n)), dim3(kPB), 0, as_stream(stream), h, n, 2.0 / (D * U0 * U0),
                       cd, cl);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_mg_fine_to_coarse(const float *f_fine, int32_t nxf, int32_t nyf, float *f_coarse, int32_t nxc,
                                     int32_t nyc, int32_t dir, void *stream) {
  VSX_REQUIRE(dir >= 0 && dir <= 3, "unknown transfer direction %d", dir);
  VSX_REQUIRE(f_fine && f_coarse && nxf >= 2 && nyf >= 2 && nxc >= 1 && nyc >= 1, "bad fine_to_coarse arguments");
  if (dir < 2) VSX_REQUIRE(nyf == 2 * nyc, "fine ny %d must be twice the coarse ny %d", nyf, nyc);
  else VSX_REQUIRE(nxf == 2 * nxc, "fine nx %d must be twice the coarse nx %d", nxf, nxc);
  const int nodes = dir < 2 ? nyc : nxc;
  hipLaunchKernelGGL(k_fine_to_coarse, dim3(blocks_for(3 * (int64_t)nodes)), dim3(kPB), 0, as_stream(stream), f_fine,
                     nxf, nyf, f_coarse, nxc, nyc, dir);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

extern "C" int vsx_mg_coarse_to_fine(const float *f_coarse, int32_t nxc, int32_t nyc, float *f_fine, int32_t nxf,
                                     int32_t nyf, int32_t dir, void *stream) {
  VSX_REQUIRE(dir >= 0 && dir <= 3, "unknown transfer direction %d", dir);
  VSX_REQUIRE(f_fine && f_coarse && nxf >= 1 && nyf >= 1 && nxc >= 1 && nyc >= 1, "bad coarse_to_fine arguments");
  if (dir < 2) VSX_REQUIRE(nyf == 2 * nyc, "fine ny %d must be twice the coarse ny %d", nyf, nyc);
  else VSX_REQUIRE(nxf == 2 * nxc, "fine nx %d must be twice the coarse nx %d", nxf, nxc);
  const int nodes = dir < 2 ? nyf : nxf;
  hipLaunchKernelGGL(k_coarse_to_fine, dim3(blocks_for(3 * (int64_t)nodes)), dim3(kPB), 0, as_stream(stream), f_coarse,
                     nxc, nyc, f_fine, nxf, nyf, dir);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}
