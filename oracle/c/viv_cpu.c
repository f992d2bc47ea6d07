/*
 * CPU restatement of the reference's D2Q9 time steps in C with OpenMP, for
 * the CPU baseline that is timed beside the GPU and for host-side parity
 * checks.  Test infrastructure: only the tests and the benchmark's CPU leg
 * load it.
 *
 * It follows the same reference functions as the NumPy oracle (oracle/*.py),
 * evaluated cell by cell in float32 (the reference's default precision),
 * with the body-force reduction and the Newmark update in float64:
 *   macroscopic / equilibrium     vivsim/lbm/basic.py:88-135
 *   BGK                            vivsim/lbm/basic.py:138-156
 *   MRT  f + op (feq - f)          vivsim/lbm/collision/mrt.py:66-88 (op passed in, 9 x 9 float32)
 *   Guo term, Guo-MRT forcing      vivsim/lbm/forcing/guo.py:6-33, 82-107 (opF passed in)
 *   streaming (periodic pull)      vivsim/lbm/basic.py:60-85
 *   NEE, equilibrium, pressure     vivsim/lbm/boundary/nee.py:19-60, eq.py:25-56,
 *                                  _helpers.py:98-132 (wrap_pressure)
 *   Peskin 4-point kernel          vivsim/ib/kernels.py:25-43
 *   stencil / interpolate / spread vivsim/ib/stencil.py:7-110 (JAX index rules)
 *   multi-direct forcing           vivsim/ib/mdf.py:10-64
 *   Newmark-beta                   vivsim/dyn.py:5-51
 *   step compositions              examples/2d/lid_driven_cavity.py:55-64 (cavity),
 *                                  examples/2d/vortex_induced_vibration.py:96-148 with the
 *                                  MRT + Guo recipe of examples/2d/poiseuille_channel.py:113-128
 *
 * Layout: f[9][nx][ny] float32, C order (y contiguous).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const float W[9] = {4.f / 9, 1.f / 9, 1.f / 9, 1.f / 9, 1.f / 9, 1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36};

#define F(f, i, x, y) (f)[((size_t)(i) * nx + (x)) * ny + (y)]

static inline void macro(const float *p, float *rho, float *ux, float *uy) {
  float r = 0.f;
  for (int i = 0; i < 9; ++i) r += p[i];
  *rho = r;
  *ux = (p[1] + p[5] + p[8] - p[3] - p[6] - p[7]) / r;
  *uy = (p[2] + p[5] + p[6] - p[4] - p[7] - p[8]) / r;
}

static inline void feq(float rho, float ux, float uy, float *out) {
  const float u2 = ux * ux + uy * uy;
  for (int i = 0; i < 9; ++i) {
    const float cu = CX[i] * ux + CY[i] * uy;
    out[i] = rho * W[i] * (1.f + 3.f * cu + 4.5f * cu * cu - 1.5f * u2);
  }
}

static inline void cell_get(const float *f, int nx, int ny, int x, int y, float *p) {
  for (int i = 0; i < 9; ++i) p[i] = F(f, i, x, y);
}
static inline void cell_set(float *f, int nx, int ny, int x, int y, const float *p) {
  for (int i = 0; i < 9; ++i) F(f, i, x, y) = p[i];
}

/* periodic pull streaming: out_i(x, y) = in_i(x - cx_i, y - cy_i) */
static void streaming(const float *in, float *out, int nx, int ny) {
#pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; ++x)
    for (int i = 0; i < 9; ++i) {
      const int xs = (x - CX[i] + nx) % nx;
      const float *src = in + ((size_t)i * nx + xs) * ny;
      float *dst = out + ((size_t)i * nx + x) * ny;
      for (int y = 0; y < ny; ++y) dst[y] = src[(y - CY[i] + ny) % ny];
    }
}

/* wall line geometry: loc 0 left, 1 right, 2 bottom, 3 top */
static void wall_cell(int loc, int j, int nx, int ny, int *x, int *y, int *xn, int *yn) {
  if (loc == 0) { *x = 0; *y = j; *xn = 1; *yn = j; }
  else if (loc == 1) { *x = nx - 1; *y = j; *xn = nx - 2; *yn = j; }
  else if (loc == 2) { *x = j; *y = 0; *xn = j; *yn = 1; }
  else { *x = j; *y = ny - 1; *xn = j; *yn = ny - 2; }
}

/* NEE on a whole wall line: wall := feq(rho_w, u_w) + (f_nb - feq(macro(f_nb))); the
 * neighbour line is read before any wall cell is written (they are different lines).
 * pressure != 0: the pressure wrapper, u_n from rho_w and the wall line's
 * sum_zero + 2 sum_out, the tangential velocity of the neighbour. */
static void wall_nee(float *f, int nx, int ny, int loc, float rho_w, float uxw, float uyw, int pressure) {
  const int n = loc < 2 ? ny : nx;
  const int axis = loc < 2 ? 0 : 1, sign = (loc == 0 || loc == 2) ? 1 : -1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) {
    int x, y, xn, yn;
    wall_cell(loc, j, nx, ny, &x, &y, &xn, &yn);
    float pw[9], pn[9], e1[9], e2[9], rn, unx, uny;
    cell_get(f, nx, ny, x, y, pw);
    cell_get(f, nx, ny, xn, yn, pn);
    macro(pn, &rn, &unx, &uny);
    float ux = uxw, uy = uyw;
    if (pressure) {
      float z = 0.f, o = 0.f;
      for (int i = 0; i < 9; ++i) {
        const int cn = (axis == 0 ? CX[i] : CY[i]) * sign;
        if (cn == 0) z += pw[i];
        if (cn < 0) o += pw[i];
      }
      const float uns = 1.f - (z + 2.f * o) / rho_w;
      ux = unx;
      uy = uny;
      if (axis == 0) ux = sign * uns;
      else uy = sign * uns;
    }
    feq(rho_w, ux, uy, e1);
    feq(rn, unx, uny, e2);
    for (int i = 0; i < 9; ++i) pw[i] = e1[i] + (pn[i] - e2[i]);
    cell_set(f, nx, ny, x, y, pw);
  }
}

static void wall_eq(float *f, int nx, int ny, int loc, float rho_w, float uxw, float uyw) {
  const int n = loc < 2 ? ny : nx;
  float e[9];
  feq(rho_w, uxw, uyw, e);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) {
    int x, y, xn, yn;
    wall_cell(loc, j, nx, ny, &x, &y, &xn, &yn);
    cell_set(f, nx, ny, x, y, e);
  }
}

/* ------------------------------------------------------------------ cavity */
/* n_steps of the lid-driven cavity: macroscopic -> equilibrium -> BGK ->
 * streaming -> NEE on left, right, bottom, top (the lid, u = u_lid, on top).
 * f is updated in place; tmp holds 9 nx ny floats. */
int vc_cavity_run(float *f, float *tmp, int nx, int ny, float omega, float u_lid, int n_steps) {
  if (nx < 3 || ny < 3 || n_steps < 0) return 1;
  for (int t = 0; t < n_steps; ++t) {
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        float p[9], e[9], rho, ux, uy;
        cell_get(f, nx, ny, x, y, p);
        macro(p, &rho, &ux, &uy);
        feq(rho, ux, uy, e);
        for (int i = 0; i < 9; ++i) p[i] = (1.f - omega) * p[i] + omega * e[i];
        cell_set(tmp, nx, ny, x, y, p);
      }
    streaming(tmp, f, nx, ny);
    wall_nee(f, nx, ny, 0, 1.f, 0.f, 0.f, 0);
    wall_nee(f, nx, ny, 1, 1.f, 0.f, 0.f, 0);
    wall_nee(f, nx, ny, 2, 1.f, 0.f, 0.f, 0);
    wall_nee(f, nx, ny, 3, 1.f, u_lid, 0.f, 0);
  }
  return 0;
}

/* ------------------------------------------------------------------ VIV */
static inline float phi4(float r) {
  const float a = fabsf(r);
  if (a > 2.f) return 0.f;
  if (a < 1.f) return (3.f - 2.f * a + sqrtf(1.f + 4.f * a - 4.f * a * a)) * 0.125f;
  return (5.f - 2.f * a - sqrtf(-7.f + 12.f * a - 4.f * a * a)) * 0.125f;
}

typedef struct VcViv {
  int nx, ny, n_markers, win, n_iter, dof_mask;
  float u_in, ds;
  double win_x0, win_y0, mass, stiffness, damping, area;
  const float *op, *opf;         /* 9 x 9 float32 MRT collision and forcing operators */
  const double *marker_x, *marker_y;
} VcViv;

/* gather rule for flat window indices: wrap a negative index once, clamp */
static inline int gather_idx(int i, int n) {
  if (i < 0) i += n;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
static inline int scatter_idx(int i, int n) {
  if (i < 0) i += n;
  return (i < 0 || i >= n) ? -1 : i;
}

/* n_steps of the VIV IB-LBM step (MRT + Guo, equilibrium inlet on the left,
 * pressure NEE outlet on the right, periodic in y).  body = [a0 a1 v0 v1 d0 d1]
 * (float64) is advanced in place; hist_d / hist_h ([n][2], may be NULL) get the
 * displacement and the body force of every step.  scratch: 9 nx ny + 2 nx ny
 * floats. */
int vc_viv_run(const VcViv *c, float *f, float *scratch, double *body, int n_steps, double *hist_d, double *hist_h) {
  const int nx = c->nx, ny = c->ny, S = c->win, SS = S * S, M = c->n_markers;
  if (nx < 3 || ny < 3 || S < 4 || S > nx || S > ny || M <= 0 || n_steps < 0) return 1;
  float *post = scratch, *g = scratch + (size_t)9 * nx * ny;
  float *uw = (float *)malloc(sizeof(float) * 2 * SS), *gw = (float *)malloc(sizeof(float) * 2 * SS);
  float *wgt = (float *)malloc(sizeof(float) * 16 * M), *ub = (float *)malloc(sizeof(float) * 4 * M);
  int *idx = (int *)malloc(sizeof(int) * 16 * M);
  if (!uw || !gw || !wgt || !ub || !idx) return 2;
  float *ub0 = ub, *Ft = ub + 2 * M;
  memset(g, 0, sizeof(float) * 2 * nx * ny);
  for (int t = 0; t < n_steps; ++t) {
    /* window origin int32(IB_X0 + d), clamped as lax.dynamic_slice clamps it */
    const int x0 = (int)trunc(c->win_x0 + body[4]), y0 = (int)trunc(c->win_y0 + body[5]);
    const int xc = x0 < 0 ? 0 : (x0 > nx - S ? nx - S : x0), yc = y0 < 0 ? 0 : (y0 > ny - S ? ny - S : y0);
    /* window velocity of the post-streaming state */
    for (int k = 0; k < SS; ++k) {
      float p[9], rho, ux, uy;
      cell_get(f, nx, ny, xc + k / S, yc + k % S, p);
      macro(p, &rho, &ux, &uy);
      uw[k] = ux;
      uw[SS + k] = uy;
    }
    /* stencils (ib/stencil.py:7-51) at marker coordinates relative to the unclamped origin */
    for (int m = 0; m < M; ++m) {
      const float rx = (float)(c->marker_x[m] + body[4] - x0), ry = (float)(c->marker_y[m] + body[5] - y0);
      const int bx = (int)floorf(rx), by = (int)floorf(ry);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          const int sx = bx - 1 + a, sy = by - 1 + b;
          wgt[16 * m + 4 * a + b] = phi4((float)sx - rx) * phi4((float)sy - ry);
          idx[16 * m + 4 * a + b] = sx * S + sy;
        }
    }
    /* multi-direct forcing (ib/mdf.py:35-64) */
    for (int m = 0; m < M; ++m) {
      float sx = 0.f, sy = 0.f;
      for (int s = 0; s < 16; ++s) {
        const int k = gather_idx(idx[16 * m + s], SS);
        sx += wgt[16 * m + s] * uw[k];
        sy += wgt[16 * m + s] * uw[SS + k];
      }
      ub0[2 * m] = sx;
      ub0[2 * m + 1] = sy;
      Ft[2 * m] = Ft[2 * m + 1] = 0.f;
    }
    const float Ux = (float)body[2], Uy = (float)body[3];
    for (int it = 0; it < c->n_iter; ++it) {
      memset(gw, 0, sizeof(float) * 2 * SS);
      for (int m = 0; m < M; ++m) {
        const float dfx = (Ux - ub0[2 * m]) * (c->ds * 2.f), dfy = (Uy - ub0[2 * m + 1]) * (c->ds * 2.f);
        Ft[2 * m] += dfx;
        Ft[2 * m + 1] += dfy;
        for (int s = 0; s < 16; ++s) {
          const int k = scatter_idx(idx[16 * m + s], SS);
          if (k >= 0) {
            gw[k] += dfx * wgt[16 * m + s];
            gw[SS + k] += dfy * wgt[16 * m + s];
          }
        }
      }
      for (int m = 0; m < M; ++m) {
        float sx = 0.f, sy = 0.f;
        for (int s = 0; s < 16; ++s) {
          const int k = gather_idx(idx[16 * m + s], SS);
          sx += wgt[16 * m + s] * (gw[k] * 0.5f);
          sy += wgt[16 * m + s] * (gw[SS + k] * 0.5f);
        }
        ub0[2 * m] += sx;
        ub0[2 * m + 1] += sy;
      }
    }
    /* force window = spread(F); body force h = sum(-F) + a * area (float64) */
    memset(gw, 0, sizeof(float) * 2 * SS);
    double h[2] = {0.0, 0.0};
    for (int m = 0; m < M; ++m) {
      for (int s = 0; s < 16; ++s) {
        const int k = scatter_idx(idx[16 * m + s], SS);
        if (k >= 0) {
          gw[k] += Ft[2 * m] * wgt[16 * m + s];
          gw[SS + k] += Ft[2 * m + 1] * wgt[16 * m + s];
        }
      }
      h[0] -= (double)Ft[2 * m];
      h[1] -= (double)Ft[2 * m + 1];
    }
    /* Newmark-beta (dyn.py:5-51), gamma 1/2, beta 1/4, dt 1 */
    for (int k = 0; k < 2; ++k) {
      const double a = body[k], v = body[2 + k], d = body[4 + k];
      h[k] += a * c->area;
      const double v1 = v + 0.5 * a, v2 = d + v + 0.25 * a;
      const double an = (h[k] - c->damping * v1 - c->stiffness * v2) / (c->mass + 0.5 * c->damping + 0.25 * c->stiffness);
      const int on = (c->dof_mask >> k) & 1;
      body[k] = on ? an : 0.0;
      body[2 + k] = on ? 0.5 * an + v1 : 0.0;
      body[4 + k] = on ? 0.25 * an + v2 : 0.0;
    }
    if (hist_d) { hist_d[2 * t] = body[4]; hist_d[2 * t + 1] = body[5]; }
    if (hist_h) { hist_h[2 * t] = h[0]; hist_h[2 * t + 1] = h[1]; }
    for (int k = 0; k < SS; ++k) {
      g[(size_t)(xc + k / S) * ny + yc + k % S] = gw[k];
      g[(size_t)nx * ny + (size_t)(xc + k / S) * ny + yc + k % S] = gw[SS + k];
    }
    /* collision + Guo forcing: u* = u + g / (2 rho), f += op (feq(u*) - f), f += opF F(g, u*) */
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        float p[9], e[9], q[9], Fi[9], rho, ux, uy;
        cell_get(f, nx, ny, x, y, p);
        macro(p, &rho, &ux, &uy);
        const float gx = g[(size_t)x * ny + y], gy = g[(size_t)nx * ny + (size_t)x * ny + y];
        ux = ux + gx * 0.5f / rho;
        uy = uy + gy * 0.5f / rho;
        feq(rho, ux, uy, e);
        for (int i = 0; i < 9; ++i) e[i] -= p[i];
        for (int i = 0; i < 9; ++i) {
          float s = 0.f;
          for (int j = 0; j < 9; ++j) s += c->op[9 * i + j] * e[j];
          q[i] = p[i] + s;
        }
        for (int i = 0; i < 9; ++i) {
          const float cu = CX[i] * ux + CY[i] * uy;
          Fi[i] = W[i] * (gx * (3.f * (CX[i] - ux) + 9.f * cu * CX[i]) + gy * (3.f * (CY[i] - uy) + 9.f * cu * CY[i]));
        }
        for (int i = 0; i < 9; ++i) {
          float s = 0.f;
          for (int j = 0; j < 9; ++j) s += c->opf[9 * i + j] * Fi[j];
          post[((size_t)i * nx + x) * ny + y] = q[i] + s;
        }
      }
    for (int k = 0; k < SS; ++k) {
      g[(size_t)(xc + k / S) * ny + yc + k % S] = 0.f;
      g[(size_t)nx * ny + (size_t)(xc + k / S) * ny + yc + k % S] = 0.f;
    }
    streaming(post, f, nx, ny);
    wall_eq(f, nx, ny, 0, 1.f, c->u_in, 0.f);
    wall_nee(f, nx, ny, 1, 1.f, 0.f, 0.f, 1);
  }
  free(uw);
  free(gw);
  free(wgt);
  free(ub);
  free(idx);
  return 0;
}

int vc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
