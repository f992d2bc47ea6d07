// Hand-written reverse mode of the 2D VIV step (MRT + Guo, equilibrium inlet,
// pressure-NEE outlet, IB multi-direct forcing, FP64 Newmark) with sqrt(T)
// checkpointing.  Objective J = (1/T) sum_{t=1..T} d_y(t)^2.
//
// Forward state of the reference: f_t (post-streaming, post-walls).  Step t:
//   P_t     = f_t + A (feq(rho, u*) - f_t) + B F(g_t, u*)       (A = M^-1 S M,
//             B = M^-1 (I - S/2) M, u* = u + g / (2 rho); forcing/guo.py:62-107)
//   s       = streaming(P_t)                                   (lbm/basic.py:44-85)
//   f_{t+1} = walls(s): left := feq(1, (u_in, 0)); right := feq(1, (u_n, u_t))
//             + s_nb - feq(macro(s_nb))                          (eq.py:25-56, nee.py:19-65,
//                                                                 _helpers.py:98-132)
//   g_t, H_t = MDF(u(f_t), W(x_m + d_t)), body_{t+1} = Newmark(body_t, H_t)
// Reverse step t, given fbar_{t+1} and the adjoint of body_{t+1}:
//   k_adj_walls  wall adjoint in place on fbar_{t+1} (-> sbar), inlet-velocity term
//                (inside k_adj_ib of step t + 1 when that step's window touches a
//                wall column)
//   k_adj_bulk   pbar_i(x) = sbar_i(x + c_i); per-cell transpose of the collision +
//                forcing Jacobian -> fbar_t, and gbar_t on the force window
//   k_adj_ib     Newmark adjoint, replay + reverse of the MDF loop (transposes of
//                interpolate / spread, kernel derivative phi' for the marker
//                positions), window-velocity adjoint added to fbar_t.  It runs
//                on its own stream beside k_adj_bulk of step t (replay first,
//                then a wait for the bulk kernel's output), see viv_adjoint.
#include <algorithm>
#include <vector>

#include "step_impl.hpp"
#include "viv2d.hpp"

namespace vsx {

int viv_run(const VsxLbm *p, const VsxIB2D *ib, const float *f_in, float *f_out, void *wa, void *wb, int n,
            void *stream);

constexpr int kMaxIter = 4;

__device__ __forceinline__ double block_sum_d(double v, double *red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) s += red[w];
  return s;
}
// Four block sums at once (red: 4 x 16 doubles); v holds the sums on return.
__device__ __forceinline__ void block_sum4_d(double (&v)[4], double *red) {
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k * 16 + (threadIdx.x >> 6)] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double s = 0.0;
    for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) s += red[k * 16 + w];
    v[k] = s;
  }
}

// feq_i(rho, u) = rho w_i (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 |u|^2)   (lbm/basic.py:114-135)
__device__ __forceinline__ void equilibrium_raw(float rho, float ux, float uy, float (&feq)[9]) {
  const float usq = ux * ux + uy * uy;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float cu = (float)D2Q9::cx[i] * ux + (float)D2Q9::cy[i] * uy;
    feq[i] = rho * D2Q9::w[i] * (1.f + 3.f * cu + 4.5f * cu * cu - 1.5f * usq);
  }
}

// d feq_i / d u_a at (rho, u): rho w_i (3 c_ia + 9 (c_i.u) c_ia - 3 u_a)
__device__ __forceinline__ float dfeq_du(int i, int a, float rho, float cu, float ua) {
  const float c = a == 0 ? (float)D2Q9::cx[i] : (float)D2Q9::cy[i];
  return rho * D2Q9::w[i] * (3.f * c + 9.f * cu * c - 3.f * ua);
}

// ------------------------------------------------------------------ walls
// Hand-off words of the reverse pass (VsxAdj2D.sync): [0] count of published
// bulk workgroups, [1] the count the IB kernels have consumed, [2] number of
// finished IB kernels, [3] error code (1: a bounded wait timed out).
constexpr long kAdjSpinLimit = 1l << 22;
__device__ __forceinline__ void adj_wait(int32_t *sync, int word, int target) {
  long spins = 0;
  while (__hip_atomic_load(&sync[word], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > kAdjSpinLimit) {
      __hip_atomic_store(&sync[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
  }
}

// Clamped window column of a body state (examples/2d/vortex_induced_vibration.py:104-105)
__device__ __forceinline__ int adj_win_col(double x0, const double *body, int nx, int S) {
  return min(max((int)trunc(x0 + body[4]), 0), nx - S);
}
// The window (columns xc .. xc + S - 1) touches none of the columns the wall
// adjoint writes (0, nx - 2, nx - 1): the walls of step t - 1 and the IB
// adjoint of step t then write disjoint cells and may run in either order.
__device__ __forceinline__ bool adj_walls_clear(int xc, int S, int nx) { return xc >= 1 && xc + S <= nx - 2; }

// Wall adjoint (rows y = threadIdx.x + k blockDim.x) in place on adj (= fbar_{t+1},
// becomes sbar, the adjoint of the streamed field before the walls); fn = f_{t+1}.
// AGENT: adj is read and written with agent-scope accesses (a hand-off between
// concurrent kernels).  Returns this thread's part of the inlet-velocity term.
template <bool AGENT>
__device__ __forceinline__ double adj_walls_rows(float *__restrict__ adj, const float *__restrict__ fn, int nx, int ny,
                                                 float u_in) {
  auto ld = [](const float *q) { return AGENT ? ld_agent(q) : *q; };
  auto st = [](float *q, float v) {
    if (AGENT) st_agent(q, v);
    else *q = v;
  };
  const size_t n = (size_t)nx * ny;
  double part = 0.0;
  for (int y = threadIdx.x; y < ny; y += blockDim.x) {
    // left, equilibrium at (1, (u_in, 0)): the wall cell forgets s; d/du_in
    {
      const size_t c0 = (size_t)y;
      float s = 0.f;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const float q = ld(&adj[i * n + c0]);
        s += q * dfeq_du(i, 0, 1.f, (float)D2Q9::cx[i] * u_in, u_in);
        st(&adj[i * n + c0], 0.f);
      }
      part += (double)s;
    }
    // right, pressure NEE with rho_w = 1
    {
      const size_t cw = (size_t)(nx - 1) * ny + y, cb = (size_t)(nx - 2) * ny + y;
      float q[9], fw[9], sb[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        q[i] = ld(&adj[i * n + cw]);
        fw[i] = fn[i * n + cw];
        sb[i] = fn[i * n + cb];
      }
      // wall velocity: (u_n, u_t) = momentum of the written wall cell (its density is 1)
      float un = 0.f, ut = 0.f, rb = 0.f, jbx = 0.f, jby = 0.f;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        un += (float)D2Q9::cx[i] * fw[i];
        ut += (float)D2Q9::cy[i] * fw[i];
        rb += sb[i];
        jbx += (float)D2Q9::cx[i] * sb[i];
        jby += (float)D2Q9::cy[i] * sb[i];
      }
      const float ubx = jbx / rb, uby = jby / rb;
      float gx = 0.f, gy = 0.f, rhobar = 0.f, uxbar = 0.f, uybar = 0.f;
      float feqb[9];
      equilibrium_raw(rb, ubx, uby, feqb);
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const float cuw = (float)D2Q9::cx[i] * un + (float)D2Q9::cy[i] * ut;
        gx += q[i] * dfeq_du(i, 0, 1.f, cuw, un);
        gy += q[i] * dfeq_du(i, 1, 1.f, cuw, ut);
        const float cub = (float)D2Q9::cx[i] * ubx + (float)D2Q9::cy[i] * uby;
        rhobar -= q[i] * feqb[i] / rb;
        uxbar -= q[i] * dfeq_du(i, 0, rb, cub, ubx);
        uybar -= q[i] * dfeq_du(i, 1, rb, cub, uby);
      }
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        // d u_n / d s_j = -sign / rho_w * coef_j (right wall: sign = -1): 1 for c_x = 0, 2 for c_x > 0
        const float coef = D2Q9::cx[j] == 0 ? 1.f : (D2Q9::cx[j] > 0 ? 2.f : 0.f);
        st(&adj[j * n + cw], gx * coef);
        const float cxj = (float)D2Q9::cx[j], cyj = (float)D2Q9::cy[j];
        const float extra = q[j] + gy * (cyj - uby) / rb + rhobar + (uxbar * (cxj - ubx) + uybar * (cyj - uby)) / rb;
        st(&adj[j * n + cb], ld(&adj[j * n + cb]) + extra);
      }
    }
  }
  return part;
}

// Wall adjoint of step t, one workgroup, on the bulk kernels' stream after the
// bulk kernel of step t + 1.  body1 (the body state of step t + 1, NULL for the
// last step of the run): when the IB window of step t + 1 touches a wall
// column, the walls run inside that step's IB kernel instead (after its
// window update), and this kernel does nothing.
__global__ __launch_bounds__(1024) void k_adj_walls(float *adj, const float *__restrict__ fn, int nx, int ny,
                                                    float u_in, double *u_in_part, const double *body1, double x0,
                                                    int S) {
  __shared__ double red[16];
  if (body1 && !adj_walls_clear(adj_win_col(x0, body1, nx, S), S, nx)) return;
  const double part = adj_walls_rows<false>(adj, fn, nx, ny, u_in);
  const double tot = block_sum_d(part, red);
  if (threadIdx.x == 0) *u_in_part = tot;
}

// ------------------------------------------------------------------ bulk
struct AdjBulk {
  int nx, ny, S;
  float rates[9];
  double win_x0, win_y0;
  int32_t *sync;   // hand-off words (adj_wait)
};

// Reverse bulk step t.  It runs while the IB kernel of step t + 1 may still
// be working on sbar: the workgroups whose pull reads a column that kernel
// writes (its window +- 1; the wall columns when it runs the walls) wait for
// it to finish (sync[2] >= need_ib) and read sbar with agent-scope loads.
// The workgroups whose output the IB kernel of step t reads (its window
// columns; the wall columns when it runs the walls) store with agent-scope
// stores and count themselves in sync[0].
__global__ __launch_bounds__(256) void k_adj_bulk(AdjBulk P, const float *__restrict__ sbar, float *__restrict__ fbar,
                                                  const float *__restrict__ ft, const float *__restrict__ gwin,
                                                  const double *__restrict__ body, const double *__restrict__ body1,
                                                  float *__restrict__ gbar, int need_ib) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
  const bool valid = y < P.ny;
  const size_t n = (size_t)P.nx * P.ny, cell = (size_t)x * P.ny + (valid ? y : 0);
  const int S = P.S;
  const int xc = adj_win_col(P.win_x0, body, P.nx, S), xc1 = adj_win_col(P.win_x0, body1, P.nx, S);
  const int yc = min(max((int)trunc(P.win_y0 + body[5]), 0), P.ny - S);
  // (the band of this step's own window too: its gbar stores must not
  // overtake the reads of the IB kernel of step t + 1)
  const bool walls_col = x == 0 || x >= P.nx - 2;
  const bool waiter = (x >= xc1 - 1 && x <= xc1 + S) || (x >= xc - 1 && x <= xc + S) ||
                      (!adj_walls_clear(xc1, S, P.nx) && (x <= 1 || x >= P.nx - 3));
  const bool publish = (x >= xc && x < xc + S) || (!adj_walls_clear(xc, S, P.nx) && walls_col);
  if (waiter) {
    if (threadIdx.x == 0) adj_wait(P.sync, 2, need_ib);
    __syncthreads();
  }
  const int r0 = x - xc, r1 = y - yc;
  const bool in_win = valid && r0 >= 0 && r0 < S && r1 >= 0 && r1 < S;
  float pb[9], f[9];
  if (valid) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      int xs = x + D2Q9::cx[i], ys = y + D2Q9::cy[i];
      xs = xs < 0 ? xs + P.nx : (xs >= P.nx ? xs - P.nx : xs);
      ys = ys < 0 ? ys + P.ny : (ys >= P.ny ? ys - P.ny : ys);
      const float *src = &sbar[i * n + (size_t)xs * P.ny + ys];
      pb[i] = waiter ? ld_agent(src) : *src;
      f[i] = ft[i * n + cell];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      pb[i] = 0.f;
      f[i] = D2Q9::w[i];
    }
  }
  float rho = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    rho += f[i];
    jx += (float)D2Q9::cx[i] * f[i];
    jy += (float)D2Q9::cy[i] * f[i];
  }
  float gx = 0.f, gy = 0.f;
  if (in_win) {
    gx = gwin[r0 * S + r1];
    gy = gwin[S * S + r0 * S + r1];
  }
  const float usx = jx / rho + gx * 0.5f / rho, usy = jy / rho + gy * 0.5f / rho;
  // moments of pbar: mb_k = |row k|^-2 (M pbar)_k; q = A^T pbar, r = B^T pbar
  float mb[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) mb[k] = mrow(k, pb) * Mrt2::inv_norm[k];
  float q[9], r[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    float sq = 0.f, sr = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const float m = (float)Mrt2::M[k][i];
      sq += m * (P.rates[k] * mb[k]);
      sr += m * ((1.f - 0.5f * P.rates[k]) * mb[k]);
    }
    q[i] = sq;
    r[i] = sr;
  }
  float feq[9];
  equilibrium_raw(rho, usx, usy, feq);
  float rhobar = 0.f, usxb = 0.f, usyb = 0.f, gxb = 0.f, gyb = 0.f;
  const float cg_x = gx, cg_y = gy;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float cx = (float)D2Q9::cx[i], cy = (float)D2Q9::cy[i];
    const float cu = cx * usx + cy * usy, cgv = cx * cg_x + cy * cg_y;
    rhobar += q[i] * feq[i] / rho;
    usxb += q[i] * dfeq_du(i, 0, rho, cu, usx) + r[i] * D2Q9::w[i] * (-3.f * cg_x + 9.f * cx * cgv);
    usyb += q[i] * dfeq_du(i, 1, rho, cu, usy) + r[i] * D2Q9::w[i] * (-3.f * cg_y + 9.f * cy * cgv);
    gxb += r[i] * D2Q9::w[i] * (3.f * (cx - usx) + 9.f * cu * cx);
    gyb += r[i] * D2Q9::w[i] * (3.f * (cy - usy) + 9.f * cu * cy);
  }
  gxb += usxb * 0.5f / rho;
  gyb += usyb * 0.5f / rho;
  const float jxb = usxb / rho, jyb = usyb / rho;
  rhobar -= (usxb * usx + usyb * usy) / rho;
  if (valid) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float v = pb[i] - q[i] + rhobar + jxb * (float)D2Q9::cx[i] + jyb * (float)D2Q9::cy[i];
      if (publish) st_agent(&fbar[i * n + cell], v);
      else fbar[i * n + cell] = v;
    }
  }
  if (in_win) {
    st_agent(&gbar[r0 * S + r1], gxb);
    st_agent(&gbar[S * S + r0 * S + r1], gyb);
  }
  if (publish) {
    wait_mem();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(&P.sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------------ IB workgroup
struct AdjIB {
  int nx, ny;
  double inv_T;
  int32_t *sync;   // hand-off words (adj_wait)
  int nby;         // workgroups per column of the reverse bulk kernel
  float u_in;      // inlet velocity (wall adjoint)
};

// LDS of k_adj_ib in 4-byte words (S = window side, SS = S * S):
//   uw [2 SS] float2 window velocity of f_t; acc [4 SS] the fixed-point spread
//   grid (two int64 planes, x then y); KEEP: A [NIT][2 SS] float2, the forward
//   replay's spreads; mark [SS bytes] the window cells the stencils read; list
//   [SS uint16] those cells in ascending order.
struct AdjLds {
  int acc, A, mark, list, total;
};
__host__ __device__ inline AdjLds adj_lds_layout(int S, int M, int nit, bool keep) {
  (void)M;
  const int SS = S * S;
  AdjLds L;
  L.acc = 2 * SS;
  L.A = 6 * SS;
  L.mark = L.A + (keep ? 2 * nit * SS : 0);
  L.list = L.mark + (SS + 3) / 4;
  L.total = L.list + (SS + 1) / 2;
  return L;
}

// Reverse IB step.  Thread t takes marker mk and stencil rows [r0, r0 + rows)
// as in the forward IB workgroup (four threads per marker when 4 M <= 1024,
// two when 2 M <= 1024); only the window cells the stencils read are pulled,
// zeroed and converted; every spread adds int64 fixed-point values in LDS (exact, independent of
// the order of the additions), with the index rule of the operation for
// stencils that leave the window.  KEEP: the forward replay's spread grids
// A_k stay in LDS for the reverse loop, else they are spread again.
template <int NIT, bool KEEP, int TPM>
__global__ __launch_bounds__(1024) void k_adj_ib(VsxIB2D ib, AdjIB P, const float *__restrict__ ft,
                                                 const float *__restrict__ gbar, float *__restrict__ fbar,
                                                 const double *__restrict__ body, const double *__restrict__ body1,
                                                 double *__restrict__ adj_body, double *__restrict__ grad_delta,
                                                 const double *__restrict__ dir_x, const double *__restrict__ dir_y,
                                                 double *__restrict__ u_in_prev, int done_value) {
  extern __shared__ float lds[];
  const int S = ib.win, SS = S * S, M = ib.n_markers;
  const int tid = threadIdx.x, nt = blockDim.x;
  const size_t n = (size_t)P.nx * P.ny;
  const AdjLds lay = adj_lds_layout(S, M, NIT, KEEP);
  float2 *uw = reinterpret_cast<float2 *>(lds);            // window velocity of f_t
  float *acc = lds + lay.acc;                              // spread grid (fixed point)
  float2 *Ak = reinterpret_cast<float2 *>(lds + lay.A);    // KEEP: [NIT] forward spreads
  uint8_t *mark = reinterpret_cast<uint8_t *>(lds + lay.mark);
  uint16_t *list = reinterpret_cast<uint16_t *>(lds + lay.list);
  __shared__ double s_body[6], s_body1[6], s_ab[7], s_misc[4];
  __shared__ double red[16], red4[4 * 16];
  __shared__ int s_org[4], s_scan[kIbThreads / 64 + 1];
  const IbLane l = ib_lane<TPM>(M);
  const bool act = l.mk < M;
  const int m = l.mk;

  // round trip 1: body state, marker inputs (what the replay needs; the body
  // adjoint comes from the IB kernel of step t + 1, read after its hand-off)
  if (tid < 6) {
    s_body[tid] = body[tid];
    s_body1[tid] = body1[tid];   // the state after the step, for the Newmark adjoint
  }
  double mxd = 0.0, myd = 0.0, dxm = 0.0, dym = 0.0;
  float ds2 = 0.f;
  if (act) {
    mxd = ib.marker_x[m];
    myd = ib.marker_y[m];
    ds2 = (ib.ds_arr ? ib.ds_arr[m] : ib.ds) * 2.0f;
    dxm = dir_x[m];
    dym = dir_y[m];
  }
  for (int k = tid; k < (SS + 3) / 4; k += nt) reinterpret_cast<uint32_t *>(mark)[k] = 0u;
  __syncthreads();
  if (tid == 0) {
    const int x0 = (int)trunc(ib.win_x0 + s_body[4]), y0 = (int)trunc(ib.win_y0 + s_body[5]);
    s_org[0] = x0;
    s_org[1] = y0;
    s_org[2] = min(max(x0, 0), P.nx - S);
    s_org[3] = min(max(y0, 0), P.ny - S);
  }
  __syncthreads();
  const int x0 = s_org[0], y0 = s_org[1], xc = s_org[2], yc = s_org[3];
  // objective term and Newmark adjoint (dyn.py:5-51), per DOF, once the body
  // adjoint of step t + 1 is in s_ab (thread 0)
  auto newmark_adjoint = [&]() {
    s_ab[5] += 2.0 * s_body1[5] * P.inv_T;
    const double dt = ib.dt, c1 = ib.gamma * dt, c2 = ib.beta * dt * dt;
    const double Mk = ib.mass + c1 * ib.damping + c2 * ib.stiffness;
    double kbar = 0.0;
    for (int k = 0; k < 2; ++k) {
      const bool on = (ib.dof_mask >> k) & 1;
      const double ab = on ? s_ab[k] : 0.0, vb = on ? s_ab[2 + k] : 0.0, db = on ? s_ab[4 + k] : 0.0;
      const double a = s_body[k], v = s_body[2 + k], d = s_body[4 + k];
      const double v2 = d + dt * v + dt * dt * (0.5 - ib.beta) * a;
      const double an = s_body1[k];
      const double anb = ab + c1 * vb + c2 * db;
      const double Hb = anb / Mk;
      kbar += anb * (-v2 - an * c2) / Mk;
      const double v1b = vb - ib.damping * anb / Mk, v2b = db - ib.stiffness * anb / Mk;
      s_ab[k] = dt * (1.0 - ib.gamma) * v1b + dt * dt * (0.5 - ib.beta) * v2b + Hb * ib.area;
      s_ab[2 + k] = v1b + dt * v2b;
      s_ab[4 + k] = v2b;
      s_misc[k] = Hb;
    }
    s_misc[2] = kbar;
  };

  // stencil of this thread's rows (weights and their derivatives); mark the cells it reads
  float wr[4] = {0, 0, 0, 0}, dwr[4] = {0, 0, 0, 0}, wy[4] = {0, 0, 0, 0}, dwy[4] = {0, 0, 0, 0};
  int base = 0;
  if (act) {
    const int r0 = l.first_row();
    const float rx = (float)(mxd + s_body[4] - (double)x0);
    const float ry = (float)(myd + s_body[5] - (double)y0);
    const int refx = (int)floorf(rx), refy = (int)floorf(ry);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < l.rows) {
        const float ax = (float)(refx - 1 + r0 + i) - rx;
        wr[i] = delta_kernel(ib.kernel, ax);
        dwr[i] = delta_kernel_deriv(ib.kernel, ax);
      }
    }
    // the four y nodes: with several threads per marker each evaluates its
    // share and takes the others from the marker's lanes
    if (l.tpm == 4) {
      const float ay = (float)(refy - 1 + r0) - ry;
      const float a = delta_kernel(ib.kernel, ay), da = delta_kernel_deriv(ib.kernel, ay);
      wy[0] = quad_get<0x00>(a);
      wy[1] = quad_get<0x55>(a);
      wy[2] = quad_get<0xAA>(a);
      wy[3] = quad_get<0xFF>(a);
      dwy[0] = quad_get<0x00>(da);
      dwy[1] = quad_get<0x55>(da);
      dwy[2] = quad_get<0xAA>(da);
      dwy[3] = quad_get<0xFF>(da);
    } else if (l.tpm == 2) {
      const float ay0 = (float)(refy - 1 + r0) - ry, ay1 = (float)(refy + r0) - ry;
      const float a0 = delta_kernel(ib.kernel, ay0), a1 = delta_kernel(ib.kernel, ay1);
      const float d0 = delta_kernel_deriv(ib.kernel, ay0), d1 = delta_kernel_deriv(ib.kernel, ay1);
      const float pa0 = quad_get<kQuadXor1>(a0), pa1 = quad_get<kQuadXor1>(a1);
      const float pd0 = quad_get<kQuadXor1>(d0), pd1 = quad_get<kQuadXor1>(d1);
      const bool lo = r0 == 0;
      wy[0] = lo ? a0 : pa0;
      wy[1] = lo ? a1 : pa1;
      wy[2] = lo ? pa0 : a0;
      wy[3] = lo ? pa1 : a1;
      dwy[0] = lo ? d0 : pd0;
      dwy[1] = lo ? d1 : pd1;
      dwy[2] = lo ? pd0 : d0;
      dwy[3] = lo ? pd1 : d1;
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float ay = (float)(refy - 1 + i) - ry;
        wy[i] = delta_kernel(ib.kernel, ay);
        dwy[i] = delta_kernel_deriv(ib.kernel, ay);
      }
    }
    base = (refx - 1 + r0) * S + (refy - 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < l.rows) {
#pragma unroll
        for (int j = 0; j < 4; ++j) mark[gather_index32(base + i * S + j, SS)] = 1;
      }
    }
  }
  __syncthreads();
  const int n_list = mark_scan(mark, list, SS, s_scan);   // ends with a barrier
  const bool inside = __all(base >= 0 && base + (l.rows - 1) * S + 3 < SS);
  const int base0 = base;

  // round trip 2: window velocity of f_t at the marked cells
  for (int idx = tid; idx < n_list; idx += nt) {
    const int k = list[idx];
    const size_t cell = (size_t)(xc + k / S) * P.ny + (yc + k % S);
    float rho = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float v = ft[i * n + cell];
      rho += v;
      jx += (float)D2Q9::cx[i] * v;
      jy += (float)D2Q9::cy[i] * v;
    }
    uw[k] = make_float2(jx / rho, jy / rho);
  }
  __syncthreads();

  const float Ux = (float)s_body[2], Uy = (float)s_body[3];
  float rxb = 0.f, ryb = 0.f;   // this thread's part of the marker-position adjoint
  auto pair_sum = [&](float &vx, float &vy) {
    vx = marker_sum(vx, l);
    vy = marker_sum(vy, l);
  };
  // gather of a 2-component grid at this thread's rows: value sums (summed over
  // the marker's threads) and the W-bar terms with coefficients (cx, cy)
  // (the stencil indices are recomputed in every call: shared between calls
  // they would stay live across the whole kernel)
  auto gather_with = [&](auto fetch, bool scatter_rule, float cx, float cy, float &vx, float &vy) {
    vx = vy = 0.f;
    int base = base0;
    asm volatile("" : "+v"(base));
    if (act) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i < l.rows) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int p = base + i * S + j;
            const int gi = inside ? p : (scatter_rule ? scatter_index32(p, SS) : gather_index32(p, SS));
            if (gi < 0) continue;
            const float w = wr[i] * wy[j];
            const float2 gv = fetch(gi);
            vx += w * gv.x;
            vy += w * gv.y;
            const float wb = cx * gv.x + cy * gv.y;   // W-bar of this stencil point
            rxb -= wb * dwr[i] * wy[j];
            ryb -= wb * wr[i] * dwy[j];
          }
        }
      }
    }
    pair_sum(vx, vy);
  };
  auto from_acc = [&](int gi) { return acc_get(acc, SS, gi); };
  // acc = spread of (vx, vy) with the given index rule (only marked cells are nonzero)
  auto spread_acc = [&](bool scatter_rule, float vx, float vy) {
    unsigned long long *g = reinterpret_cast<unsigned long long *>(acc);
    int base = base0;
    asm volatile("" : "+v"(base));
    for (int idx = tid; idx < n_list; idx += nt) {
      const int k = list[idx];
      g[k] = 0ull;
      g[SS + k] = 0ull;
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i < l.rows) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int p = base + i * S + j;
            const int gi = inside ? p : (scatter_rule ? scatter_index32(p, SS) : gather_index32(p, SS));
            if (gi >= 0) {
              const float w = wr[i] * wy[j];
              acc_add(acc, SS, gi, vx * w, vy * w);
            }
          }
        }
      }
    }
    __syncthreads();
  };

  // ---- forward replay of the MDF loop (ib/mdf.py:45-57)
  float ubx, uby, dFx[NIT], dFy[NIT], Fx = 0.f, Fy = 0.f;
  gather_with([&](int gi) { return uw[gi]; }, false, 0.f, 0.f, ubx, uby);
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    dFx[k] = (Ux - ubx) * ds2;
    dFy[k] = (Uy - uby) * ds2;
    Fx += dFx[k];
    Fy += dFy[k];
    spread_acc(true, dFx[k], dFy[k]);
    if constexpr (KEEP) {
      float2 *A = Ak + SS * k;
      for (int idx = tid; idx < n_list; idx += nt) {
        const int c = list[idx];
        A[c] = acc_get(acc, SS, c);
      }
    }
    float ax = 0.f, ay = 0.f;
    if (act) {
      int base = base0;
      asm volatile("" : "+v"(base));
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i < l.rows) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float2 av = acc_get(acc, SS, gather_index32(base + i * S + j, SS));
            const float w = wr[i] * wy[j];
            ax += w * (av.x * 0.5f);
            ay += w * (av.y * 0.5f);
          }
        }
      }
    }
    pair_sum(ax, ay);
    ubx += ax;
    uby += ay;
    __syncthreads();
  }
  // ---- hand-offs: the IB kernel of step t + 1 has finished (sync[2]; it ran
  // on the other IB stream and wrote the body adjoint, the marker gradients
  // and sync[1]), then the reverse bulk kernel of this step has published the
  // force-window adjoint and fbar_t on the window and wall columns
  double gd0 = 0.0;
  {
    // published columns: the window; the wall columns 0, nx - 2, nx - 1 when
    // this kernel runs the walls of step t - 1
    int npub = S;
    if (!adj_walls_clear(xc, S, P.nx)) npub += (xc > 0) + (xc + S <= P.nx - 2) + (xc + S <= P.nx - 1);
    if (tid == 0) {
      adj_wait(P.sync, 2, done_value - 1);
      const int target = __hip_atomic_load(&P.sync[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + P.nby * npub;
      adj_wait(P.sync, 0, target);
      __hip_atomic_store(&P.sync[1], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (tid < 7) s_ab[tid] = ld_agent(&adj_body[tid]);   // [6]: the stiffness accumulator
    if (act && l.first_row() == 0) gd0 = ld_agent(&grad_delta[m]);
    __syncthreads();
    if (tid == 0) newmark_adjoint();
    __syncthreads();
  }
  // ---- reverse
  // the force-window adjoint at the marked cells, staged in LDS over the spread
  // grid (its marked cells, the only ones read, are zeroed by every spread)
  {
    float2 *gst = reinterpret_cast<float2 *>(acc);
    for (int idx = tid; idx < n_list; idx += nt) {
      const int k = list[idx];
      gst[k] = make_float2(ld_agent(&gbar[k]), ld_agent(&gbar[SS + k]));
    }
    __syncthreads();
  }
  // g = W^T F (scatter rule); H = -sum F + a area
  float Fbx, Fby;
  gather_with([&](int gi) { return reinterpret_cast<const float2 *>(acc)[gi]; }, true, Fx, Fy, Fbx, Fby);
  __syncthreads();   // the staged values are read before a spread zeroes the grid
  Fbx -= (float)s_misc[0];
  Fby -= (float)s_misc[1];
  float ubbx = 0.f, ubby = 0.f, Vbx = 0.f, Vby = 0.f;
#pragma unroll
  for (int k = NIT - 1; k >= 0; --k) {
    // ub_{k+1} = ub_k + W (0.5 A_k), A_k = W^T dF_k
    {
      float vx, vy;
      if constexpr (KEEP) {
        const float2 *A = Ak + SS * k;
        gather_with([&](int gi) { return A[gi]; }, false, 0.5f * ubbx, 0.5f * ubby, vx, vy);
      } else {
        spread_acc(true, dFx[k], dFy[k]);
        gather_with(from_acc, false, 0.5f * ubbx, 0.5f * ubby, vx, vy);
      }
    }
    __syncthreads();
    spread_acc(false, 0.5f * ubbx, 0.5f * ubby);   // Abar_k = 0.5 W^T ubbar_{k+1} (gather rule)
    float abx, aby;
    gather_with(from_acc, true, dFx[k], dFy[k], abx, aby);   // dFbar += W Abar_k (scatter rule)
    __syncthreads();
    const float dFbx = Fbx + abx, dFby = Fby + aby;
    ubbx -= ds2 * dFbx;
    ubby -= ds2 * dFby;
    Vbx += ds2 * dFbx;
    Vby += ds2 * dFby;
  }
  // ub_0 = W u: W-bar term and the window-velocity adjoint ubar = W^T ubbar_0 (gather rule)
  {
    float vx, vy;
    gather_with([&](int gi) { return uw[gi]; }, false, ubbx, ubby, vx, vy);
  }
  spread_acc(false, ubbx, ubby);
  // fbar_t += (c_i - u) . ubar / rho on the window cells the stencils reach
  // two cells per thread and pass, every load issued before the first store
  // (agent-scope loads and stores stay in program order: a load after a
  // store would wait for it)
  for (int i0 = tid; i0 < n_list; i0 += 2 * nt) {
    float fv[2][9], fb[2][9];
    float2 ub[2];
    size_t cell[2];
    bool on[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int idx = i0 + h * nt;
      const int k = idx < n_list ? list[idx] : 0;
      ub[h] = idx < n_list ? acc_get(acc, SS, k) : make_float2(0.f, 0.f);
      on[h] = ub[h].x != 0.f || ub[h].y != 0.f;
      cell[h] = (size_t)(xc + k / S) * P.ny + (yc + k % S);
      if (on[h]) {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          fv[h][i] = ft[i * n + cell[h]];
          fb[h][i] = ld_agent(&fbar[i * n + cell[h]]);
        }
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (!on[h]) continue;
      const int k = list[i0 + h * nt];
      float rho = 0.f;
#pragma unroll
      for (int i = 0; i < 9; ++i) rho += fv[h][i];
      const float2 u = uw[k];
#pragma unroll
      for (int i = 0; i < 9; ++i)
        st_agent(&fbar[i * n + cell[h]],
                 fb[h][i] + (ub[h].x * ((float)D2Q9::cx[i] - u.x) + ub[h].y * ((float)D2Q9::cy[i] - u.y)) / rho);
    }
  }
  if (u_in_prev && !adj_walls_clear(xc, S, P.nx)) {
    // the wall adjoint of step t - 1 when the window touches a wall column
    // (else k_adj_walls on the bulk stream): after the window update
    wait_mem();
    __syncthreads();
    const double part = adj_walls_rows<true>(fbar, ft, P.nx, P.ny, P.u_in);
    const double tot = block_sum_d(part, red);
    if (tid == 0) *u_in_prev = tot;
  }
  // marker-position adjoints: delta (radial direction), displacement; the
  // marker-level values V are held by every thread of a marker, counted once
  const bool lead = act && l.first_row() == 0;
  float rxm = rxb, rym = ryb;
  pair_sum(rxm, rym);
  if (lead) st_agent(&grad_delta[m], gd0 + ((double)rxm * dxm + (double)rym * dym));
  double sums[4] = {act ? (double)rxb : 0.0, act ? (double)ryb : 0.0, lead ? (double)Vbx : 0.0,
                    lead ? (double)Vby : 0.0};
  block_sum4_d(sums, red4);
  const double srx = sums[0], sry = sums[1], svx = sums[2], svy = sums[3];
  if (tid == 0) {
    // agent-scope stores: read by the IB kernel of step t - 1 on the other IB stream
    st_agent(&adj_body[0], s_ab[0]);
    st_agent(&adj_body[1], s_ab[1]);
    st_agent(&adj_body[2], s_ab[2] + svx);
    st_agent(&adj_body[3], s_ab[3] + svy);
    st_agent(&adj_body[4], s_ab[4] + srx);
    st_agent(&adj_body[5], s_ab[5] + sry);
    st_agent(&adj_body[6], s_ab[6] + s_misc[2]);
  }
  // every agent-scope store of fbar_t has completed: the reverse bulk kernel of
  // step t - 1 may read it
  wait_mem();
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&P.sync[2], done_value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_adj_final(const double *hist_d, int T, const double *adj_body, const double *u_in_part, double *out) {
  __shared__ double red[16];
  double j = 0.0, u = 0.0;
  for (int t = threadIdx.x; t < T; t += blockDim.x) {
    j += hist_d[2 * t + 1] * hist_d[2 * t + 1];
    u += u_in_part[t];
  }
  j = block_sum_d(j, red);
  u = block_sum_d(u, red);
  if (threadIdx.x == 0) {
    out[0] = j / T;
    out[1] = adj_body[6];
    out[2] = u;
  }
}

// ------------------------------------------------------------------ driver

int validate_adj(const VsxLbm *p, const VsxIB2D *ib, const VsxAdj2D *a) {
  VSX_REQUIRE(a, "null adjoint parameters");
  VSX_REQUIRE(p->collision == VSX_COLL_MRT && p->forcing == VSX_FORCE_GUO && p->storage == VSX_STORE_F32,
              "the adjoint supports MRT collision with Guo forcing in FP32 storage");
  VSX_REQUIRE(p->n_walls == 2 && p->walls[0].kind == VSX_WALL_EQ && p->walls[0].mode == VSX_MODE_PLAIN &&
                  p->walls[0].loc == VSX_LOC_LEFT && !p->walls[0].ux_arr && !p->walls[0].uy_arr &&
                  !p->walls[0].rho_arr && p->walls[0].rho == 1.f && p->walls[0].uy == 0.f &&
                  p->walls[1].kind == VSX_WALL_NEE && p->walls[1].mode == VSX_MODE_PRESSURE &&
                  p->walls[1].loc == VSX_LOC_RIGHT && !p->walls[1].rho_arr && p->walls[1].rho == 1.f,
              "the adjoint supports walls [equilibrium inlet (left, rho 1, uy 0), pressure NEE outlet (right, rho 1)]");
  VSX_REQUIRE(!p->g && p->g_uniform[0] == 0.f && p->g_uniform[1] == 0.f, "no extra body force in the adjoint");
  VSX_REQUIRE(ib->n_markers <= kIbThreads, "the adjoint takes at most %d markers", kIbThreads);
  VSX_REQUIRE(ib->n_iter >= 1 && ib->n_iter <= kMaxIter, "the adjoint takes 1..%d MDF iterations", kMaxIter);
  VSX_REQUIRE(a->n_steps >= 1 && a->seg >= 1, "bad step count / segment");
  VSX_REQUIRE(a->dir_x && a->dir_y && a->ckpt && a->ckpt_body && a->tape_f && a->tape_g && a->tape_body && a->adj_a &&
                  a->adj_b && a->grid && a->adj_body && a->hist_d && a->u_in_part && a->grad_delta && a->out &&
                  a->sync,
              "null adjoint buffer");
  VSX_REQUIRE(p->nx >= 4, "grid too small");
  return VSX_OK;
}

template <int NIT>
void launch_adj_ib(const VsxIB2D &ib, const AdjIB &P, const float *ft, const float *gbar, float *fbar,
                   const double *body, const double *body1, const VsxAdj2D &a, double *u_in_prev, int done_value,
                   int max_lds, hipStream_t s) {
  const size_t keep = sizeof(float) * (size_t)adj_lds_layout(ib.win, ib.n_markers, NIT, true).total;
  const size_t drop = sizeof(float) * (size_t)adj_lds_layout(ib.win, ib.n_markers, NIT, false).total;
  // one kernel per lane mapping (threads per marker): a kernel holds the code of one only
  auto go = [&](auto tpm) {
    constexpr int T = decltype(tpm)::value;
    if (keep + 2048 <= (size_t)max_lds)
      hipLaunchKernelGGL((k_adj_ib<NIT, true, T>), dim3(1), dim3(kIbThreads), keep, s, ib, P, ft, gbar, fbar, body,
                         body1, a.adj_body, a.grad_delta, a.dir_x, a.dir_y, u_in_prev, done_value);
    else
      hipLaunchKernelGGL((k_adj_ib<NIT, false, T>), dim3(1), dim3(kIbThreads), drop, s, ib, P, ft, gbar, fbar, body,
                         body1, a.adj_body, a.grad_delta, a.dir_x, a.dir_y, u_in_prev, done_value);
  };
  switch (ib_tpm(ib.n_markers, kIbThreads)) {
    case 4: go(std::integral_constant<int, 4>()); break;
    case 2: go(std::integral_constant<int, 2>()); break;
    default: go(std::integral_constant<int, 1>()); break;
  }
}

int viv_adjoint(const VsxLbm *p, const VsxIB2D *ib, const VsxAdj2D *a, const float *f0, void *wa, void *wb,
                float *f_tmp, void *stream) {
  int rc = validate_lbm<D2Q9>(p);
  if (rc) return rc;
  rc = validate_ib2d(p, ib);
  if (rc) return rc;
  rc = validate_adj(p, ib, a);
  if (rc) return rc;
  VSX_REQUIRE(f0 && wa && wb && f_tmp, "null buffer");
  int max_lds = 0;
  if (int rc_ = max_lds_per_block(&max_lds)) return rc_;
  VSX_REQUIRE(sizeof(float) * (size_t)adj_lds_layout(ib->win, ib->n_markers, 1, false).total + 2048 <= (size_t)max_lds,
              "IB window %d too large for the adjoint workgroup", ib->win);
  hipStream_t s = as_stream(stream);
  const int T = a->n_steps, K = a->seg, nseg = (T + K - 1) / K;
  // the fused forward pass waits on ready flags that must start below its first step
  VSX_HIP(hipMemsetAsync(ib->win_origin, 0, sizeof(int32_t) * 8, s));
  const size_t n = (size_t)p->nx * p->ny, fsz = 9 * n, gsz = 2 * (size_t)ib->win * ib->win;
  VsxIB2D w = *ib;
  // forward pass: checkpoints at segment starts
  VSX_HIP(hipMemcpyAsync(a->ckpt, f0, sizeof(float) * fsz, hipMemcpyDeviceToDevice, s));
  VSX_HIP(hipMemcpyAsync(a->ckpt_body, ib->body, sizeof(double) * 6, hipMemcpyDeviceToDevice, s));
  for (int sg = 0; sg < nseg; ++sg) {
    const int t0 = sg * K, nn = std::min(K, T - t0);
    w.step0 = t0;
    w.fused = w.halo ? 2 : 1;   // the forward pass as vsx_viv2d_run runs it (pipelined when it can be)
    w.g_win = ib->g_win;
    w.hist_d = a->hist_d + 2 * (size_t)t0;
    w.hist_h = nullptr;
    float *dst = sg + 1 < nseg ? a->ckpt + (size_t)(sg + 1) * fsz : f_tmp;
    rc = viv_run(p, &w, a->ckpt + (size_t)sg * fsz, dst, wa, wb, nn, stream);
    if (rc) return rc;
    if (sg + 1 < nseg)
      VSX_HIP(hipMemcpyAsync(a->ckpt_body + 6 * (size_t)(sg + 1), ib->body + 6 * ((t0 + nn) & 1), sizeof(double) * 6,
                             hipMemcpyDeviceToDevice, s));
  }
  // backward pass
  VSX_HIP(hipMemsetAsync(a->adj_a, 0, sizeof(float) * fsz, s));
  VSX_HIP(hipMemsetAsync(a->adj_body, 0, sizeof(double) * 8, s));
  VSX_HIP(hipMemsetAsync(a->grad_delta, 0, sizeof(double) * ib->n_markers, s));
  VSX_HIP(hipMemsetAsync(a->u_in_part, 0, sizeof(double) * T, s));
  float *cur = a->adj_a, *nxt = a->adj_b;
  AdjBulk PB;
  PB.nx = p->nx;
  PB.ny = p->ny;
  PB.S = ib->win;
  for (int k = 0; k < 9; ++k) PB.rates[k] = p->mrt_rates[k];
  PB.win_x0 = ib->win_x0;
  PB.win_y0 = ib->win_y0;
  const dim3 bgrid((p->ny + 255) / 256, p->nx);
  VSX_HIP(hipMemsetAsync(a->sync, 0, sizeof(int32_t) * 4, s));
  PB.sync = a->sync;
  AdjIB PI;
  PI.nx = p->nx;
  PI.ny = p->ny;
  PI.inv_T = 1.0 / T;
  PI.sync = a->sync;
  PI.nby = (int)bgrid.x;
  PI.u_in = p->walls[0].ux;
  // Three streams: the recompute of segment sg - 1 (s2) overlaps the reverse
  // pass of segment sg; in the reverse pass the bulk and wall kernels run on
  // the caller's stream and the IB kernels on s3, each IB kernel replaying the
  // forward MDF while the bulk kernel of its step runs and taking that
  // kernel's output through the hand-off words a->sync (k_adj_bulk,
  // k_adj_ib).  The IB kernels alternate between two streams, s3 and s4:
  // the one of step t replays its forward MDF while the one of step t + 1
  // runs its reverse half, and waits for it (sync[2]) before it reads the
  // body adjoint.  Every kernel that waits on another is enqueued after it,
  // so the waits end whether or not the streams share a hardware queue.
  // s2, s3 and s4 are the library's streams of the device (the 3D bulk
  // stream, the one the pipelined forward pass runs its IB kernels on, and a
  // third), created once: a stream created per call may land on the caller's
  // hardware queue, where the two kernels of a step would run one after the
  // other.  Tapes are double-buffered by segment parity.
  PipeRes *res = nullptr;
  if ((rc = pipe_res(&res))) return rc;
  hipStream_t s2 = res->b, s3 = res->s, s4 = res->x;
  std::vector<hipEvent_t> ev_rec(nseg, nullptr), ev_rev(nseg, nullptr), ev_rev3(nseg, nullptr),
      ev_rev4(nseg, nullptr);
  hipEvent_t ev_fwd = nullptr, ev_end3 = nullptr, ev_end4 = nullptr;
  auto cleanup = [&]() {
    for (auto e : ev_rec) if (e) (void)hipEventDestroy(e);
    for (auto e : ev_rev) if (e) (void)hipEventDestroy(e);
    for (auto e : ev_rev3) if (e) (void)hipEventDestroy(e);
    for (auto e : ev_rev4) if (e) (void)hipEventDestroy(e);
    if (ev_fwd) (void)hipEventDestroy(ev_fwd);
    if (ev_end3) (void)hipEventDestroy(ev_end3);
    if (ev_end4) (void)hipEventDestroy(ev_end4);
  };
  auto check = [&](hipError_t e) { return e == hipSuccess ? VSX_OK : fail(VSX_ERR_HIP, "adjoint: %s", hipGetErrorString(e)); };
  auto tf = [&](int sg, int i) { return a->tape_f + ((size_t)(sg & 1) * (K + 1) + i) * fsz; };
  auto tg = [&](int sg, int i) { return a->tape_g + ((size_t)(sg & 1) * K + i) * gsz; };
  auto tb = [&](int sg, int i) { return a->tape_body + ((size_t)(sg & 1) * (K + 1) + i) * 6; };
  auto recompute = [&](int sg) -> int {
    const int t0 = sg * K, nn = std::min(K, T - t0);
    int r = check(hipMemcpyAsync(tf(sg, 0), a->ckpt + (size_t)sg * fsz, sizeof(float) * fsz, hipMemcpyDeviceToDevice, s2));
    if (!r) r = check(hipMemcpyAsync(tb(sg, 0), a->ckpt_body + 6 * (size_t)sg, sizeof(double) * 6, hipMemcpyDeviceToDevice, s2));
    // the one-launch form's ready flags count up from zero within a segment
    if (!r) r = check(hipMemsetAsync(ib->win_origin, 0, sizeof(int32_t) * 8, s2));
    for (int i = 0; i < nn && !r; ++i) {
      // one step in the one-launch form (the IB workgroup as block 0 of the
      // bulk launch, then the streaming pass into the tape), with the tape as
      // its body ring: an even step0 makes the IB read the body state from
      // body[0..5] = tape row i and write the next one to body[6..11] = tape
      // row i + 1 (no copy per step); step0 = 2 i keeps the flags increasing
      w.step0 = 2 * i;
      w.body = tb(sg, i);
      w.fused = 1;
      w.hist_d = nullptr;
      w.hist_h = nullptr;
      w.g_win = tg(sg, i);   // the IB workgroup writes this step's force window into the tape
      r = viv_run(p, &w, tf(sg, i), tf(sg, i + 1), wa, wb, 1, s2);
    }
    if (!r) r = check(hipEventCreateWithFlags(&ev_rec[sg], hipEventDisableTiming));
    if (!r) r = check(hipEventRecord(ev_rec[sg], s2));
    return r;
  };
  auto event = [&](hipEvent_t &e, hipStream_t st) {
    int r = check(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return r ? r : check(hipEventRecord(e, st));
  };
  rc = event(ev_fwd, s);
  if (!rc) rc = check(hipStreamWaitEvent(s2, ev_fwd, 0));
  if (!rc) rc = check(hipStreamWaitEvent(s3, ev_fwd, 0));
  if (!rc) rc = check(hipStreamWaitEvent(s4, ev_fwd, 0));
  if (!rc) rc = recompute(nseg - 1);
  if (!rc) {
    // the walls of the last step (those of every other step run inside the IB
    // kernel of the step after it)
    const int tl = (nseg - 1) * K, nl = T - tl;
    rc = check(hipStreamWaitEvent(s, ev_rec[nseg - 1], 0));
    hipLaunchKernelGGL(k_adj_walls, dim3(1), dim3(1024), 0, s, cur, (const float *)tf(nseg - 1, nl), p->nx, p->ny,
                       p->walls[0].ux, a->u_in_part + (T - 1), (const double *)nullptr, ib->win_x0, ib->win);
  }
  for (int sg = nseg - 1; sg >= 0 && !rc; --sg) {
    const int t0 = sg * K, nn = std::min(K, T - t0);
    rc = check(hipStreamWaitEvent(s, ev_rec[sg], 0));
    if (!rc) rc = check(hipStreamWaitEvent(s3, ev_rec[sg], 0));
    if (!rc) rc = check(hipStreamWaitEvent(s4, ev_rec[sg], 0));
    for (int i = nn - 1; i >= 0 && !rc; --i) {
      const int t = t0 + i;
      const float *ft = tf(sg, i);
      const double *bt = tb(sg, i), *bt1 = tb(sg, i + 1);
      hipLaunchKernelGGL(k_adj_bulk, bgrid, dim3(256), 0, s, PB, (const float *)cur, nxt, ft,
                         (const float *)tg(sg, i), bt, bt1, a->grid, T - 1 - t);
      double *u_prev = t > 0 ? a->u_in_part + (t - 1) : nullptr;
      if (u_prev)
        hipLaunchKernelGGL(k_adj_walls, dim3(1), dim3(1024), 0, s, nxt, ft, p->nx, p->ny, p->walls[0].ux, u_prev, bt,
                           ib->win_x0, ib->win);
      hipStream_t si = (t & 1) ? s3 : s4;
      switch (ib->n_iter) {
        case 1: launch_adj_ib<1>(*ib, PI, ft, a->grid, nxt, bt, bt1, *a, u_prev, T - t, max_lds, si); break;
        case 2: launch_adj_ib<2>(*ib, PI, ft, a->grid, nxt, bt, bt1, *a, u_prev, T - t, max_lds, si); break;
        case 3: launch_adj_ib<3>(*ib, PI, ft, a->grid, nxt, bt, bt1, *a, u_prev, T - t, max_lds, si); break;
        default: launch_adj_ib<4>(*ib, PI, ft, a->grid, nxt, bt, bt1, *a, u_prev, T - t, max_lds, si); break;
      }
      rc = check(hipGetLastError());
      std::swap(cur, nxt);
    }
    if (!rc) rc = event(ev_rev[sg], s);
    if (!rc) rc = event(ev_rev3[sg], s3);
    if (!rc) rc = event(ev_rev4[sg], s4);
    if (!rc && sg > 0) {
      // tape parity (sg - 1) & 1 was last read by the reverse pass of segment sg + 1
      if (sg + 1 < nseg) rc = check(hipStreamWaitEvent(s2, ev_rev[sg + 1], 0));
      if (!rc && sg + 1 < nseg) rc = check(hipStreamWaitEvent(s2, ev_rev3[sg + 1], 0));
      if (!rc && sg + 1 < nseg) rc = check(hipStreamWaitEvent(s2, ev_rev4[sg + 1], 0));
      if (!rc) rc = recompute(sg - 1);
    }
  }
  if (!rc) rc = event(ev_end3, s3);
  if (!rc) rc = event(ev_end4, s4);
  if (!rc) rc = check(hipStreamWaitEvent(s, ev_end3, 0));
  if (!rc) rc = check(hipStreamWaitEvent(s, ev_end4, 0));
  cleanup();
  if (rc) return rc;
  hipLaunchKernelGGL(k_adj_final, dim3(1), dim3(1024), 0, s, (const double *)a->hist_d, T, (const double *)a->adj_body,
                     (const double *)a->u_in_part, a->out);
  VSX_CHECK_LAUNCH();
  return VSX_OK;
}

}  // namespace vsx

using namespace vsx;

extern "C" int vsx_viv2d_adjoint(const VsxLbm *p, const VsxIB2D *ib, const VsxAdj2D *adj, const float *f0,
                                 void *work_a, void *work_b, float *f_tmp, void *stream) {
  return viv_adjoint(p, ib, adj, f0, work_a, work_b, f_tmp, stream);
}
