/*
 * vivsim_amd — C ABI of the MI355X (gfx950) IB-LBM kernels.
 *
 * The reference (haimingz/vivsim) is a pure-Python/JAX library whose public
 * surface is the module functions of vivsim.lbm, vivsim.lbm3d, vivsim.ib,
 * vivsim.ib3d and vivsim.dyn.  Each entry point below replaces one of those
 * functions (cited as reference file:line) or one of the user-composed step
 * functions of the reference examples.  The Python modules in
 * vivsim_amd/{lbm,lbm3d,ib,ib3d,dyn} bind these symbols through ctypes.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers to contiguous C-order buffers in
 *    the reference layout: f[Q][NX][NY] (D2Q9) / f[Q][NX][NY][NZ] (D3Q19),
 *    rho[...], u[D][...], marker arrays [M][C], stencils [M][S].
 *  - Field dtype float32, stencil indices int32, structural state float64.
 *  - `stream` is a hipStream_t (NULL = default stream).  No entry point
 *    synchronises the host or allocates device memory; buffers are owned by
 *    the caller.
 *  - Every function returns a status code (VSX_OK on success).  The message
 *    of the last failure on the calling thread is vsx_last_error().
 *    VSX_ERR_VALUE maps to Python ValueError, VSX_ERR_KEY to KeyError (the
 *    reference raises KeyError for an unknown wall `loc`, e.g.
 *    lbm/boundary/nee.py:43), VSX_ERR_HIP to RuntimeError.
 */
#ifndef VIVSIM_AMD_H
#define VIVSIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ enums */
enum { VSX_OK = 0, VSX_ERR_VALUE = 1, VSX_ERR_KEY = 2, VSX_ERR_HIP = 3 };

/* wall locations: lbm/lattice.py:64-109, lbm3d/lattice.py:64-125 */
enum { VSX_LOC_LEFT = 0, VSX_LOC_RIGHT = 1, VSX_LOC_BOTTOM = 2, VSX_LOC_TOP = 3,
       VSX_LOC_BACK = 4, VSX_LOC_FRONT = 5 };

/* wall schemes: boundary/{nee,nebb,eq,bb}.py */
enum { VSX_WALL_NEE = 0, VSX_WALL_NEBB = 1, VSX_WALL_EQ = 2, VSX_WALL_BB = 3,
       VSX_WALL_SPECULAR = 4 };

/* wall wrappers: boundary/_helpers.py:180-239 (plain core, velocity, pressure, force-corrected) */
enum { VSX_MODE_PLAIN = 0, VSX_MODE_VELOCITY = 1, VSX_MODE_PRESSURE = 2,
       VSX_MODE_FORCE_CORRECTED = 3 };

enum { VSX_COLL_BGK = 0, VSX_COLL_MRT = 1, VSX_COLL_KBC = 2, VSX_COLL_REG = 3 };
enum { VSX_FORCE_NONE = 0, VSX_FORCE_GUO = 1, VSX_FORCE_EDM = 2 };

/* ib/kernels.py:4-61 (+ a 2-point hat kernel, not in the reference) */
enum { VSX_KERNEL_PESKIN_2PT = 0, VSX_KERNEL_PESKIN_3PT = 1, VSX_KERNEL_PESKIN_4PT = 2,
       VSX_KERNEL_COSINE_4PT = 3 };

/* population storage of the fused steps: FP32, or 16-bit deviations f_i - feq_i(1, store_u) as
 * FP16, BF16 or E3M12 (sign, 15-bit magnitude code: 12-13 significant bits, |deviation| < 0.249,
 * larger values saturate) */
enum { VSX_STORE_F32 = 0, VSX_STORE_F16S = 1, VSX_STORE_BF16S = 2, VSX_STORE_E3M12S = 3 };

/* phases of the fused pull-scheme step (the stored state is post-collision) */
enum { VSX_PHASE_COLLIDE = 0,   /* read f locally (post-stream state), collide, write          */
       VSX_PHASE_FULL = 1,      /* pull-stream + walls, collide, write                          */
       VSX_PHASE_STREAM = 2 };  /* pull-stream + walls, write (recovers the reference's state)  */

/* ------------------------------------------------------------------ structs */

/* One domain wall.  Scalars apply when the matching *_arr pointer is NULL.
 * Per-node arrays have the wall's shape: (N,) in 2D, (N1, N2) in 3D with the
 * two remaining axes in x, y, z order (boundary/_helpers.py:53-77). */
typedef struct VsxWall {
  int32_t kind, mode, loc, reserved0;
  float rho, ux, uy, uz;
  float gx, gy, gz, reserved1;
  const float *rho_arr, *ux_arr, *uy_arr, *uz_arr;
  const float *gx_arr, *gy_arr, *gz_arr;
} VsxWall;

/* Parameters of a fused LBM step (2D: nz = 1). */
typedef struct VsxLbm {
  int32_t nx, ny, nz;
  int32_t collision, forcing, storage;
  float omega;
  float mrt_rates[19];     /* diag(S) in the reference's moment basis (collision/mrt.py:28-44) */
  int32_t n_walls;         /* walls[0..n_walls) are applied in order after streaming           */
  VsxWall walls[6];
  const float *g;          /* optional full body-force field [D][NX][NY(][NZ)] (device)        */
  float g_uniform[3];      /* uniform body force added to g                                     */
  float store_u[3];        /* 16-bit storages keep f_i - feq_i(1, store_u) (0: f_i - w_i); a
                              reference velocity near the flow's keeps the stored numbers small */
  int32_t reserved;
} VsxLbm;

/* 2D immersed body (rigid cylinder on springs) coupled to a VsxLbm step.
 * Step composition: examples/2d/vortex_induced_vibration.py:96-148. */
typedef struct VsxIB2D {
  int32_t n_markers, kernel, n_iter, win;   /* win: side of the square IB window            */
  const double *marker_x, *marker_y;        /* initial marker coordinates (device, FP64)    */
  const float *ds_arr;                      /* per-marker ds (device) or NULL -> ds          */
  float ds;
  int32_t dof_mask;                         /* bit0: x DOF moves, bit1: y DOF moves          */
  double win_x0, win_y0;                    /* window origin before displacement            */
  double mass, stiffness, damping, area;    /* Newmark (dyn.py:5-51), h += a * area          */
  double dt, gamma, beta;
  double *body;        /* device FP64 [2][6]: slot (step0 + t) & 1 holds [a0 a1 v0 v1 d0 d1]
                          before step t; step t writes the other slot                       */
  float *g_win;        /* device [2][win][win]: force window of the last step (within a fused
                          step it first carries the window velocity to the IB workgroup)   */
  double *hist_d, *hist_h;   /* optional device histories [n_steps][2]                      */
  int32_t *win_origin; /* device int32 [8]: [0..1] window origin x, y of the last step;
                          [2] ready flag (= step0 + t + 1 once step t's force window is
                          written); [3] error flag (1 if a bounded wait ran out); [4..5]
                          window-tile counters (batched step; [4] also the hand-off tile
                          counter of the pipelined step); [6] = step0 + t + 1 once the IB
                          workgroup of step t has pulled its window (fused == 1: the bulk
                          tiles wait for it) or once the bulk pass of step t has handed the
                          cells around the window to the next IB kernel (fused == 2);
                          [7] reserved                                                       */
  int32_t step0;       /* global index of the first step of this run (0 after a reset of the
                          body state, which must also zero win_origin)                       */
  int32_t fused;       /* 0: separate IB and bulk kernels; 1: one launch per step, the IB
                          workgroup overlapped with the bulk tiles; 2: pipelined, the IB
                          kernel of step t + 1 (on a second stream) overlapped with the bulk
                          pass of step t (needs `halo`, and the window >= 2 cells inside
                          the grid)                                                          */
  int32_t reserved;
  float *halo;         /* device [2][9][win + 4][win + 4] FP32 (fused == 2): post-collision
                          populations around the window, handed from the bulk pass of a step
                          to the IB kernel of the next (half (step0 + t) & 1 by step t)     */
  int32_t *status;     /* optional pinned host int32 (hipHostMalloc) or NULL: the kernels also
                          store their error codes there (1 = a bounded wait ran out, 2 = the
                          pipelined window left the hand-off buffer), so a host can check a
                          run without copying win_origin[3] back                            */
} VsxIB2D;

/* 3D immersed body held fixed in a D3Q19 VsxLbm step (flow past a sphere).
 * Step composition: examples/3d/flow_past_sphere.py:139-171; the stencil of the
 * fixed markers is computed once, relative to the window origin, as in
 * flow_past_sphere.py:114-121 (vsx_ib_get_ib_stencil3d on window-relative
 * coordinates), and transposed once with vsx_ib3d_transpose_stencil. */
typedef struct VsxIB3D {
  int32_t n_markers, n_iter;
  int32_t origin[3];          /* window origin in the grid (window inside the grid)            */
  int32_t win[3];             /* window sizes (S0, S1, S2); cells = S0 * S1 * S2               */
  const float *W;             /* device [M][64] stencil weights                                */
  const int32_t *I;           /* device [M][64] flat window indices                            */
  const int32_t *cell_ptr;    /* device [cells + 1]: stencil slots of window cell k are         */
  const int32_t *cell_slot;   /*   cell_slot[cell_ptr[k] .. cell_ptr[k+1]) (slot = m * 64 + s) */
  const float *u_target;      /* device [M][3] target marker velocity                          */
  const float *ds_arr;        /* device [M] or NULL -> ds                                      */
  float ds;
  int32_t pipelined;          /* 1: the IB of step t + 1 runs on a second stream while the bulk
                                 pass of step t streams (needs `halo`, the window 2+ cells inside
                                 the grid, and walls without shared corners)                   */
  float *g_win;               /* device [3][cells]: force window of the last step              */
  float *work;                /* device scratch, vsx_ib3d_work_floats(M, cells) floats           */
  double *hist_h;             /* optional device [n_steps][3]: total marker force (FP64)        */
  float *halo;                /* device [2][19][S0 + 2][S1 + 2][S2 + 2] FP32 (pipelined): the
                                 post-collision populations around the window, handed from the
                                 bulk pass of a step to the IB of the next                       */
} VsxIB3D;

/* One case of a batched 2D VIV sweep.  All cases of a batch advance together,
 * one launch per step for the whole batch (IB workgroups of every case plus
 * the bulk tiles of every case); the step of each case is the one of
 * vsx_viv2d_run (examples/2d/vortex_induced_vibration.py:96-157). */
typedef struct VsxVivCase {
  VsxLbm lbm;          /* per-case grid, omega, walls; collision, forcing and storage equal across the batch */
  VsxIB2D ib;          /* per-case body; ib.step0 and ib.fused are not used; the histories are indexed
                          by the batch clock (global step index)                                       */
  int64_t f_offset;    /* offset, in populations, of the case's f[9][nx][ny] in the batch buffers     */
  int32_t tile_y;      /* bulk tile = (1024 / tile_y) x tile_y cells; tile_y in {64, 128, 256, 512, 1024} */
  int32_t hist_cap;    /* steps that ib.hist_d / ib.hist_h can hold                                    */
} VsxVivCase;

/* A prepared batch (filled by vsx_viv2d_batch_prepare). */
typedef struct VsxVivBatch {
  int32_t n_cases, collision, storage, lds_bytes;
  int64_t n_blocks;    /* workgroups of one step launch                                                 */
  int64_t total_pops;  /* populations in each batch buffer (f_in, f_out, work_a, work_b)                */
  void *table;         /* device table, vsx_viv2d_batch_table_bytes bytes, caller-owned                 */
  int32_t *clock;      /* device int32, caller-owned and zeroed before the first run: global index of
                          the next step; every run advances it by n_steps                               */
} VsxVivBatch;

/* Reverse-mode (adjoint) run of the 2D VIV step with sqrt(T) checkpointing.
 * The reference differentiates its jitted step with jax.grad; here the
 * backward pass is hand written.  Objective J = (1/T) sum_{t=1..T} d_y(t)^2;
 * parameters: marker radial offsets delta_m (the markers of the VsxIB2D are
 * X_m = X0_m + delta_m (dir_x_m, dir_y_m)), the spring stiffness and the inlet
 * velocity of the equilibrium inlet.  Supported step: D2Q9 MRT + Guo forcing,
 * walls [equilibrium inlet at left, pressure-NEE outlet at right], one marker
 * per thread of the IB workgroup (M <= 1024).  All buffers are caller-owned. */
typedef struct VsxAdj2D {
  int32_t n_steps, seg;     /* T and the checkpoint interval (steps per recomputed segment)       */
  const double *dir_x, *dir_y;   /* device [M]: d X_m / d delta_m                                  */
  float *ckpt;              /* device [ceil(T/seg)][9][n]: post-streaming checkpoints              */
  double *ckpt_body;        /* device [ceil(T/seg)][6]                                              */
  float *tape_f;            /* device [2][seg + 1][9][n]: post-streaming states of a segment (two
                               segments in flight: recompute of one overlaps the reverse of the other) */
  float *tape_g;            /* device [2][seg][2][S][S]: force windows of a segment                 */
  double *tape_body;        /* device [2][seg + 1][6]: body states of a segment                     */
  float *adj_a, *adj_b;     /* device [9][n]: adjoint fields                                       */
  float *grid;              /* device [2][S][S]: adjoint of the force window                        */
  double *adj_body;         /* device [8]: adjoint of (a, v, d) and the stiffness accumulator        */
  double *hist_d;           /* device [T][2]: displacement history of the forward run              */
  double *u_in_part;        /* device [T]: per-step inlet-velocity gradient terms                    */
  double *grad_delta;       /* device [M] output: dJ / d delta_m                                    */
  double *out;              /* device [3] output: J, dJ / d stiffness, dJ / d u_in                    */
  int32_t *sync;            /* device [4] scratch of the reverse pass (hand-off counters between its
                               bulk and IB kernels); sync[3] != 0 after the call: a bounded wait
                               timed out and the gradient is invalid                                */
} VsxAdj2D;

/* ------------------------------------------------------------------ library */
const char *vsx_version(void);
const char *vsx_last_error(void);
int vsx_device_sync(void);
/* Instrumentation (no counterpart in the reference): between start and stop the
 * fused drivers record a HIP event pair, on their stream, around every
 * VSX_PHASE_FULL bulk kernel (at most max_pairs); stop synchronises on the last
 * pair and returns the per-launch durations in milliseconds. */
int vsx_timer_start(int32_t max_pairs);
int vsx_timer_stop(float *ms, int32_t max, int32_t *n_out);

/* ------------------------------------------------------------------ D2Q9 / D3Q19 pure functions
 * n = number of cells (product of the spatial shape); arrays are [Q][n], [D][n], [n].
 * Suffix d2q9 <-> vivsim/lbm, d3q19 <-> vivsim/lbm3d. */

/* lbm/basic.py:60-85, lbm3d/basic.py:52-87 */
int vsx_d2q9_streaming(const float *f, float *out, int32_t nx, int32_t ny, void *stream);
int vsx_d3q19_streaming(const float *f, float *out, int32_t nx, int32_t ny, int32_t nz, void *stream);
/* lbm/basic.py:88-111, lbm3d/basic.py:90-106 */
int vsx_d2q9_get_macroscopic(const float *f, float *rho, float *u, int64_t n, void *stream);
int vsx_d3q19_get_macroscopic(const float *f, float *rho, float *u, int64_t n, void *stream);
/* lbm/basic.py:175-194, lbm3d/basic.py:155-167 : out = g * 0.5 / rho ; g [comps][n],
 * rho [n] or NULL -> rho_scalar */
int vsx_velocity_correction(const float *g, const float *rho, float rho_scalar, float *out, int32_t comps,
                            int64_t n, void *stream);
/* lbm/basic.py:114-135, lbm3d/basic.py:109-130 */
int vsx_d2q9_get_equilibrium(const float *rho, const float *u, float *feq, int64_t n, void *stream);
int vsx_d3q19_get_equilibrium(const float *rho, const float *u, float *feq, int64_t n, void *stream);
/* lbm/basic.py:138-156, lbm3d/basic.py:133-146 */
int vsx_d2q9_collision_bgk(const float *f, const float *feq, float omega, float *out, int64_t n, void *stream);
int vsx_d3q19_collision_bgk(const float *f, const float *feq, float omega, float *out, int64_t n, void *stream);
/* lbm/collision/mrt.py:66-88, lbm3d/collision/mrt.py:83-98 ; op = host float[Q*Q] */
int vsx_d2q9_collision_mrt(const float *f, const float *feq, const float *op, float *out, int64_t n, void *stream);
int vsx_d3q19_collision_mrt(const float *f, const float *feq, const float *op, float *out, int64_t n, void *stream);
/* lbm/collision/kbc.py:10-59, lbm3d/collision/kbc.py:8-42 */
int vsx_d2q9_collision_kbc(const float *f, const float *feq, float omega, float *out, int64_t n, void *stream);
int vsx_d3q19_collision_kbc(const float *f, const float *feq, float omega, float *out, int64_t n, void *stream);
/* lbm/collision/reg.py:4-49, lbm3d/collision/reg.py:45-62 */
int vsx_d2q9_collision_reg(const float *f, const float *feq, float omega, float *out, int64_t n, void *stream);
int vsx_d3q19_collision_reg(const float *f, const float *feq, float omega, float *out, int64_t n, void *stream);
/* lbm/forcing/guo.py:6-33, lbm3d/forcing/guo.py:13-38 */
int vsx_d2q9_get_guo_forcing_term(const float *g, const float *u, float *out, int64_t n, void *stream);
int vsx_d3q19_get_guo_forcing_term(const float *g, const float *u, float *out, int64_t n, void *stream);
/* lbm/forcing/guo.py:38-57, lbm3d/forcing/guo.py:41-56 */
int vsx_d2q9_forcing_guo_bgk(const float *f, const float *g, const float *u, float omega, float *out, int64_t n, void *stream);
int vsx_d3q19_forcing_guo_bgk(const float *f, const float *g, const float *u, float omega, float *out, int64_t n, void *stream);
/* lbm/forcing/guo.py:82-107, lbm3d/forcing/guo.py:71-87 ; op = host float[Q*Q] */
int vsx_d2q9_forcing_guo_mrt(const float *f, const float *g, const float *u, const float *op, float *out, int64_t n, void *stream);
int vsx_d3q19_forcing_guo_mrt(const float *f, const float *g, const float *u, const float *op, float *out, int64_t n, void *stream);
/* lbm/forcing/edm.py:4-31, lbm3d/forcing/edm.py:6-26 */
int vsx_d2q9_forcing_edm(const float *f, const float *g, const float *u, float *out, int64_t n, void *stream);
int vsx_d3q19_forcing_edm(const float *f, const float *g, const float *u, float *out, int64_t n, void *stream);
/* lbm/boundary/bb.py:98-110, lbm3d/boundary/bb.py:56-59 ; mask = uint8 [n] */
int vsx_d2q9_obstacle_bounce_back(const float *f, const uint8_t *mask, float *out, int64_t n, void *stream);
int vsx_d3q19_obstacle_bounce_back(const float *f, const uint8_t *mask, float *out, int64_t n, void *stream);

/* One wall of any kind/mode on a full field (out = f with the wall line rewritten;
 * out may alias f).  f_pre is the pre-streaming field (bounce-back / specular only).
 *   boundary_{nee,nebb,equilibrium}        lbm/boundary/{nee,nebb,eq}.py (+ 3D)
 *   boundary_{velocity,pressure,force_corrected}_*   boundary/_helpers.py:180-239
 *   boundary_bounce_back / specular_reflection        lbm/boundary/bb.py:15-95, lbm3d/boundary/bb.py:9-53 */
int vsx_d2q9_boundary(const VsxWall *wall, const float *f_pre, const float *f, float *out,
                      int32_t nx, int32_t ny, void *stream);
int vsx_d3q19_boundary(const VsxWall *wall, const float *f_pre, const float *f, float *out,
                       int32_t nx, int32_t ny, int32_t nz, void *stream);

/* lbm/boundary/cbc.py:14-53, lbm3d/boundary/cbc.py:16-51 : rho_next [N], u_next [D][N] */
int vsx_d2q9_boundary_characteristic(const float *rho, const float *u, int32_t loc, float *rho_next,
                                     float *u_next, int32_t nx, int32_t ny, void *stream);
int vsx_d3q19_boundary_characteristic(const float *rho, const float *u, int32_t loc, float *rho_next,
                                      float *u_next, int32_t nx, int32_t ny, int32_t nz, void *stream);

/* ------------------------------------------------------------------ fused steps
 * A user step "macroscopic -> [IB] -> collision -> forcing -> streaming -> walls"
 * (README.md:105-121; examples/2d/lid_driven_cavity.py:55-64;
 * examples/2d/poiseuille_channel.py:80-128) as one pull-scheme pass.
 * src / dst use the storage of p->storage, except that VSX_PHASE_COLLIDE reads
 * float32 when src_f32 != 0 and VSX_PHASE_STREAM writes float32 when dst_f32 != 0. */
int vsx_d2q9_step(const VsxLbm *p, const void *src, int32_t src_f32, void *dst, int32_t dst_f32,
                  int32_t phase, void *stream);
int vsx_d3q19_step(const VsxLbm *p, const void *src, int32_t src_f32, void *dst, int32_t dst_f32,
                   int32_t phase, void *stream);

/* n_steps reference steps (the jitted lax.scan chunk driver, e.g.
 * examples/2d/lid_driven_cavity.py:67-73): f_out = step^n(f_in).
 * work_a / work_b hold NX*NY(*NZ)*Q populations in p->storage; f_out may alias f_in. */
int vsx_d2q9_run(const VsxLbm *p, const float *f_in, float *f_out, void *work_a, void *work_b,
                 int32_t n_steps, void *stream);
int vsx_d3q19_run(const VsxLbm *p, const float *f_in, float *f_out, void *work_a, void *work_b,
                  int32_t n_steps, void *stream);

/* n_steps of the 2D IB-LBM VIV step (examples/2d/vortex_induced_vibration.py:96-157):
 * per step: macroscopic -> IB window -> stencil -> multi-direct forcing -> body
 * force (FP64) -> Newmark (FP64) -> collision + forcing -> streaming -> walls. */
int vsx_viv2d_run(const VsxLbm *p, const VsxIB2D *ib, const float *f_in, float *f_out,
                  void *work_a, void *work_b, int32_t n_steps, void *stream);

/* n_steps of the 3D IB-LBM step past a fixed body (examples/3d/flow_past_sphere.py:139-171):
 * per step: macroscopic -> IB window -> multi-direct forcing (ib/mdf.py:10-64 with the
 * precomputed stencil) -> total marker force (FP64) -> collision + forcing -> streaming
 * -> walls.  p must be a D3Q19 step. */
int vsx_sphere3d_run(const VsxLbm *p, const VsxIB3D *ib, const float *f_in, float *f_out,
                     void *work_a, void *work_b, int32_t n_steps, void *stream);
/* Size in floats of VsxIB3D.work. */
int64_t vsx_ib3d_work_floats(int32_t n_markers, int64_t cells);
/* Transpose of a stencil (I [m][s] flat indices into a grid of n_grid cells):
 * cell_ptr [n_grid + 1], cell_slot [m * s] list, for each cell in ascending
 * order, the slots m * s_len + j whose index lands on it under the scatter
 * rule of ib/stencil.py:106-110 (out-of-range entries dropped).  tmp: device
 * scratch of n_grid + 1 int32. */
int vsx_ib3d_transpose_stencil(const int32_t *I, int32_t m, int32_t s, int64_t n_grid, int32_t *cell_ptr,
                               int32_t *cell_slot, int32_t *tmp, void *stream);

/* Batched VIV sweep (no counterpart in the reference, which runs one case per
 * jitted program; each case is the step of vsx_viv2d_run).  prepare validates
 * the cases, orders the bulk tiles (tiles in the x-band of each case's IB window
 * last) and uploads the table (a synchronous copy: call it outside graph
 * capture); run advances every case n_steps; graph captures run into a HIP
 * graph on a private stream and returns its executable (replay with
 * vsx_graph_launch: every replay continues from the batch clock). */
int64_t vsx_viv2d_batch_table_bytes(const VsxVivCase *cases, int32_t n_cases);
int vsx_viv2d_batch_prepare(const VsxVivCase *cases, int32_t n_cases, void *table, int32_t *clock,
                            VsxVivBatch *out);
int vsx_viv2d_batch_run(const VsxVivBatch *b, const float *f_in, float *f_out, void *work_a, void *work_b,
                        int32_t n_steps, void *stream);
int vsx_viv2d_batch_graph(const VsxVivBatch *b, const float *f_in, float *f_out, void *work_a, void *work_b,
                          int32_t n_steps, void **graph_exec);
int vsx_graph_launch(void *graph_exec, void *stream);
int vsx_graph_destroy(void *graph_exec);

/* Gradient of the VIV objective (see VsxAdj2D) for the run f0 -> step^T(f0) that
 * vsx_viv2d_run would perform with p and ib (ib->step0 and ib->fused are set by the
 * driver; ib->body slot 0 holds the initial body state).  work_a / work_b as for
 * vsx_viv2d_run; f_tmp: device [9][n] float32 scratch. */
int vsx_viv2d_adjoint(const VsxLbm *p, const VsxIB2D *ib, const VsxAdj2D *adj, const float *f0, void *work_a,
                      void *work_b, float *f_tmp, void *stream);

/* ------------------------------------------------------------------ immersed boundary */

/* ib/kernels.py:4-61 : out[i] = phi(r[i]) */
int vsx_ib_kernel(int32_t kernel, const float *r, float *out, int64_t n, void *stream);
/* ib/stencil.py:7-51 : W, I [m][(2 radius)^2] */
int vsx_ib_get_ib_stencil2d(const float *marker_x, const float *marker_y, int32_t m, int32_t ny,
                            int32_t kernel, int32_t radius, float *W, int32_t *I, void *stream);
/* ib3d/stencil.py:7-59 : coords [m][3]; W, I [m][(2 radius)^3] */
int vsx_ib_get_ib_stencil3d(const float *coords, int32_t m, int32_t nx, int32_t ny, int32_t nz,
                            int32_t kernel, int32_t radius, float *W, int32_t *I, void *stream);
/* ib/stencil.py:54-78 : out[m][c] = sum_s W[m][s] grid[c][I[m][s]]  (grid [c][n_grid]) */
int vsx_ib_interpolate(const float *grid, int32_t c, int64_t n_grid, const float *W, const int32_t *I,
                       int32_t m, int32_t s, float *out, void *stream);
/* ib/stencil.py:81-110 : out = grid_in + scatter(marker[m][c] W[m][s] -> I[m][s]); out may alias grid_in */
int vsx_ib_spread(const float *marker, const float *grid_in, float *out, int32_t c, int64_t n_grid,
                  const float *W, const int32_t *I, int32_t m, int32_t s, void *stream);
/* ib/mdf.py:10-64 : grid_force [c][n_grid], marker_force [m][c]; ds_arr NULL -> scalar ds.
 * work: device scratch of 3*m*c floats. */
int vsx_ib_multi_direct_forcing(const float *grid_u, int32_t c, int64_t n_grid, const float *W,
                                const int32_t *I, int32_t m, int32_t s, const float *u_target,
                                const float *ds_arr, float ds, int32_t n_iter, float *grid_force,
                                float *marker_force, float *work, void *stream);
/* No counterpart in the reference: vsx_ib_multi_direct_forcing without the work
 * argument and without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops).  Its scratch (ub, F, dF)
 * is a fixed per-device array of the library sized for m * c <= 65536 marker
 * components (768 KiB, held until the process ends); larger inputs return
 * VSX_ERR_VALUE.  Every call on one device shares that array, so calls that
 * overlap on one device (concurrent streams or threads) race. */
int vsx_ib_multi_direct_forcing_noalloc(const float *grid_u, int32_t c, int64_t n_grid, const float *W,
                                        const int32_t *I, int32_t m, int32_t s, const float *u_target,
                                        const float *ds_arr, float ds, int32_t n_iter, float *grid_force,
                                        float *marker_force, void *stream);
/* ib/geometry.py:6-43 : coords [m][2] */
int vsx_ib_get_area(const float *coords, int32_t m, float *out, void *stream);
int vsx_ib_get_ds(const float *coords, int32_t m, int32_t closed, float *out, void *stream);
/* ib3d/geometry.py:6-67 : verts [nv][3], faces [nf][3] int32 (vertex indices: gathers clamp,
 * the ds scatter-add drops out-of-range vertices, as JAX indexing does) */
int vsx_ib3d_get_triangle_areas(const float *verts, int32_t nv, const int32_t *faces, int32_t nf, float *out,
                                void *stream);
int vsx_ib3d_get_surface_area(const float *verts, int32_t nv, const int32_t *faces, int32_t nf, float *out,
                              void *stream);
int vsx_ib3d_get_volume(const float *verts, int32_t nv, const int32_t *faces, int32_t nf, float *out, void *stream);
int vsx_ib3d_get_ds(const float *verts, int32_t nv, const int32_t *faces, int32_t nf, float *out, void *stream);

/* ------------------------------------------------------------------ diagnostics (post.py)
 * u is a velocity field [D][NX][NY] (D = 2, nz = 1) or [D][NX][NY][NZ] (D = 3); derivatives are
 * jnp.gradient's (unit spacing; central inside, one-sided first order at the ends). */
enum { VSX_POST_GRADIENT = 0,                  /* post.py:18-31  out [D][D][n]: du_i / dx_j        */
       VSX_POST_VORTICITY = 1,                 /* post.py:34-57  out [n] (2D) or [3][n] (3D)       */
       VSX_POST_VORTICITY_MAGNITUDE = 2,       /* post.py:60-72                                    */
       VSX_POST_DIVERGENCE = 3,                /* post.py:76-87                                    */
       VSX_POST_STRAIN_RATE = 4,               /* post.py:90-99  out [D][D][n]                     */
       VSX_POST_STRAIN_RATE_MAGNITUDE = 5,     /* post.py:102-111                                  */
       VSX_POST_Q_CRITERION = 6,               /* post.py:176-191                                  */
       VSX_POST_ENSTROPHY = 7,                 /* post.py:148-159                                  */
       VSX_POST_VELOCITY_MAGNITUDE = 8,        /* post.py:6-15                                     */
       VSX_POST_KINETIC_ENERGY = 9 };          /* post.py:114-122                                  */
int vsx_post_field(const float *u, int32_t dim, int32_t nx, int32_t ny, int32_t nz, int32_t op, float *out,
                   void *stream);
/* post.py:125-133 (op = KINETIC_ENERGY) and 162-170 (op = ENSTROPHY): out[0] = mean over the
 * cells, summed in FP64 in a fixed order; partial: device scratch of 1024 doubles */
int vsx_post_mean(const float *u, int32_t dim, int32_t nx, int32_t ny, int32_t nz, int32_t op, double *partial,
                  float *out, void *stream);
/* post.py:136-145 */
int vsx_post_pressure(const float *rho, int64_t n, float cs2, float *out, void *stream);
/* examples/2d/vortex_induced_vibration.py:228-229 : cd[t], cl[t] = 2 h[t][0..1] / (D U0^2) */
int vsx_force_coefficients(const double *h, int32_t n, double D, double U0, double *cd, double *cl, void *stream);

/* ------------------------------------------------------------------ grid refinement (multigrid.py)
 * dir: 0 left, 1 right, 2 up, 3 down; the fine grid has twice the coarse nodes along the
 * interface.  Written in place into the destination field. */
/* multigrid.py:58-98 */
int vsx_mg_fine_to_coarse(const float *f_fine, int32_t nx_fine, int32_t ny_fine, float *f_coarse,
                          int32_t nx_coarse, int32_t ny_coarse, int32_t dir, void *stream);
/* multigrid.py:100-130 */
int vsx_mg_coarse_to_fine(const float *f_coarse, int32_t nx_coarse, int32_t ny_coarse, float *f_fine,
                          int32_t nx_fine, int32_t ny_fine, int32_t dir, void *stream);

/* ------------------------------------------------------------------ structure (dyn.py), FP64 */
/* dyn.py:5-51 : a, v, d, h [ndof]; m, k, c scalars (matrix == 0) or [ndof][ndof] (matrix != 0) */
int vsx_dyn_newmark(const double *a, const double *v, const double *d, const double *h, int32_t ndof,
                    const double *m, const double *k, const double *c, int32_t matrix, double dt,
                    double gamma, double beta, double *a_out, double *v_out, double *d_out, void *stream);
/* dyn.py:126-136 : out[c] = sum_m h[m][c] (FP64 accumulation of float32 forces) */
int vsx_dyn_get_force_to_obj(const float *h, int32_t m, int32_t c, double *out, void *stream);
/* dyn.py:139-154 */
int vsx_dyn_get_torque_to_obj(const float *x, const float *y, double xc, double yc, const double *d,
                              const float *h, int32_t m, double *out, void *stream);
/* dyn.py:68-100 : 2-DOF translation (rotate == 0) or 3-DOF rigid motion (rotate != 0) */
int vsx_dyn_markers_coords(const float *x0, const float *y0, int32_t m, double xc, double yc,
                           const double *d, int32_t rotate, float *x, float *y, void *stream);
/* dyn.py:103-120 : v_markers [m][2] */
int vsx_dyn_markers_velocity_3dof(const float *x, const float *y, int32_t m, double xc, double yc,
                                  const double *d, const double *v, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* VIVSIM_AMD_H */
