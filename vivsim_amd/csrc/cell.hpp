// Per-cell LBM arithmetic shared by the standalone operator kernels and the
// fused step kernels.  Everything is FP32 and lives in registers; the lattice
// tables are constexpr, so the fully unrolled loops fold the +-1/+-2 direction
// components into adds and the zero entries away.
#pragma once
#include "lattice.hpp"

namespace vsx {

// ---------------------------------------------------------------- macroscopic
// rho = sum f, u = sum c f / rho (lbm/basic.py:88-111, lbm3d/basic.py:90-106)
template <class L>
__device__ __forceinline__ void macroscopic(const float (&f)[L::Q], float &rho, float (&u)[3]) {
  float r = 0.f, jp[3] = {0.f, 0.f, 0.f}, jn[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    r += f[i];
#pragma unroll
    for (int a = 0; a < L::D; ++a) {
      const int c = cax<L>(i, a);
      if (c > 0) jp[a] += f[i];
      if (c < 0) jn[a] += f[i];
    }
  }
  rho = r;
#pragma unroll
  for (int a = 0; a < 3; ++a) u[a] = (a < L::D) ? (jp[a] - jn[a]) / r : 0.f;
}

// ---------------------------------------------------------------- equilibrium
// feq_i = rho w_i (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 |u|^2)
// (lbm/basic.py:114-135; the D3Q19 division form of lbm3d/basic.py:109-130 is the same polynomial)
template <class L>
__device__ __forceinline__ float cdot(int i, const float (&v)[3]) {
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < L::D; ++a) {
    const int c = cax<L>(i, a);
    if (c == 1) s += v[a];
    if (c == -1) s -= v[a];
  }
  return s;
}

template <class L>
__device__ __forceinline__ float usq(const float (&u)[3]) {
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < L::D; ++a) s += u[a] * u[a];
  return s;
}

template <class L>
__device__ __forceinline__ float feq_i(int i, float rho, const float (&u)[3], float u2) {
  const float uc = cdot<L>(i, u);
  return (rho * L::w[i]) * (1.f + 3.f * uc + 4.5f * uc * uc - 1.5f * u2);
}

template <class L>
__device__ __forceinline__ void equilibrium(float rho, const float (&u)[3], float (&feq)[L::Q]) {
  const float u2 = usq<L>(u);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) feq[i] = feq_i<L>(i, rho, u, u2);
}

// ---------------------------------------------------------------- Guo forcing term
// F_i = w_i (3 (c_i.g - u.g) + 9 (c_i.u)(c_i.g))   (lbm/forcing/guo.py:6-33, lbm3d/forcing/guo.py:13-38)
template <class L>
__device__ __forceinline__ float guo_i(int i, const float (&u)[3], const float (&g)[3], float ug) {
  const float cu = cdot<L>(i, u), cg = cdot<L>(i, g);
  return L::w[i] * (3.f * (cg - ug) + 9.f * cg * cu);
}

template <class L>
__device__ __forceinline__ float dot3(const float (&a)[3], const float (&b)[3]) {
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < L::D; ++k) s += a[k] * b[k];
  return s;
}

// ---------------------------------------------------------------- second-order projections
// D3Q19: s_q = w_q / (2 cs^4) H(c_q) : Pi with Pi = sum_r c_r c_r fneq_r and
// H(c) = c c - cs^2 I  — the structured form of the 19x19 table of
// lbm3d/collision/reg.py:10-33 (6 stress sums instead of a dense matvec).
template <class L>
__device__ __forceinline__ void project2_3d(const float (&fn)[L::Q], float (&s)[L::Q]) {
  float P[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
  for (int r = 0; r < L::Q; ++r) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) {
        const int cc = cax<L>(r, a) * cax<L>(r, b);
        if (cc == 1) P[a][b] += fn[r];
        if (cc == -1) P[a][b] -= fn[r];
      }
  }
  const float tr = (P[0][0] + P[1][1] + P[2][2]) * (1.f / 3.f);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int cc = cax<L>(q, a) * cax<L>(q, b);
        const float pab = a <= b ? P[a][b] : P[b][a];
        if (cc == 1) acc += pab;
        if (cc == -1) acc -= pab;
      }
    s[q] = (4.5f * L::w[q]) * (acc - tr);
  }
}

// ---------------------------------------------------------------- collisions (in place)
// 1 / (feq_i + 1e-20) of the entropic inner products: the hardware reciprocal
// (v_rcp_f32, 1 ulp), one per direction, in place of two IEEE divisions per
// direction; the inner products only scale the stabilizer, so the ulp does
// not show beyond float32 rounding, and the divisions had made the KBC step
// bound by its arithmetic rather than by HBM
__device__ __forceinline__ float kbc_inv_feq(float feq) { return __builtin_amdgcn_rcpf(feq + 1e-20f); }

__device__ __forceinline__ float kbc_half_gamma(float omega, float ish, float ihh) {
  return 1.0f / omega - (1.0f - 1.0f / omega) * ish / (ihh + 1e-20f);
}

// D2Q9 KBC: shear part from N = Pxx - Pyy and Pxy (lbm/collision/kbc.py:10-59)
__device__ __forceinline__ void kbc_d2q9(float (&f)[9], const float (&feq)[9], float omega) {
  float fn[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) fn[i] = f[i] - feq[i];
  const float nsd = fn[1] - fn[2] + fn[3] - fn[4];
  const float ss = fn[5] - fn[6] + fn[7] - fn[8];
  const float a = nsd * 0.25f, b = ss * 0.25f;
  const float s[9] = {0.f, a, -a, a, -a, b, -b, b, -b};
  float ish = 0.f, ihh = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float h = fn[i] - s[i];
    const float inv = kbc_inv_feq(feq[i]);
    ish += h * s[i] * inv;
    ihh += h * h * inv;
  }
  const float hg = kbc_half_gamma(omega, ish, ihh);
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] -= omega * (s[i] + hg * (fn[i] - s[i]));
}

// D3Q19 KBC: s = P fneq (lbm3d/collision/kbc.py:8-42)
template <class L>
__device__ __forceinline__ void kbc_d3q19(float (&f)[L::Q], const float (&feq)[L::Q], float omega) {
  float fn[L::Q], s[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) fn[i] = f[i] - feq[i];
  project2_3d<L>(fn, s);
  float ish = 0.f, ihh = 0.f;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const float h = fn[i] - s[i];
    const float inv = kbc_inv_feq(feq[i]);
    ish += h * s[i] * inv;
    ihh += h * h * inv;
  }
  const float hg = kbc_half_gamma(omega, ish, ihh);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) f[i] -= omega * (s[i] + hg * (fn[i] - s[i]));
}

// D2Q9 regularized: closed-form Hermite-2 projection (lbm/collision/reg.py:4-49)
__device__ __forceinline__ void reg_d2q9(float (&f)[9], const float (&feq)[9], float omega) {
  float fn[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) fn[i] = f[i] - feq[i];
  const float ds = fn[5] + fn[6] + fn[7] + fn[8];
  const float pxx = fn[1] + fn[3] + ds, pyy = fn[2] + fn[4] + ds;
  const float pxy = fn[5] - fn[6] + fn[7] - fn[8];
  const float tr = (pxx + pyy) / 12.0f;
  const float ax = (2.0f * pxx - pyy) / 6.0f, ay = (2.0f * pyy - pxx) / 6.0f;
  const float dp = tr + pxy / 4.0f, dn = tr - pxy / 4.0f;
  const float reg[9] = {-8.0f * tr, ax, ay, ax, ay, dp, dn, dp, dn};
  const float k = 1.0f - omega;
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = feq[i] + k * reg[i];
}

template <class L>
__device__ __forceinline__ void reg_d3q19(float (&f)[L::Q], const float (&feq)[L::Q], float omega) {
  float fn[L::Q], s[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) fn[i] = f[i] - feq[i];
  project2_3d<L>(fn, s);
  const float k = 1.0f - omega;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) f[i] = feq[i] + k * s[i];
}

// ---------------------------------------------------------------- D2Q9 MRT in moment space
// Lallemand-Luo basis of lbm/collision/mrt.py:10-22.  Its rows are orthogonal,
// so M^-1 = M^T diag(1 / |row_k|^2): the reference's dense f + M^-1 S M (feq - f)
// (mrt.py:88) and f + M^-1 (I - S/2) M F (forcing/guo.py:107) become one sparse
// forward transform, a diagonal scaling and one sparse back transform.
struct Mrt2 {
  static constexpr int M[9][9] = {
      {1, 1, 1, 1, 1, 1, 1, 1, 1},      {-4, -1, -1, -1, -1, 2, 2, 2, 2}, {4, -2, -2, -2, -2, 1, 1, 1, 1},
      {0, 1, 0, -1, 0, 1, -1, -1, 1},   {0, -2, 0, 2, 0, 1, -1, -1, 1},   {0, 0, 1, 0, -1, 1, 1, -1, -1},
      {0, 0, -2, 0, 2, 1, 1, -1, -1},   {0, 1, -1, 1, -1, 0, 0, 0, 0},    {0, 0, 0, 0, 0, 1, -1, 1, -1}};
  static constexpr float inv_norm[9] = {1.f / 9, 1.f / 36, 1.f / 36, 1.f / 6, 1.f / 12,
                                        1.f / 6, 1.f / 12, 1.f / 4, 1.f / 4};
};

__device__ __forceinline__ float mrow(int k, const float (&x)[9]) {
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int m = Mrt2::M[k][i];
    if (m == 1) s += x[i];
    if (m == -1) s -= x[i];
    if (m == 2) s += 2.f * x[i];
    if (m == -2) s -= 2.f * x[i];
    if (m == 4) s += 4.f * x[i];
    if (m == -4) s -= 4.f * x[i];
  }
  return s;
}

// f += M^-1 [ S M (feq - f) + (I - S/2) M F ]   (F may be null -> collision only)
template <bool HAS_F>
__device__ __forceinline__ void mrt_d2q9(float (&f)[9], const float (&feq)[9], const float (&F)[9],
                                         const float (&rates)[9]) {
  float d[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = feq[i] - f[i];
  float a[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    float r = rates[k] * mrow(k, d);
    if (HAS_F) r += (1.f - 0.5f * rates[k]) * mrow(k, F);
    a[k] = r * Mrt2::inv_norm[k];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int m = Mrt2::M[k][i];
      if (m == 1) s += a[k];
      if (m == -1) s -= a[k];
      if (m == 2) s += 2.f * a[k];
      if (m == -2) s -= 2.f * a[k];
      if (m == 4) s += 4.f * a[k];
      if (m == -4) s -= 4.f * a[k];
    }
    f[i] += s;
  }
}


// Unforced D2Q9 MRT in moment space: f += M^-1 S (meq - M f), with the
// equilibrium moments of this basis in closed form,
//   meq = (rho, rho (3|u|^2 - 2), rho (1 - 3|u|^2), rho ux, -rho ux, rho uy,
//          -rho uy, rho (ux^2 - uy^2), rho ux uy)  (= M feq(rho, u) exactly),
// and the conserved rows (rho, jx, jy: meq - m = 0) left out.  Equal to
// mrt_d2q9<false>(f, feq(rho, u), ...) up to float32 rounding.
__device__ __forceinline__ void mrt_d2q9_unforced(float (&f)[9], float rho, const float (&u)[3],
                                                  const float (&rates)[9]) {
  const float ux = u[0], uy = u[1], u2 = ux * ux + uy * uy;
  const float jx = rho * ux, jy = rho * uy;
  const float meq[9] = {rho, rho * (3.f * u2 - 2.f), rho * (1.f - 3.f * u2), jx, -jx, jy, -jy,
                        rho * (ux * ux - uy * uy), jx * uy};
  float a[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (k == 0 || k == 3 || k == 5) {
      a[k] = 0.f;
      continue;
    }
    a[k] = rates[k] * (meq[k] - mrow(k, f)) * Mrt2::inv_norm[k];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      if (k == 0 || k == 3 || k == 5) continue;
      const int m = Mrt2::M[k][i];
      if (m == 1) s += a[k];
      if (m == -1) s -= a[k];
      if (m == 2) s += 2.f * a[k];
      if (m == -2) s -= 2.f * a[k];
      if (m == 4) s += 4.f * a[k];
      if (m == -4) s -= 4.f * a[k];
    }
    f[i] += s;
  }
}

// ---------------------------------------------------------------- D3Q19 MRT in moment space
// Raw-moment basis of lbm3d/collision/mrt.py:7-30 (rows: rho; j; |c|^2;
// 3cx^2-|c|^2; cy^2-cz^2; cxcy, cxcz, cycz; six third-order; three fourth-order).
// Its inverse is exactly N72 / 72 with the integer matrix N72 below (M N72 = 72 I,
// checked on the host at library load), so the reference's dense
// f + M^-1 S M (feq - f) (mrt.py:96-98) and f + M^-1 (I - S/2) M F
// (forcing/guo.py:85-87) become sparse integer transforms.
struct Mrt3 {
  static constexpr int M[19][19] = {{1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
        {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
        {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
        {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
        {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2},
        {0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
        {0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1},
        {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1},
        {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1}};
  static constexpr int N72[19][19] = {{72, 0, 0, 0, -72, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 72, 72, 72},
        {0, 36, 0, 0, 12, 12, 0, 0, 0, 0, 0, 0, -36, 0, -36, 0, -36, -36, 0},
        {0, -36, 0, 0, 12, 12, 0, 0, 0, 0, 0, 0, 36, 0, 36, 0, -36, -36, 0},
        {0, 0, 36, 0, 12, -6, 18, 0, 0, 0, -36, 0, 0, 0, 0, -36, -36, 0, -36},
        {0, 0, -36, 0, 12, -6, 18, 0, 0, 0, 36, 0, 0, 0, 0, 36, -36, 0, -36},
        {0, 0, 0, 36, 12, -6, -18, 0, 0, 0, 0, -36, 0, -36, 0, 0, 0, -36, -36},
        {0, 0, 0, -36, 12, -6, -18, 0, 0, 0, 0, 36, 0, 36, 0, 0, 0, -36, -36},
        {0, 0, 0, 0, 0, 0, 0, 18, 0, 0, 18, 0, 18, 0, 0, 0, 18, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, -18, 0, 0, 18, 0, -18, 0, 0, 0, 18, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, -18, 0, 0, -18, 0, 18, 0, 0, 0, 18, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 18, 0, 0, -18, 0, -18, 0, 0, 0, 18, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 18, 0, 0, 18, 0, 0, 18, 0, 0, 18, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, -18, 0, 0, 18, 0, 0, -18, 0, 0, 18, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, -18, 0, 0, -18, 0, 0, 18, 0, 0, 18, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 18, 0, 0, -18, 0, 0, -18, 0, 0, 18, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 18, 0, 0, 0, 18, 0, 18, 0, 0, 18},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, -18, 0, 0, 0, 18, 0, -18, 0, 0, 18},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, -18, 0, 0, 0, -18, 0, 18, 0, 0, 18},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 18, 0, 0, 0, -18, 0, -18, 0, 0, 18}};
  
};

template <int K>
__device__ __forceinline__ float int_row(const int (&row)[K], const float (&x)[K]) {
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int m = row[i];
    if (m == 1) s += x[i];
    else if (m == -1) s -= x[i];
    else if (m != 0) s += (float)m * x[i];
  }
  return s;
}

// f += M^-1 [ S M (feq - f) + (I - S/2) M F ]  (F == nullptr -> collision only)
__device__ __forceinline__ void mrt_d3q19(float (&f)[19], const float (&feq)[19], const float (&F)[19], bool has_f,
                                          const float *rates) {
  float d[19], r[19];
#pragma unroll
  for (int i = 0; i < 19; ++i) d[i] = feq[i] - f[i];
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    float v = rates[k] * int_row<19>(Mrt3::M[k], d);
    if (has_f) v += (1.f - 0.5f * rates[k]) * int_row<19>(Mrt3::M[k], F);
    r[k] = v;
  }
#pragma unroll
  for (int i = 0; i < 19; ++i) f[i] += int_row<19>(Mrt3::N72[i], r) * (1.f / 72.f);
}

}  // namespace vsx
