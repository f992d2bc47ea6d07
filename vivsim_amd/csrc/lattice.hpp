// Lattice descriptors for the HIP kernels (D2Q9, D3Q19).
//
// The direction numbering is the data layout of the reference's population
// arrays f[9][NX][NY] and f[19][NX][NY][NZ] and must match it exactly:
//   D2Q9  : vivsim/lbm/lattice.py:39-110
//   D3Q19 : vivsim/lbm3d/lattice.py:39-126
// Wall groups (in/out/tangential/zero directions per face) follow the same
// files; the reference pairs in_dirs[k] with out_dirs[k] by position in its
// Zou/He and bounce-back formulas, so their order is kept as well.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace vsx {

enum Loc : int32_t { LOC_LEFT = 0, LOC_RIGHT = 1, LOC_BOTTOM = 2, LOC_TOP = 3, LOC_BACK = 4, LOC_FRONT = 5 };

struct D2Q9 {
  static constexpr int Q = 9;
  static constexpr int D = 2;
  static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr float w[9] = {4.f / 9, 1.f / 9, 1.f / 9, 1.f / 9, 1.f / 9,
                                 1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36};
  // per wall (left, right, bottom, top): axis, inward sign, in/out dirs, tangential dirs, +/- side dirs
  static constexpr int n_walls = 4;
  static constexpr int n_in = 3;
  static constexpr int w_axis[4] = {0, 0, 1, 1};
  static constexpr int w_sign[4] = {1, -1, 1, -1};
  static constexpr int w_in[4][3] = {{1, 5, 8}, {3, 7, 6}, {2, 5, 6}, {4, 7, 8}};
  static constexpr int w_out[4][3] = {{3, 7, 6}, {1, 5, 8}, {4, 7, 8}, {2, 5, 6}};
  static constexpr int w_tan[4][2] = {{2, 4}, {2, 4}, {1, 3}, {1, 3}};
  static constexpr int w_pos[4][3] = {{2, 5, 6}, {2, 5, 6}, {1, 5, 8}, {1, 5, 8}};
  static constexpr int w_neg[4][3] = {{4, 7, 8}, {4, 7, 8}, {3, 7, 6}, {3, 7, 6}};
};

struct D3Q19 {
  static constexpr int Q = 19;
  static constexpr int D = 3;
  static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr float w[19] = {1.f / 3,
                                  1.f / 18, 1.f / 18, 1.f / 18, 1.f / 18, 1.f / 18, 1.f / 18,
                                  1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36,
                                  1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36};
  // faces: left, right, bottom, top, back, front
  static constexpr int n_walls = 6;
  static constexpr int n_in = 5;
  static constexpr int w_axis[6] = {0, 0, 1, 1, 2, 2};
  static constexpr int w_sign[6] = {1, -1, 1, -1, 1, -1};
  static constexpr int w_in[6][5] = {{1, 7, 9, 11, 13}, {2, 8, 10, 12, 14}, {3, 7, 8, 15, 17},
                                     {4, 9, 10, 16, 18}, {5, 11, 12, 15, 16}, {6, 13, 14, 17, 18}};
  static constexpr int w_out[6][5] = {{2, 8, 10, 12, 14}, {1, 7, 9, 11, 13}, {4, 9, 10, 16, 18},
                                      {3, 7, 8, 15, 17}, {6, 13, 14, 17, 18}, {5, 11, 12, 15, 16}};
  static constexpr int w_tan_axes[6][2] = {{1, 2}, {1, 2}, {0, 2}, {0, 2}, {0, 1}, {0, 1}};
};

// c_i along an axis (0, 1, 2)
template <class L>
__host__ __device__ constexpr int cax(int i, int a) {
  return a == 0 ? L::cx[i] : (a == 1 ? L::cy[i] : L::cz[i]);
}

}  // namespace vsx
