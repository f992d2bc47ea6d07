"""Oracle restatement of ``vivsim.multigrid`` (block grid-refinement transfers
between levels that differ by a factor 2).  Test infrastructure only: the
product never imports it."""

import numpy as np

RIGHT_DIRS = [1, 5, 8]     # lbm/lattice.py:58-61 (via lbm/basic.py:36-39)
LEFT_DIRS = [3, 7, 6]
UP_DIRS = [2, 5, 6]
DOWN_DIRS = [4, 7, 8]


def init_grid(width, height, level=0, buffer_x=0, buffer_y=0):
    """(f, rho, u) of a block of width x height at refinement `level` plus buffer cells — multigrid.py:21-55."""
    nx = int(width * 2 ** level) + buffer_x
    ny = int(height * 2 ** level) + buffer_y
    return np.zeros((9, nx, ny), np.float32), np.ones((nx, ny), np.float32), np.zeros((2, nx, ny), np.float32)


def fine_to_coarse(f_fine, f_coarse, dir):
    """Mean of the 2 x 2 fine cells next to the interface into the coarse buffer line — multigrid.py:58-98."""
    out = f_coarse.copy()
    if dir == "left":
        d = LEFT_DIRS
        out[d, -1] = 0.25 * (f_fine[d, 0, 0::2] + f_fine[d, 0, 1::2] + f_fine[d, 1, 0::2] + f_fine[d, 1, 1::2])
    if dir == "right":
        d = RIGHT_DIRS
        out[d, 0] = 0.25 * (f_fine[d, -1, 0::2] + f_fine[d, -1, 1::2] + f_fine[d, -2, 0::2] + f_fine[d, -2, 1::2])
    if dir == "up":
        d = UP_DIRS
        out[d, :, 0] = 0.25 * (f_fine[d, 0::2, -1] + f_fine[d, 1::2, -1] + f_fine[d, 0::2, -2] + f_fine[d, 1::2, -2])
    if dir == "down":
        d = DOWN_DIRS
        out[d, :, -1] = 0.25 * (f_fine[d, 0::2, 0] + f_fine[d, 1::2, 0] + f_fine[d, 0::2, 1] + f_fine[d, 1::2, 1])
    return out


def coarse_to_fine(f_coarse, f_fine, dir):
    """Each coarse interface value copied to the two fine cells it covers — multigrid.py:100-130."""
    out = f_fine.copy()
    if dir == "left":
        d = LEFT_DIRS
        out[d, -1, 0::2] = f_coarse[d, 0]
        out[d, -1, 1::2] = f_coarse[d, 0]
    if dir == "right":
        d = RIGHT_DIRS
        out[d, 0, 0::2] = f_coarse[d, -1]
        out[d, 0, 1::2] = f_coarse[d, -1]
    if dir == "up":
        d = UP_DIRS
        out[d, 0::2, 0] = f_coarse[d, :, -1]
        out[d, 1::2, 0] = f_coarse[d, :, -1]
    if dir == "down":
        d = DOWN_DIRS
        out[d, 0::2, -1] = f_coarse[d, :, 0]
        out[d, 1::2, -1] = f_coarse[d, :, 0]
    return out


def get_omega(nu, level=0):
    """Level-scaled relaxation rate — multigrid.py:133-147."""
    om0 = 1 / (3 * nu + 0.5)
    return 2 * om0 / (2 ** (level + 1) + (1 - 2 ** level) * om0)


def coord_to_indices(x, y, grid_start_x, grid_start_y, level=0):
    """multigrid.py:150-171."""
    return int((x - grid_start_x) * 2 ** level), int((y - grid_start_y) * 2 ** level)
