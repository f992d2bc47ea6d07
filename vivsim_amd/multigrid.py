"""Block grid-refinement helpers with the names and signatures of
``vivsim.multigrid`` (vivsim/multigrid.py:21-171).  The two line transfers run
as HIP kernels (``vsx_mg_*``); the level-scaled relaxation rate and the
index conversion are host arithmetic, as in the reference."""

import torch

from . import _lib as L
from ._util import f32, require

_DIRS = {"left": 0, "right": 1, "up": 2, "down": 3}


def init_grid(width, height, level=0, buffer_x=0, buffer_y=0):
    """(f, rho, u) of a width x height block at refinement `level` plus buffer cells — multigrid.py:21-55."""
    nx = int(width * 2 ** level) + buffer_x
    ny = int(height * 2 ** level) + buffer_y
    dev = L.device()
    return (torch.zeros((9, nx, ny), dtype=torch.float32, device=dev),
            torch.ones((nx, ny), dtype=torch.float32, device=dev),
            torch.zeros((2, nx, ny), dtype=torch.float32, device=dev))


def fine_to_coarse(f_fine, f_coarse, dir):
    """The coarse buffer line of the interface in direction `dir` set to the mean of the
    2 x 2 fine cells next to it — multigrid.py:58-98."""
    src, dst = f32(f_fine), f32(f_coarse)
    require(src.ndim == 3 and src.shape[0] == 9 and dst.ndim == 3 and dst.shape[0] == 9,
            "f_fine and f_coarse must have shape (9, NX, NY)")
    out = dst.clone()
    if dir not in _DIRS:          # the reference matches none of its branches and returns f_coarse
        return out
    L.call("vsx_mg_fine_to_coarse", L.ptr(src), src.shape[1], src.shape[2], L.ptr(out), out.shape[1], out.shape[2],
           _DIRS[dir])
    return out


def coarse_to_fine(f_coarse, f_fine, dir):
    """The fine buffer line of the interface in direction `dir` set from the coarse
    interface line, each coarse value on the two fine cells it covers — multigrid.py:100-130."""
    src, dst = f32(f_coarse), f32(f_fine)
    require(src.ndim == 3 and src.shape[0] == 9 and dst.ndim == 3 and dst.shape[0] == 9,
            "f_fine and f_coarse must have shape (9, NX, NY)")
    out = dst.clone()
    if dir not in _DIRS:          # the reference matches none of its branches and returns f_fine
        return out
    L.call("vsx_mg_coarse_to_fine", L.ptr(src), src.shape[1], src.shape[2], L.ptr(out), out.shape[1], out.shape[2],
           _DIRS[dir])
    return out


def get_omega(nu, level=0):
    """Relaxation rate at refinement `level` (equal to 1 / (3 nu 2^level + 0.5)) — multigrid.py:133-147."""
    omega_l0 = 1 / (3 * nu + 0.5)
    return 2 * omega_l0 / (2 ** (level + 1) + (1 - 2 ** level) * omega_l0)


def coord_to_indices(x, y, grid_start_x, grid_start_y, level=0):
    """Local indices of a global coordinate in a block at `level` — multigrid.py:150-171."""
    return int((x - grid_start_x) * 2 ** level), int((y - grid_start_y) * 2 ** level)


__all__ = ["init_grid", "fine_to_coarse", "coarse_to_fine", "get_omega", "coord_to_indices"]
