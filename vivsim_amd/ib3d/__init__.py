"""3D immersed-boundary operators with the signatures of ``vivsim.ib3d``
(vivsim/ib3d/__init__.py:1-12): 3D stencil and mesh geometry, plus the
dimension-agnostic kernels / interpolate / spread / MDF of ``vivsim_amd.ib``."""

import torch

from .. import _lib as L
from .._util import empty, f32, require
from ..ib import (kernel_cosine_4pt, kernel_name, kernel_peskin_2pt, kernel_peskin_3pt, kernel_peskin_4pt,
                  interpolate, multi_direct_forcing, spread)


def get_ib_stencil(marker_coords, grid_shape, kernel=kernel_peskin_4pt, stencil_radius=2):
    """ib3d/stencil.py:7-59 (same ValueErrors) -> (W [M, S], I [M, S] int32)."""
    if len(grid_shape) != 3:
        raise ValueError(f"grid_shape must be a 3-tuple, got {grid_shape}.")
    p = f32(marker_coords)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError(f"marker_coords must have shape (n_markers, 3), got {tuple(p.shape)}.")
    nx, ny, nz = (int(v) for v in grid_shape)
    m = p.shape[0]
    s = (2 * int(stencil_radius)) ** 3
    W = empty((m, s), p)
    I = empty((m, s), p, dtype=torch.int32)
    L.call("vsx_ib_get_ib_stencil3d", L.ptr(p), m, nx, ny, nz, L.KERNEL[kernel_name(kernel)], int(stencil_radius),
           L.ptr(W), L.ptr(I))
    return W, I


def _mesh(vertex_coords, faces):
    v = f32(vertex_coords)
    f = faces.to(L.device(), torch.int32).contiguous() if isinstance(faces, torch.Tensor) else \
        torch.as_tensor(faces, dtype=torch.int32, device=L.device()).contiguous()
    require(v.ndim == 2 and v.shape[1] == 3 and f.ndim == 2 and f.shape[1] == 3, "need vertices [V,3], faces [F,3]")
    return v, f


def get_triangle_areas(vertex_coords, faces):
    """ib3d/geometry.py:6-18"""
    v, f = _mesh(vertex_coords, faces)
    out = empty((f.shape[0],), v)
    L.call("vsx_ib3d_get_triangle_areas", L.ptr(v), v.shape[0], L.ptr(f), f.shape[0], L.ptr(out))
    return out


def get_surface_area(vertex_coords, faces):
    """ib3d/geometry.py:21-23"""
    v, f = _mesh(vertex_coords, faces)
    out = empty((), v)
    L.call("vsx_ib3d_get_surface_area", L.ptr(v), v.shape[0], L.ptr(f), f.shape[0], L.ptr(out))
    return out


def get_volume(vertex_coords, faces):
    """ib3d/geometry.py:26-45"""
    v, f = _mesh(vertex_coords, faces)
    out = empty((), v)
    L.call("vsx_ib3d_get_volume", L.ptr(v), v.shape[0], L.ptr(f), f.shape[0], L.ptr(out))
    return out


def get_ds(vertex_coords, faces):
    """ib3d/geometry.py:48-67"""
    v, f = _mesh(vertex_coords, faces)
    out = empty((v.shape[0],), v)
    L.call("vsx_ib3d_get_ds", L.ptr(v), v.shape[0], L.ptr(f), f.shape[0], L.ptr(out))
    return out


__all__ = ["get_ds", "get_surface_area", "get_triangle_areas", "get_volume", "get_ib_stencil", "kernel_cosine_4pt",
           "kernel_peskin_3pt", "kernel_peskin_4pt", "kernel_peskin_2pt", "multi_direct_forcing", "interpolate", "spread"]
