"""Immersed-boundary operators with the signatures of ``vivsim.ib``
(vivsim/ib/__init__.py:1-6).  The delta kernels are passed, as in the
reference, as the function objects defined here; they select the HIP
implementation (a Python callable cannot run inside a HIP kernel)."""

import ctypes

import torch

from .. import _lib as L
from .._util import empty, f32, is_scalar, numel, require, scalar


def _kernel_eval(kind, r):
    r = f32(r)
    out = torch.empty_like(r)
    L.call("vsx_ib_kernel", L.KERNEL[kind], L.ptr(r), L.ptr(out), numel(r.shape))
    return out


def kernel_peskin_3pt(r):
    """ib/kernels.py:4-22"""
    return _kernel_eval("peskin_3pt", r)


def kernel_peskin_4pt(r):
    """ib/kernels.py:25-43"""
    return _kernel_eval("peskin_4pt", r)


def kernel_cosine_4pt(r):
    """ib/kernels.py:46-61"""
    return _kernel_eval("cosine_4pt", r)


def kernel_peskin_2pt(r):
    """Linear hat max(0, 1 - |r|) — not in the reference; an extension."""
    return _kernel_eval("peskin_2pt", r)


_KERNEL_ID = {kernel_peskin_2pt: "peskin_2pt", kernel_peskin_3pt: "peskin_3pt",
              kernel_peskin_4pt: "peskin_4pt", kernel_cosine_4pt: "cosine_4pt"}


def kernel_name(kernel):
    """Map a kernel function (or its name) to the HIP kernel id."""
    if isinstance(kernel, str):
        if kernel not in L.KERNEL:
            raise ValueError(f"unknown kernel {kernel!r}")
        return kernel
    try:
        return _KERNEL_ID[kernel]
    except (KeyError, TypeError):
        pass
    name = getattr(kernel, "__name__", "")
    for k in L.KERNEL:
        if name == "kernel_" + k:
            return k
    raise ValueError("kernel must be one of kernel_peskin_2pt/3pt/4pt or kernel_cosine_4pt")


def get_ib_stencil(marker_x, marker_y, ny, kernel=kernel_peskin_4pt, stencil_radius=2):
    """ib/stencil.py:7-51 -> (weights float32 [M, S], indices int32 [M, S])."""
    mx, my = f32(marker_x).reshape(-1), f32(marker_y).reshape(-1)
    require(mx.shape == my.shape, "marker_x and marker_y must have the same length")
    m = mx.shape[0]
    s = (2 * int(stencil_radius)) ** 2
    W = empty((m, s), mx)
    I = empty((m, s), mx, dtype=torch.int32)
    L.call("vsx_ib_get_ib_stencil2d", L.ptr(mx), L.ptr(my), m, int(ny), L.KERNEL[kernel_name(kernel)],
           int(stencil_radius), L.ptr(W), L.ptr(I))
    return W, I


def _stencil(stencil_weights, stencil_indices):
    W = f32(stencil_weights)
    I = stencil_indices.to(L.device(), torch.int32).contiguous() if isinstance(stencil_indices, torch.Tensor) \
        else torch.as_tensor(stencil_indices, dtype=torch.int32, device=L.device()).contiguous()
    require(W.ndim == 2 and W.shape == I.shape, "stencil weights and indices must be [M, S] of equal shape")
    return W, I


def interpolate(grid_values, stencil_weights, stencil_indices):
    """ib/stencil.py:54-78 -> [M, C]."""
    g = f32(grid_values)
    W, I = _stencil(stencil_weights, stencil_indices)
    c = g.shape[0]
    out = empty((W.shape[0], c), g)
    L.call("vsx_ib_interpolate", L.ptr(g), c, numel(g.shape[1:]), L.ptr(W), L.ptr(I), W.shape[0], W.shape[1], L.ptr(out))
    return out


def spread(marker_values, grid_values, stencil_weights, stencil_indices):
    """ib/stencil.py:81-110 -> grid_values + scatter-add of marker_values."""
    g = f32(grid_values)
    mv = f32(marker_values)
    W, I = _stencil(stencil_weights, stencil_indices)
    require(mv.shape == (W.shape[0], g.shape[0]), "marker_values must be [M, C]")
    out = torch.empty_like(g)
    L.call("vsx_ib_spread", L.ptr(mv), L.ptr(g), L.ptr(out), g.shape[0], numel(g.shape[1:]), L.ptr(W), L.ptr(I),
           W.shape[0], W.shape[1])
    return out


def multi_direct_forcing(grid_u, stencil_weights, stencil_indices, marker_u_target, marker_ds, n_iter=5):
    """ib/mdf.py:10-64 -> (grid_force like grid_u, marker_reaction_force [M, C])."""
    g = f32(grid_u)
    W, I = _stencil(stencil_weights, stencil_indices)
    m, c = W.shape[0], g.shape[0]
    U = f32(marker_u_target)
    require(U.shape == (m, c), "marker_u_target must be [M, C]")
    if is_scalar(marker_ds):
        ds_arr, ds = None, scalar(marker_ds)
    else:
        ds_arr, ds = f32(marker_ds).reshape(-1), 0.0
        if ds_arr.numel() == 1:
            ds_arr, ds = None, float(ds_arr.item())
    gf = torch.empty_like(g)
    mf = empty((m, c), g)
    work = empty((3 * m * c,), g)
    L.call("vsx_ib_multi_direct_forcing", L.ptr(g), c, numel(g.shape[1:]), L.ptr(W), L.ptr(I), m, W.shape[1],
           L.ptr(U), L.ptr(ds_arr), ctypes.c_float(ds), int(n_iter), L.ptr(gf), L.ptr(mf), L.ptr(work))
    return gf, mf


def get_area(marker_coords):
    """ib/geometry.py:6-19 -> 0-d tensor."""
    p = f32(marker_coords)
    require(p.ndim == 2 and p.shape[1] == 2, "marker_coords must be [M, 2]")
    out = empty((), p)
    L.call("vsx_ib_get_area", L.ptr(p), p.shape[0], L.ptr(out))
    return out


def get_ds(marker_coords, closed=True):
    """ib/geometry.py:22-43 -> [M]."""
    p = f32(marker_coords)
    require(p.ndim == 2 and p.shape[1] == 2, "marker_coords must be [M, 2]")
    out = empty((p.shape[0],), p)
    L.call("vsx_ib_get_ds", L.ptr(p), p.shape[0], int(bool(closed)), L.ptr(out))
    return out


__all__ = ["get_area", "get_ds", "kernel_cosine_4pt", "kernel_peskin_3pt", "kernel_peskin_4pt", "kernel_peskin_2pt",
           "multi_direct_forcing", "get_ib_stencil", "interpolate", "spread"]
