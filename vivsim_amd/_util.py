"""Tensor plumbing for the Python API: device placement, dtype, wall structs."""

import numbers

import numpy as np
import torch

from . import _lib as L


def is_scalar(v):
    """Scalar wall data (Python / NumPy number or 0-d array), cf. ``jnp.isscalar``
    in lbm/boundary/_helpers.py:70-75."""
    if isinstance(v, numbers.Number) or np.isscalar(v):
        return True
    return hasattr(v, "ndim") and v.ndim == 0


def scalar(v):
    return float(v.item() if hasattr(v, "item") else v)


def dev(x, dtype=torch.float32):
    """``x`` as a contiguous tensor of ``dtype`` on the HIP device (host data is copied there)."""
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.as_tensor(np.asarray(x))
    if t.device.type != "cuda":
        t = t.to(L.device())
    if t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def f32(x):
    return dev(x, torch.float32)


def empty(shape, like=None, dtype=torch.float32):
    return torch.empty(tuple(shape), dtype=dtype, device=L.device() if like is None else like.device)


def numel(shape):
    n = 1
    for s in shape:
        n *= int(s)
    return n


def require(cond, msg):
    if not cond:
        raise ValueError(msg)


def check_pop(f, q, name="f"):
    require(f.ndim >= 2 and f.shape[0] == q, f"{name} must have shape ({q}, ...), got {tuple(f.shape)}")


class WallBuilder:
    """Builds a ``VsxWall`` (scalars inline, per-node arrays as device pointers).
    Keeps the array tensors alive for as long as the builder lives."""

    def __init__(self):
        self.keep = []

    def __call__(self, kind, mode, loc, nodes_shape, rho=1.0, ux=0.0, uy=0.0, uz=0.0, gx=0.0, gy=0.0, gz=0.0):
        w = L.VsxWall()
        w.kind = L.WALL[kind]
        w.mode = L.MODE[mode]
        w.loc = L.LOC[loc]
        for name, v in (("rho", rho), ("ux", ux), ("uy", uy), ("uz", uz), ("gx", gx), ("gy", gy), ("gz", gz)):
            if is_scalar(v):
                setattr(w, name, scalar(v))
                setattr(w, name + "_arr", None)
            else:
                t = f32(v)
                t = t.expand(tuple(nodes_shape)).contiguous() if tuple(t.shape) != tuple(nodes_shape) else t
                self.keep.append(t)
                setattr(w, name, 0.0)
                setattr(w, name + "_arr", t.data_ptr())
        return w


def wall_nodes_shape(shape, loc, walls3d=False):
    """Node shape of a wall on a grid of ``shape`` (spatial only)."""
    axis = {"left": 0, "right": 0, "bottom": 1, "top": 1, "back": 2, "front": 2}[loc]
    return tuple(s for a, s in enumerate(shape) if a != axis)
