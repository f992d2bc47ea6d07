"""Rigid-body structural update with the signatures of ``vivsim.dyn``
(vivsim/dyn.py:5-154).  The Newmark update and the force / torque reductions
run in FP64 on the device (the reference runs them in float32 under JAX's
default precision)."""

import torch

from . import _lib as L
from ._util import empty, f32, is_scalar, require, scalar


def _d64(x):
    if isinstance(x, torch.Tensor):
        return x.to(device=L.device(), dtype=torch.float64).contiguous()
    return torch.as_tensor(x, dtype=torch.float64, device=L.device()).contiguous()


def newmark(a, v, d, h, m, k, c, dt=1, gamma=0.5, beta=0.25):
    """Newmark-beta (dyn.py:5-51): scalar m, k, c (per-DOF) or matrices (multi-DOF)."""
    scalar_state = all(is_scalar(x) for x in (a, v, d, h))
    A, V, D, H = (_d64(x).reshape(-1) for x in (a, v, d, h))
    n = max(A.numel(), V.numel(), D.numel(), H.numel())
    A, V, D, H = (x.expand(n).contiguous() for x in (A, V, D, H))
    matrix = not is_scalar(m) and _d64(m).ndim > 0
    if matrix:
        Mt, Kt, Ct = (_d64(x).reshape(n, n).contiguous() for x in (m, k, c))
    else:
        Mt, Kt, Ct = (_d64(scalar(x)).reshape(1) for x in (m, k, c))
    ao, vo, do = (torch.empty(n, dtype=torch.float64, device=A.device) for _ in range(3))
    L.call("vsx_dyn_newmark", L.ptr(A), L.ptr(V), L.ptr(D), L.ptr(H), n, L.ptr(Mt), L.ptr(Kt), L.ptr(Ct), int(matrix),
           float(dt), float(gamma), float(beta), L.ptr(ao), L.ptr(vo), L.ptr(do))
    if scalar_state:
        return ao.reshape(()), vo.reshape(()), do.reshape(())
    return ao, vo, do


def newmark_2dof(a, v, d, h, m, k, c, dt=1, gamma=0.5, beta=0.25):
    """dyn.py:55-57"""
    return newmark(a, v, d, h, m, k, c, dt, gamma, beta)


def newmark_3dof(a, v, d, h, m, k, c, dt=1, gamma=0.5, beta=0.25):
    """dyn.py:60-62"""
    return newmark(a, v, d, h, m, k, c, dt, gamma, beta)


def _coords(x0, y0, xc, yc, d, rotate):
    x0, y0 = f32(x0).reshape(-1), f32(y0).reshape(-1)
    require(x0.shape == y0.shape, "marker arrays differ in length")
    dd = _d64(d).reshape(-1)
    require(dd.numel() >= (3 if rotate else 2), "displacement too short")
    x, y = torch.empty_like(x0), torch.empty_like(y0)
    L.call("vsx_dyn_markers_coords", L.ptr(x0), L.ptr(y0), x0.numel(), float(xc), float(yc), L.ptr(dd), int(rotate),
           L.ptr(x), L.ptr(y))
    return x, y


def get_markers_coords_2dof(x_markers_init, y_markers_init, d):
    """dyn.py:68-81"""
    return _coords(x_markers_init, y_markers_init, 0.0, 0.0, d, False)


def get_markers_coords_3dof(x_markers_init, y_markers_init, x_center_init, y_center_init, d):
    """dyn.py:84-100"""
    return _coords(x_markers_init, y_markers_init, scalar(x_center_init), scalar(y_center_init), d, True)


def get_markers_velocity_3dof(x_markers, y_markers, x_center_init, y_center_init, d, v):
    """dyn.py:103-120 -> [M, 2]"""
    x, y = f32(x_markers).reshape(-1), f32(y_markers).reshape(-1)
    dd, vv = _d64(d).reshape(-1), _d64(v).reshape(-1)
    out = empty((x.numel(), 2), x)
    L.call("vsx_dyn_markers_velocity_3dof", L.ptr(x), L.ptr(y), x.numel(), scalar(x_center_init),
           scalar(y_center_init), L.ptr(dd), L.ptr(vv), L.ptr(out))
    return out


def get_force_to_obj(h_markers):
    """dyn.py:126-136 (FP64 accumulation) -> [C] float64"""
    h = f32(h_markers)
    require(h.ndim == 2, "h_markers must be [M, C]")
    out = torch.empty(h.shape[1], dtype=torch.float64, device=h.device)
    L.call("vsx_dyn_get_force_to_obj", L.ptr(h), h.shape[0], h.shape[1], L.ptr(out))
    return out


def get_torque_to_obj(x_markers, y_markers, x_center_init, y_center_init, d, h_markers):
    """dyn.py:139-154 (FP64 accumulation) -> 0-d float64"""
    x, y = f32(x_markers).reshape(-1), f32(y_markers).reshape(-1)
    h = f32(h_markers)
    dd = _d64(d).reshape(-1)
    out = torch.empty((), dtype=torch.float64, device=h.device)
    L.call("vsx_dyn_get_torque_to_obj", L.ptr(x), L.ptr(y), scalar(x_center_init), scalar(y_center_init), L.ptr(dd),
           L.ptr(h), x.numel(), L.ptr(out))
    return out
