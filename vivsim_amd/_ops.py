"""Shared implementation of the lattice operators for ``vivsim_amd.lbm`` (D2Q9)
and ``vivsim_amd.lbm3d`` (D3Q19): argument checking, output allocation and
the C-ABI calls.  Every array result is computed by a HIP kernel."""

import ctypes

import numpy as np
import torch

from . import _lib as L
from ._util import WallBuilder, check_pop, empty, f32, is_scalar, numel, require, scalar, wall_nodes_shape

LOCS2 = ("left", "right", "top", "bottom")
LOCS3 = ("left", "right", "bottom", "top", "back", "front")


class Lattice:
    def __init__(self, pfx, q, d, locs):
        self.pfx, self.q, self.d, self.locs = pfx, q, d, locs

    def fn(self, name):
        return f"vsx_{self.pfx}_{name}"

    # ------------------------------------------------------------ pointwise
    def streaming(self, f):
        f = f32(f)
        check_pop(f, self.q)
        require(f.ndim == 1 + self.d, f"f must have {1 + self.d} dimensions")
        out = torch.empty_like(f)
        L.call(self.fn("streaming"), L.ptr(f), L.ptr(out), *[int(s) for s in f.shape[1:]])
        return out

    def get_macroscopic(self, f):
        f = f32(f)
        check_pop(f, self.q)
        S = tuple(f.shape[1:])
        rho = empty(S, f)
        u = empty((self.d,) + S, f)
        L.call(self.fn("get_macroscopic"), L.ptr(f), L.ptr(rho), L.ptr(u), numel(S))
        return rho, u

    def get_equilibrium(self, rho, u):
        u = f32(u)
        rho = f32(rho)
        require(u.shape[0] == self.d, f"u must have shape ({self.d}, ...)")
        S = tuple(u.shape[1:])
        rho = rho.expand(S).contiguous() if tuple(rho.shape) != S else rho
        feq = empty((self.q,) + S, u)
        L.call(self.fn("get_equilibrium"), L.ptr(rho), L.ptr(u), L.ptr(feq), numel(S))
        return feq

    def _pair(self, f, feq):
        f, feq = f32(f), f32(feq)
        check_pop(f, self.q)
        require(f.shape == feq.shape, "f and feq must have the same shape")
        return f, feq, torch.empty_like(f), numel(f.shape[1:])

    def collision_bgk(self, f, feq, omega):
        f, feq, out, n = self._pair(f, feq)
        L.call(self.fn("collision_bgk"), L.ptr(f), L.ptr(feq), ctypes.c_float(float(omega)), L.ptr(out), n)
        return out

    def _op(self, op):
        a = op.detach().cpu().numpy() if isinstance(op, torch.Tensor) else np.asarray(op)
        a = np.ascontiguousarray(a, dtype=np.float32)
        require(a.shape == (self.q, self.q), f"operator must be {self.q}x{self.q}")
        return a

    def collision_mrt(self, f, feq, op):
        f, feq, out, n = self._pair(f, feq)
        a = self._op(op)
        L.call(self.fn("collision_mrt"), L.ptr(f), L.ptr(feq), a.ctypes.data_as(ctypes.c_void_p), L.ptr(out), n)
        return out

    def collision_kbc(self, f, feq, omega):
        f, feq, out, n = self._pair(f, feq)
        L.call(self.fn("collision_kbc"), L.ptr(f), L.ptr(feq), ctypes.c_float(float(omega)), L.ptr(out), n)
        return out

    def collision_reg(self, f, feq, omega):
        f, feq, out, n = self._pair(f, feq)
        L.call(self.fn("collision_reg"), L.ptr(f), L.ptr(feq), ctypes.c_float(float(omega)), L.ptr(out), n)
        return out

    def _gu(self, g, u):
        u = f32(u)
        require(u.shape[0] == self.d, f"u must have shape ({self.d}, ...)")
        g = f32(g)
        g = g.expand(u.shape).contiguous() if g.shape != u.shape else g
        return g, u, numel(u.shape[1:])

    def get_guo_forcing_term(self, g, u):
        g, u, n = self._gu(g, u)
        out = empty((self.q,) + tuple(u.shape[1:]), u)
        L.call(self.fn("get_guo_forcing_term"), L.ptr(g), L.ptr(u), L.ptr(out), n)
        return out

    def forcing_guo_bgk(self, f, g, u, omega):
        f = f32(f)
        g, u, n = self._gu(g, u)
        out = torch.empty_like(f)
        L.call(self.fn("forcing_guo_bgk"), L.ptr(f), L.ptr(g), L.ptr(u), ctypes.c_float(float(omega)), L.ptr(out), n)
        return out

    def forcing_guo_mrt(self, f, g, u, op):
        f = f32(f)
        g, u, n = self._gu(g, u)
        a = self._op(op)
        out = torch.empty_like(f)
        L.call(self.fn("forcing_guo_mrt"), L.ptr(f), L.ptr(g), L.ptr(u), a.ctypes.data_as(ctypes.c_void_p), L.ptr(out), n)
        return out

    def forcing_edm(self, f, g, u):
        f = f32(f)
        g, u, n = self._gu(g, u)
        out = torch.empty_like(f)
        L.call(self.fn("forcing_edm"), L.ptr(f), L.ptr(g), L.ptr(u), L.ptr(out), n)
        return out

    def obstacle_bounce_back(self, f, mask):
        f = f32(f)
        check_pop(f, self.q)
        m = f32(mask).ne(0).to(torch.uint8) if not isinstance(mask, torch.Tensor) or mask.device.type != "cuda" \
            else mask.to(torch.uint8)
        m = m.expand(f.shape[1:]).contiguous()
        out = torch.empty_like(f)
        L.call(self.fn("obstacle_bounce_back"), L.ptr(f), L.ptr(m), L.ptr(out), numel(f.shape[1:]))
        return out

    # ------------------------------------------------------------ walls
    def wall(self, f, kind, mode, loc, f_pre=None, **vals):
        if loc not in self.locs:
            raise KeyError(loc)
        f = f32(f)
        check_pop(f, self.q)
        require(f.ndim == 1 + self.d, f"f must have {1 + self.d} dimensions")
        shape = tuple(f.shape[1:])
        wb = WallBuilder()
        w = wb(kind, mode, loc, wall_nodes_shape(shape, loc), **vals)
        pre = None
        if kind in ("bb", "specular"):
            pre = f32(f_pre)
            require(pre.shape == f.shape, "f_before_stream must have the shape of f")
        out = torch.empty_like(f)
        L.call(self.fn("boundary"), ctypes.byref(w), L.ptr(pre), L.ptr(f), L.ptr(out), *shape)
        return out

    def boundary_characteristic(self, rho, u, loc):
        if loc not in self.locs:
            raise ValueError(f"loc must be one of: {', '.join(repr(x) for x in self.locs)}")
        rho, u = f32(rho), f32(u)
        shape = tuple(rho.shape)
        require(len(shape) == self.d and tuple(u.shape) == (self.d,) + shape, "rho / u shapes do not match")
        ns = wall_nodes_shape(shape, loc)
        rn = empty(ns, rho)
        un = empty((self.d,) + ns, rho)
        L.call(self.fn("boundary_characteristic"), L.ptr(rho), L.ptr(u), L.LOC[loc], L.ptr(rn), L.ptr(un), *shape)
        return rn, un


def velocity_correction(g, rho=1):
    """0.5 g / rho — lbm/basic.py:175-194.  Scalar inputs give a scalar."""
    if is_scalar(g) and is_scalar(rho):
        return scalar(g) * 0.5 / scalar(rho)
    if is_scalar(rho):
        return _vc(f32(g), None, scalar(rho))
    rt = f32(rho)
    gt = torch.full(tuple(rt.shape), scalar(g), dtype=torch.float32, device=rt.device) if is_scalar(g) else f32(g)
    return _vc(gt, rt, 1.0)


def _vc(g, rho, rho_s):
    out = torch.empty_like(g)
    S = g.shape[1:] if rho is not None and g.ndim == rho.ndim + 1 else g.shape
    if rho is not None and g.ndim == rho.ndim + 1:
        comps = g.shape[0]
        r = rho.expand(S).contiguous()
    else:
        comps = 1
        r = rho.expand(g.shape).contiguous() if rho is not None else None
    L.call("vsx_velocity_correction", L.ptr(g), L.ptr(r), ctypes.c_float(rho_s), L.ptr(out), comps, numel(S))
    return out
