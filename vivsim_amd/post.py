"""Field diagnostics with the names and signatures of ``vivsim.post``
(vivsim/post.py:6-211), computed on the device by the HIP library
(``vsx_post_*``): a run can monitor vorticity, Q-criterion or energies
without copying its fields to the host.  Derivatives are ``jnp.gradient``'s
(unit spacing, central inside, first-order one-sided at the ends); FP32
arithmetic, means summed in FP64."""

import ctypes

import torch

from . import _lib as L
from ._util import empty, f32, numel, require

GRADIENT, VORTICITY, VORTICITY_MAGNITUDE, DIVERGENCE, STRAIN_RATE, STRAIN_RATE_MAGNITUDE, Q_CRITERION, ENSTROPHY, \
    VELOCITY_MAGNITUDE, KINETIC_ENERGY = range(10)


def _velocity(u):
    u = f32(u)
    require(u.ndim in (3, 4) and u.shape[0] == u.ndim - 1,
            f"u must have shape (2, NX, NY) or (3, NX, NY, NZ), got {tuple(u.shape)}")
    dims = tuple(u.shape[1:]) + ((1,) if u.ndim == 3 else ())
    return u, u.shape[0], dims


def _field(u, op, comps):
    u, dim, (nx, ny, nz) = _velocity(u)
    spatial = tuple(u.shape[1:])
    out = empty(tuple(comps(dim)) + spatial, u)
    L.call("vsx_post_field", L.ptr(u), dim, nx, ny, nz, op, L.ptr(out))
    return out


def velocity_magnitude(u):
    """post.py:6-14."""
    return _field(u, VELOCITY_MAGNITUDE, lambda d: ())


def velocity_gradient(u):
    """G[i, j] = du_i / dx_j, shape (dim, dim, *spatial) — post.py:17-29."""
    return _field(u, GRADIENT, lambda d: (d, d))


def vorticity(u):
    """2D scalar dv/dx - du/dy, 3D curl — post.py:32-54."""
    return _field(u, VORTICITY, lambda d: () if d == 2 else (3,))


def vorticity_magnitude(u):
    """post.py:57-68."""
    return _field(u, VORTICITY_MAGNITUDE, lambda d: ())


def divergence(u):
    """post.py:72-83."""
    return _field(u, DIVERGENCE, lambda d: ())


def strain_rate(u):
    """post.py:86-95."""
    return _field(u, STRAIN_RATE, lambda d: (d, d))


def strain_rate_magnitude(u):
    """post.py:98-107."""
    return _field(u, STRAIN_RATE_MAGNITUDE, lambda d: ())


def kinetic_energy(u):
    """post.py:110-118."""
    return _field(u, KINETIC_ENERGY, lambda d: ())


def _mean(u, op):
    u, dim, (nx, ny, nz) = _velocity(u)
    partial = empty((1024,), u, torch.float64)
    out = empty((), u)
    L.call("vsx_post_mean", L.ptr(u), dim, nx, ny, nz, op, L.ptr(partial), L.ptr(out))
    return out


def mean_kinetic_energy(u):
    """post.py:121-129."""
    return _mean(u, KINETIC_ENERGY)


def pressure(rho, cs2=1.0 / 3.0):
    """post.py:132-141."""
    r = f32(rho)
    out = torch.empty_like(r)
    L.call("vsx_post_pressure", L.ptr(r), numel(r.shape), ctypes.c_float(cs2), L.ptr(out))
    return out


def enstrophy(u):
    """post.py:144-155."""
    return _field(u, ENSTROPHY, lambda d: ())


def mean_enstrophy(u):
    """post.py:158-166."""
    return _mean(u, ENSTROPHY)


def q_criterion(u):
    """Q = -0.5 tr(G G) — post.py:172-187."""
    return _field(u, Q_CRITERION, lambda d: ())


def calculate_curl(u):
    """post.py:194-196."""
    return vorticity(u)


def calculate_vorticity(u):
    """post.py:199-201."""
    return vorticity(u)


def calculate_vorticity_dimensionless(u, l, u0):
    """post.py:204-206."""
    return vorticity(u) * l / u0


def calculate_velocity_magnitude(u):
    """post.py:209-211."""
    return velocity_magnitude(u)


def force_coefficients(h_hist, D, U0):
    """(Cd, Cl) = 2 h / (D U0^2) of a device force history [n, 2] (FP64), as the
    VIV example evaluates them (examples/2d/vortex_induced_vibration.py:228-229)."""
    h = h_hist.to(device=L.device(), dtype=torch.float64).contiguous()
    require(h.ndim == 2 and h.shape[1] == 2, "h_hist must have shape (n, 2)")
    cd = torch.empty(h.shape[0], dtype=torch.float64, device=h.device)
    cl = torch.empty_like(cd)
    L.call("vsx_force_coefficients", L.ptr(h), h.shape[0], float(D), float(U0), L.ptr(cd), L.ptr(cl))
    return cd, cl


__all__ = ["velocity_magnitude", "velocity_gradient", "vorticity", "vorticity_magnitude", "divergence", "strain_rate",
           "strain_rate_magnitude", "kinetic_energy", "mean_kinetic_energy", "pressure", "enstrophy",
           "mean_enstrophy", "q_criterion", "calculate_curl", "calculate_vorticity",
           "calculate_vorticity_dimensionless", "calculate_velocity_magnitude", "force_coefficients"]
