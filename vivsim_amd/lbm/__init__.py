"""D2Q9 lattice Boltzmann operators with the signatures of ``vivsim.lbm``
(vivsim/lbm/__init__.py:1-42).  Arrays are torch tensors on the HIP device in
the reference layout (``f[9, NX, NY]``, ``u[2, NX, NY]``); host arrays are
copied to the device.  Every function returns new arrays (pure, like the
reference)."""

import numpy as np
import torch

from .. import _ops
from ..lattice import D2Q9

_L = _ops.Lattice("d2q9", 9, 2, _ops.LOCS2)

# ------------------------------------------------------------------ basic (lbm/basic.py)


def streaming(f):
    """lbm/basic.py:60-85"""
    return _L.streaming(f)


def get_macroscopic(f):
    """lbm/basic.py:88-111"""
    return _L.get_macroscopic(f)


def get_equilibrium(rho, u):
    """lbm/basic.py:114-135"""
    return _L.get_equilibrium(rho, u)


def collision_bgk(f, feq, omega):
    """lbm/basic.py:138-156"""
    return _L.collision_bgk(f, feq, omega)


def get_omega(nu):
    """lbm/basic.py:159-172"""
    return 1 / (3 * nu + 0.5)


def get_velocity_correction(g, rho=1):
    """lbm/basic.py:175-194"""
    return _ops.velocity_correction(g, rho)


# ------------------------------------------------------------------ collisions

# Lallemand-Luo basis (lbm/collision/mrt.py:10-22)
_M = np.array([
    [1, 1, 1, 1, 1, 1, 1, 1, 1], [-4, -1, -1, -1, -1, 2, 2, 2, 2], [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1], [0, -2, 0, 2, 0, 1, -1, -1, 1], [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1], [0, 1, -1, 1, -1, 0, 0, 0, 0], [0, 0, 0, 0, 0, 1, -1, 1, -1]])


def mrt_relaxation_rates(omega):
    """diag(S) of lbm/collision/mrt.py:28-44."""
    return np.array([0, 1.4, 1.4, 0, 1.2, 0, 1.2, omega, omega])


def get_mrt_collision_operator(omega):
    """M^-1 S M as float32, formed as the reference does under JAX's default
    precision (lbm/collision/mrt.py:47-63)."""
    minv = np.linalg.inv(_M.astype(np.float64)).astype(np.float32)
    S = np.diag(mrt_relaxation_rates(omega)).astype(np.float32)
    return torch.from_numpy((minv @ S) @ _M.astype(np.float32))


def get_mrt_forcing_operator(omega):
    """M^-1 (I - S/2) M (lbm/forcing/guo.py:62-79)."""
    minv = np.linalg.inv(_M.astype(np.float64)).astype(np.float32)
    S = np.diag(mrt_relaxation_rates(omega)).astype(np.float32)
    return torch.from_numpy((minv @ (np.eye(9, dtype=np.float32) - np.float32(0.5) * S)) @ _M.astype(np.float32))


def collision_mrt(f, feq, mrt_collision_matrix):
    """lbm/collision/mrt.py:66-88"""
    return _L.collision_mrt(f, feq, mrt_collision_matrix)


def collision_kbc(f, feq, omega):
    """lbm/collision/kbc.py:10-59"""
    return _L.collision_kbc(f, feq, omega)


def collision_reg(f, feq, omega):
    """lbm/collision/reg.py:4-49"""
    return _L.collision_reg(f, feq, omega)


# ------------------------------------------------------------------ forcing


def get_guo_forcing_term(g, u):
    """lbm/forcing/guo.py:6-33"""
    return _L.get_guo_forcing_term(g, u)


def forcing_guo_bgk(f, g, u, omega):
    """lbm/forcing/guo.py:38-57"""
    return _L.forcing_guo_bgk(f, g, u, omega)


def forcing_guo_mrt(f, g, u, mrt_forcing_operator):
    """lbm/forcing/guo.py:82-107"""
    return _L.forcing_guo_mrt(f, g, u, mrt_forcing_operator)


def forcing_edm(f, g, u):
    """lbm/forcing/edm.py:4-31"""
    return _L.forcing_edm(f, g, u)


# ------------------------------------------------------------------ boundaries


def boundary_nee(f, loc, rho_wall=1, ux_wall=0, uy_wall=0):
    """lbm/boundary/nee.py:19-60"""
    return _L.wall(f, "nee", "plain", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall)


def boundary_velocity_nee(f, loc, ux_wall=0, uy_wall=0):
    """lbm/boundary/nee.py:62 (+ _helpers.py:180-197)"""
    return _L.wall(f, "nee", "velocity", loc, ux=ux_wall, uy=uy_wall)


def boundary_pressure_nee(f, loc, rho_wall=1):
    """lbm/boundary/nee.py:63 (+ _helpers.py:200-217)"""
    return _L.wall(f, "nee", "pressure", loc, rho=rho_wall)


def boundary_force_corrected_nee(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, gx_wall=0, gy_wall=0):
    """lbm/boundary/nee.py:64 (+ _helpers.py:220-239)"""
    return _L.wall(f, "nee", "force_corrected", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, gx=gx_wall, gy=gy_wall)


def boundary_nebb(f, loc, rho_wall=1, ux_wall=0, uy_wall=0):
    """lbm/boundary/nebb.py:20-58"""
    return _L.wall(f, "nebb", "plain", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall)


def boundary_velocity_nebb(f, loc, ux_wall=0, uy_wall=0):
    """lbm/boundary/nebb.py:60"""
    return _L.wall(f, "nebb", "velocity", loc, ux=ux_wall, uy=uy_wall)


def boundary_pressure_nebb(f, loc, rho_wall=1):
    """lbm/boundary/nebb.py:61"""
    return _L.wall(f, "nebb", "pressure", loc, rho=rho_wall)


def boundary_force_corrected_nebb(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, gx_wall=0, gy_wall=0):
    """lbm/boundary/nebb.py:62"""
    return _L.wall(f, "nebb", "force_corrected", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, gx=gx_wall, gy=gy_wall)


def boundary_equilibrium(f, loc, rho_wall=1, ux_wall=0, uy_wall=0):
    """lbm/boundary/eq.py:25-56"""
    return _L.wall(f, "eq", "plain", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall)


def boundary_velocity_equilibrium(f, loc, ux_wall=0, uy_wall=0):
    """lbm/boundary/eq.py:59"""
    return _L.wall(f, "eq", "velocity", loc, ux=ux_wall, uy=uy_wall)


def boundary_pressure_equilibrium(f, loc, rho_wall=1):
    """lbm/boundary/eq.py:60"""
    return _L.wall(f, "eq", "pressure", loc, rho=rho_wall)


def boundary_force_corrected_equilibrium(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, gx_wall=0, gy_wall=0):
    """lbm/boundary/eq.py:61"""
    return _L.wall(f, "eq", "force_corrected", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, gx=gx_wall, gy=gy_wall)


def boundary_bounce_back(f_before_stream, f, loc, ux_wall=0, uy_wall=0):
    """lbm/boundary/bb.py:15-59"""
    return _L.wall(f, "bb", "plain", loc, f_pre=f_before_stream, ux=ux_wall, uy=uy_wall)


def boundary_specular_reflection(f_before_stream, f, loc, ux_wall=0, uy_wall=0):
    """lbm/boundary/bb.py:62-95"""
    return _L.wall(f, "specular", "plain", loc, f_pre=f_before_stream, ux=ux_wall, uy=uy_wall)


def obstacle_bounce_back(f, mask):
    """lbm/boundary/bb.py:98-110"""
    return _L.obstacle_bounce_back(f, mask)


def boundary_characteristic(rho, u, loc="right"):
    """lbm/boundary/cbc.py:14-53"""
    return _L.boundary_characteristic(rho, u, loc)


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("np", "torch")]
