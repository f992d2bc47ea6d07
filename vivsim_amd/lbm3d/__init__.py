"""D3Q19 lattice Boltzmann operators with the signatures of ``vivsim.lbm3d``
(vivsim/lbm3d/__init__.py:1-44).  Layout ``f[19, NX, NY, NZ]``, ``u[3, ...]``."""

import numpy as np
import torch

from .. import _ops
from ..lattice import D3Q19

_L = _ops.Lattice("d3q19", 19, 3, _ops.LOCS3)


def streaming(f):
    """lbm3d/basic.py:52-87"""
    return _L.streaming(f)


def get_macroscopic(f):
    """lbm3d/basic.py:90-106"""
    return _L.get_macroscopic(f)


def get_equilibrium(rho, u):
    """lbm3d/basic.py:109-130"""
    return _L.get_equilibrium(rho, u)


def collision_bgk(f, feq, omega):
    """lbm3d/basic.py:133-146"""
    return _L.collision_bgk(f, feq, omega)


def get_omega(nu):
    """lbm3d/basic.py:149-152"""
    return 1 / (3 * nu + 0.5)


def get_velocity_correction(g, rho=1):
    """lbm3d/basic.py:155-167"""
    return _ops.velocity_correction(g, rho)


def _basis():
    """Raw-moment basis of lbm3d/collision/mrt.py:7-30 (float32 in the reference)."""
    c = D3Q19.c.astype(np.int64)
    cx, cy, cz = c[:, 0], c[:, 1], c[:, 2]
    c2 = cx * cx + cy * cy + cz * cz
    rows = [np.ones(19, dtype=np.int64), cx, cy, cz, c2, 3 * cx * cx - c2, cy * cy - cz * cz,
            cx * cy, cx * cz, cy * cz, cx * cx * cy, cx * cx * cz, cx * cy * cy, cy * cy * cz, cx * cz * cz,
            cy * cz * cz, cx * cx * cy * cy, cx * cx * cz * cz, cy * cy * cz * cz]
    return np.array(rows, dtype=np.float32)


_M = _basis()
_MINV = np.linalg.inv(_M)          # float32 as in lbm3d/collision/mrt.py:31


def mrt_relaxation_rates(omega):
    """diag(S) of lbm3d/collision/mrt.py:34-73."""
    return np.array([0, 0, 0, 0, 1.1] + [omega] * 5 + [1.2] * 6 + [1.4] * 3)


def get_mrt_collision_operator(omega):
    """lbm3d/collision/mrt.py:76-80"""
    S = np.diag(mrt_relaxation_rates(omega)).astype(np.float32)
    return torch.from_numpy((_MINV @ S) @ _M)


def get_mrt_forcing_operator(omega):
    """lbm3d/forcing/guo.py:59-68"""
    S = np.diag(mrt_relaxation_rates(omega)).astype(np.float32)
    return torch.from_numpy((_MINV @ (np.eye(19, dtype=np.float32) - np.float32(0.5) * S)) @ _M)


def collision_mrt(f, feq, mrt_collision_matrix):
    """lbm3d/collision/mrt.py:83-98"""
    return _L.collision_mrt(f, feq, mrt_collision_matrix)


def collision_kbc(f, feq, omega):
    """lbm3d/collision/kbc.py:8-42"""
    return _L.collision_kbc(f, feq, omega)


def collision_reg(f, feq, omega):
    """lbm3d/collision/reg.py:45-62"""
    return _L.collision_reg(f, feq, omega)


def get_guo_forcing_term(g, u):
    """lbm3d/forcing/guo.py:13-38"""
    return _L.get_guo_forcing_term(g, u)


def forcing_guo_bgk(f, g, u, omega):
    """lbm3d/forcing/guo.py:41-56"""
    return _L.forcing_guo_bgk(f, g, u, omega)


def forcing_guo_mrt(f, g, u, mrt_forcing_operator):
    """lbm3d/forcing/guo.py:71-87"""
    return _L.forcing_guo_mrt(f, g, u, mrt_forcing_operator)


def forcing_edm(f, g, u):
    """lbm3d/forcing/edm.py:6-26"""
    return _L.forcing_edm(f, g, u)


def _w(f, kind, mode, loc, **kw):
    return _L.wall(f, kind, mode, loc, **kw)


def boundary_nee(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/nee.py:13-33"""
    return _w(f, "nee", "plain", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def boundary_velocity_nee(f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/nee.py:35 (+ _helpers.py:91-102)"""
    return _w(f, "nee", "velocity", loc, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def boundary_pressure_nee(f, loc, rho_wall=1):
    """lbm3d/boundary/nee.py:36 (+ _helpers.py:105-116)"""
    return _w(f, "nee", "pressure", loc, rho=rho_wall)


def boundary_force_corrected_nee(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0, gx_wall=0, gy_wall=0, gz_wall=0):
    """lbm3d/boundary/nee.py:37 (+ _helpers.py:119-133)"""
    return _w(f, "nee", "force_corrected", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, uz=uz_wall,
              gx=gx_wall, gy=gy_wall, gz=gz_wall)


def boundary_nebb(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/nebb.py:13-30"""
    return _w(f, "nebb", "plain", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def boundary_velocity_nebb(f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/nebb.py:33"""
    return _w(f, "nebb", "velocity", loc, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def boundary_pressure_nebb(f, loc, rho_wall=1):
    """lbm3d/boundary/nebb.py:34"""
    return _w(f, "nebb", "pressure", loc, rho=rho_wall)


def boundary_force_corrected_nebb(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0, gx_wall=0, gy_wall=0, gz_wall=0):
    """lbm3d/boundary/nebb.py:35"""
    return _w(f, "nebb", "force_corrected", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, uz=uz_wall,
              gx=gx_wall, gy=gy_wall, gz=gz_wall)


def boundary_equilibrium(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/eq.py:13-26"""
    return _w(f, "eq", "plain", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def boundary_velocity_equilibrium(f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/eq.py:29"""
    return _w(f, "eq", "velocity", loc, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def boundary_pressure_equilibrium(f, loc, rho_wall=1):
    """lbm3d/boundary/eq.py:30"""
    return _w(f, "eq", "pressure", loc, rho=rho_wall)


def boundary_force_corrected_equilibrium(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0, gx_wall=0, gy_wall=0,
                                         gz_wall=0):
    """lbm3d/boundary/eq.py:31"""
    return _w(f, "eq", "force_corrected", loc, rho=rho_wall, ux=ux_wall, uy=uy_wall, uz=uz_wall,
              gx=gx_wall, gy=gy_wall, gz=gz_wall)


def boundary_bounce_back(f_before_stream, f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/bb.py:20-34"""
    return _w(f, "bb", "plain", loc, f_pre=f_before_stream, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def boundary_specular_reflection(f_before_stream, f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/bb.py:37-53 (same update as bounce-back in the reference)"""
    return _w(f, "specular", "plain", loc, f_pre=f_before_stream, ux=ux_wall, uy=uy_wall, uz=uz_wall)


def obstacle_bounce_back(f, mask):
    """lbm3d/boundary/bb.py:56-59"""
    return _L.obstacle_bounce_back(f, mask)


def boundary_characteristic(rho, u, loc="right"):
    """lbm3d/boundary/cbc.py:16-51"""
    return _L.boundary_characteristic(rho, u, loc)


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("np", "torch")]
