"""Differentiable float64 restatement of the 2D VIV step (test infrastructure).

The reference differentiates its jitted step with ``jax.grad``; JAX is absent
here, so gradients of the VIV objective are pinned by torch.autograd through
this float64 restatement (CPU), itself pinned against the NumPy oracle's
forward (oracle/steps.py) and against central finite differences.

The composition is the one of ``oracle.steps.viv_step`` (MRT + Guo, equilibrium
inlet, pressure-NEE outlet, periodic y, Peskin 4-point stencil, MDF, FP64
Newmark; examples/2d/vortex_induced_vibration.py:96-148) with three design
parameters made explicit:

* ``delta`` [M]: radial offsets of the markers, X_m = c + (R + delta_m) e_m
  (the marker arc length ds stays the case's constant, as MARKER_DS is in
  vortex_induced_vibration.py:44);
* ``stiffness``: the spring constant K of the Newmark update (dyn.py:5-51);
* ``u_in``: the inlet wall velocity of boundary_equilibrium (lbm/boundary/eq.py:25-56).

Objective: J = (1/T) sum_{t=1..T} d_y(t)^2 (d after step t).
"""

import math

import numpy as np
import torch

from . import lattice as LT
from . import lbm2d as L2

C = torch.tensor(LT.C2, dtype=torch.float64)        # [9, 2]
W = torch.tensor(LT.W2, dtype=torch.float64)        # [9]


def macroscopic(f):
    rho = f.sum(0)
    ux = (f[1] + f[5] + f[8]) - (f[3] + f[7] + f[6])
    uy = (f[2] + f[5] + f[6]) - (f[4] + f[7] + f[8])
    return rho, torch.stack([ux / rho, uy / rho])


def equilibrium(rho, u):
    """lbm/basic.py:114-135."""
    shp = (9,) + (1,) * rho.dim()
    cx, cy, w = C[:, 0].reshape(shp), C[:, 1].reshape(shp), W.reshape(shp)
    uc = u[0][None] * cx + u[1][None] * cy
    usq = (u * u).sum(0)
    return rho[None] * w * (1 + 3 * uc + 4.5 * uc ** 2 - 1.5 * usq[None])


def guo_term(g, u):
    """forcing/guo.py:6-33."""
    shp = (9,) + (1,) * (u.dim() - 1)
    cx, cy, w = C[:, 0].reshape(shp), C[:, 1].reshape(shp), W.reshape(shp)
    uc = u[0][None] * cx + u[1][None] * cy
    return w * (g[0][None] * (3 * (cx - u[0][None]) + 9 * (uc * cx)) + g[1][None] * (3 * (cy - u[1][None]) + 9 * (uc * cy)))


def streaming(f):
    """lbm/basic.py:44-85: f_i(x) <- f_i(x - c_i), periodic."""
    return torch.stack([torch.roll(f[i], shifts=(int(LT.C2[i, 0]), int(LT.C2[i, 1])), dims=(0, 1)) for i in range(9)])


def walls_inlet_eq_outlet_pressure_nee(fs, u_in):
    """boundary_equilibrium(left, ux_wall=u_in), then boundary_pressure_nee(right, rho_wall=1)
    (lbm/boundary/eq.py:25-56, nee.py:19-65, _helpers.py:98-132)."""
    ny = fs.shape[2]
    left = equilibrium(torch.ones(ny, dtype=fs.dtype), torch.stack([u_in * torch.ones(ny, dtype=fs.dtype),
                                                                    torch.zeros(ny, dtype=fs.dtype)]))
    fs = torch.cat([left[:, None, :], fs[:, 1:, :]], dim=1)
    s = LT.WALLS2["right"]
    wall, nb = fs[:, -1, :], fs[:, -2, :]
    t0, t1 = s["tan"]
    o0, o1, o2 = s["out_dirs"]
    num = wall[0] + wall[t0] + wall[t1] + 2 * (wall[o0] + wall[o1] + wall[o2])
    un = s["sign"] * (1 - num / 1.0)
    rho_nb = nb.sum(0)
    p, n = s["pos"], s["neg"]
    ut = (nb[p[0]] - nb[n[0]] + nb[p[1]] - nb[n[1]] + nb[p[2]] - nb[n[2]]) / rho_nb
    r_nb, u_nb = macroscopic(nb)
    newwall = equilibrium(torch.ones(ny, dtype=fs.dtype), torch.stack([un, ut])) + (nb - equilibrium(r_nb, u_nb))
    return torch.cat([fs[:, :-1, :], newwall[:, None, :]], dim=1)


def phi4(r):
    """ib/kernels.py:25-43 (Peskin 4-point)."""
    a = r.abs()
    inner = (3 - 2 * a + torch.sqrt(torch.clamp(1 + 4 * a - 4 * a ** 2, min=0))) * 0.125
    outer = (5 - 2 * a - torch.sqrt(torch.clamp(-7 + 12 * a - 4 * a ** 2, min=0))) * 0.125
    return torch.where(a > 2, torch.zeros_like(a), torch.where(a < 1, inner, outer))


def stencil(rx, ry, S):
    """ib/stencil.py:7-51 with radius 2: weights [M, 16], flat indices [M, 16] (x-major)."""
    off = torch.arange(-1, 3, dtype=torch.float64)
    refx, refy = torch.floor(rx.detach()), torch.floor(ry.detach())
    sx = refx[:, None] + off[None]                      # [M, 4]
    sy = refy[:, None] + off[None]
    wx, wy = phi4(sx - rx[:, None]), phi4(sy - ry[:, None])
    Wm = (wx[:, :, None] * wy[:, None, :]).reshape(-1, 16)
    I = (sx[:, :, None] * S + sy[:, None, :]).reshape(-1, 16).long()
    assert int(I.min()) >= 0 and int(I.max()) < S * S, "stencil leaves the window"
    return Wm, I


def interp(grid, Wm, I):
    """ib/stencil.py:54-78: grid [2, S*S] -> [M, 2]."""
    return torch.stack([(Wm * grid[c][I]).sum(1) for c in range(2)], dim=1)


def spread(F, Wm, I, n):
    """ib/stencil.py:81-110: F [M, 2] -> grid [2, n]."""
    out = []
    for c in range(2):
        v = (F[:, c:c + 1] * Wm).reshape(-1)
        out.append(torch.zeros(n, dtype=F.dtype).index_add(0, I.reshape(-1), v))
    return torch.stack(out)


class VIVModel:
    """The VIV case of ``oracle.steps.VIV2D`` with differentiable (delta, K, u_in)."""

    def __init__(self, case):
        self.case = case
        self.op = torch.tensor(L2.get_mrt_collision_operator(case.omega))
        self.fop = torch.tensor(L2.get_mrt_forcing_operator(case.omega))
        th = 2 * np.pi * np.arange(case.n_markers) / case.n_markers
        self.cos, self.sin = torch.tensor(np.cos(th)), torch.tensor(np.sin(th))

    def markers(self, delta):
        c = self.case
        R = 0.5 * c.D + delta
        return c.cx + R * self.cos, c.cy + R * self.sin

    def step(self, f, d, v, a, delta, stiffness, u_in):
        c = self.case
        stiffness = torch.as_tensor(stiffness).to(torch.float64)
        u_in = torch.as_tensor(u_in).to(torch.float64)
        S = c.ib_size
        rho, u = macroscopic(f)
        # IB window (origin truncated from the displacement, no gradient through it)
        x0 = int(math.trunc(c.ib_x0 + float(d[0].detach())))
        y0 = int(math.trunc(c.ib_y0 + float(d[1].detach())))
        xc, yc = min(max(x0, 0), c.nx - S), min(max(y0, 0), c.ny - S)
        mx, my = self.markers(delta)
        Wm, I = stencil(mx + d[0] - x0, my + d[1] - y0, S)
        uw = u[:, xc:xc + S, yc:yc + S].reshape(2, S * S)
        ub = interp(uw, Wm, I)
        target = v[None, :].expand(c.n_markers, 2)
        ds2 = 2 * c.ds
        F = torch.zeros_like(ub)
        for _ in range(c.n_iter):
            dF = (target - ub) * ds2
            F = F + dF
            ub = ub + interp(0.5 * spread(dF, Wm, I, S * S), Wm, I)
        gw = spread(F, Wm, I, S * S).reshape(2, S, S)
        H = -F.sum(0) + a * c.area
        # Newmark (dyn.py:5-51), damping 0, transverse DOF only
        dt, gam, bet = 1.0, 0.5, 0.25
        v1 = v + dt * (1 - gam) * a
        v2 = d + dt * v + dt * dt * (0.5 - bet) * a
        an = (H - c.damping * v1 - stiffness * v2) / (c.mass + gam * dt * c.damping + bet * dt * dt * stiffness)
        mask = torch.tensor([float(c.dof[0]), float(c.dof[1])], dtype=torch.float64)
        a_n = an * mask
        v_n = (gam * dt * an + v1) * mask
        d_n = (bet * dt * dt * an + v2) * mask
        # collision + Guo forcing (MRT), streaming, walls
        g = torch.nn.functional.pad(gw, (yc, c.ny - yc - S, xc, c.nx - xc - S))   # window placed in the grid
        uc = u + g * 0.5 / rho
        feq = equilibrium(rho, uc)
        fp = f + torch.tensordot(self.op, feq - f, dims=([1], [0]))
        fp = fp + torch.tensordot(self.fop, guo_term(g, uc), dims=([1], [0]))
        fn = walls_inlet_eq_outlet_pressure_nee(streaming(fp), u_in)
        return fn, d_n, v_n, a_n, H

    def objective(self, f0, steps, delta, stiffness, u_in, v0=(0.0, 0.0)):
        """J = (1/T) sum_t d_y(t)^2 and the displacement history."""
        f = f0
        d = torch.zeros(2, dtype=torch.float64)
        v = torch.tensor(list(v0), dtype=torch.float64)
        a = torch.zeros(2, dtype=torch.float64)
        J = torch.zeros((), dtype=torch.float64)
        hist = []
        for _ in range(steps):
            f, d, v, a, H = self.step(f, d, v, a, delta, stiffness, u_in)
            J = J + d[1] ** 2
            hist.append(d)
        return J / steps, torch.stack(hist)
