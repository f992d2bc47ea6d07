"""Oracle restatement of ``vivsim.post`` (field diagnostics) and of the force
coefficients of the VIV example.  Test infrastructure only: the product never
imports it.

The reference differentiates with ``jnp.gradient`` (default edge_order=1:
central differences (u[k+1] - u[k-1]) / 2 inside, one-sided u[1] - u[0] and
u[n-1] - u[n-2] at the ends, unit spacing), which ``np.gradient`` restates.
"""

import numpy as np


def velocity_magnitude(u):
    """post.py:6-14."""
    return np.linalg.norm(u, axis=0)


def velocity_gradient(u):
    """G[i, j] = du_i / dx_j — post.py:17-29."""
    dim = u.shape[0]
    grads = np.gradient(u, axis=tuple(range(1, dim + 1)))
    return np.stack(grads, axis=1)


def vorticity(u):
    """2D: dv/dx - du/dy; 3D: curl — post.py:32-54."""
    if u.shape[0] == 2:
        return np.gradient(u[1], axis=0) - np.gradient(u[0], axis=1)
    dudy, dudz = np.gradient(u[0], axis=1), np.gradient(u[0], axis=2)
    dvdx, dvdz = np.gradient(u[1], axis=0), np.gradient(u[1], axis=2)
    dwdx, dwdy = np.gradient(u[2], axis=0), np.gradient(u[2], axis=1)
    return np.stack([dwdy - dvdz, dudz - dwdx, dvdx - dudy])


def vorticity_magnitude(u):
    """post.py:57-68."""
    w = vorticity(u)
    return np.abs(w) if u.shape[0] == 2 else np.linalg.norm(w, axis=0)


def divergence(u):
    """post.py:72-83."""
    return sum(np.gradient(u[i], axis=i) for i in range(u.shape[0]))


def strain_rate(u):
    """post.py:86-95."""
    G = velocity_gradient(u)
    return 0.5 * (G + np.swapaxes(G, 0, 1))


def strain_rate_magnitude(u):
    """post.py:98-107."""
    S = strain_rate(u)
    return np.sqrt(np.einsum("ij...,ij...->...", S, S))


def kinetic_energy(u):
    """post.py:110-118."""
    return 0.5 * np.sum(u ** 2, axis=0)


def mean_kinetic_energy(u):
    """post.py:121-129."""
    return np.mean(kinetic_energy(u))


def pressure(rho, cs2=1.0 / 3.0):
    """post.py:132-141."""
    return rho * cs2


def enstrophy(u):
    """post.py:144-155."""
    w = vorticity(u)
    return 0.5 * w ** 2 if u.shape[0] == 2 else 0.5 * np.sum(w ** 2, axis=0)


def mean_enstrophy(u):
    """post.py:158-166."""
    return np.mean(enstrophy(u))


def q_criterion(u):
    """Q = -0.5 tr(G G) — post.py:172-187."""
    G = velocity_gradient(u)
    return -0.5 * np.einsum("ij...,ji...->...", G, G)


def force_coefficients(h_hist, D, U0):
    """Cd, Cl = 2 h / (D U0^2) of a [n, 2] force history —
    examples/2d/vortex_induced_vibration.py:228-229."""
    h = np.asarray(h_hist)
    return h[:, 0] * 2 / (D * U0 ** 2), h[:, 1] * 2 / (D * U0 ** 2)
