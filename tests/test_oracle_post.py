"""Pins the diagnostics and grid-refinement oracles (oracle/post.py,
oracle/multigrid.py) by closed forms and brute-force loops."""

import numpy as np
import pytest

from oracle import multigrid as MG
from oracle import post as P

rng = np.random.default_rng(5)


def _grid2(nx=7, ny=6):
    return np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float), indexing="ij")


def test_rotation_and_strain_2d():
    """Solid-body rotation u = W (-(y - yc), x - xc): vorticity 2 W, Q = W^2,
    zero divergence and strain; pure strain u = (a x, -a y): Q = -a^2 < 0,
    |S| = sqrt(2) a, zero vorticity.  Linear fields make every difference
    (central and one-sided) exact."""
    x, y = _grid2()
    W, a = 0.3, 0.07
    rot = np.stack([-W * (y - 2.5), W * (x - 3.0)])
    assert np.allclose(P.vorticity(rot), 2 * W, atol=1e-14)
    assert np.allclose(P.q_criterion(rot), W * W, atol=1e-14)
    assert np.allclose(P.divergence(rot), 0, atol=1e-14)
    assert np.allclose(P.strain_rate_magnitude(rot), 0, atol=1e-14)
    assert np.allclose(P.enstrophy(rot), 2 * W * W, atol=1e-14)
    assert np.isclose(P.mean_enstrophy(rot), 2 * W * W)
    st = np.stack([a * x, -a * y])
    assert np.allclose(P.q_criterion(st), -a * a, atol=1e-14)
    assert np.allclose(P.strain_rate_magnitude(st), np.sqrt(2) * a, atol=1e-14)
    assert np.allclose(P.vorticity(st), 0, atol=1e-14) and np.allclose(P.divergence(st), 0, atol=1e-14)
    G = P.velocity_gradient(st)
    assert np.allclose(G[0, 0], a) and np.allclose(G[1, 1], -a) and np.allclose(G[0, 1], 0) and np.allclose(G[1, 0], 0)


def test_rotation_3d():
    """Rotation with angular velocity W about axis n: curl = 2 W n, Q = |W|^2."""
    x, y, z = np.meshgrid(np.arange(5.0), np.arange(6.0), np.arange(4.0), indexing="ij")
    Wv = np.array([0.1, -0.2, 0.25])
    r = np.stack([x - 2, y - 2.5, z - 1.5])
    u = np.cross(Wv[:, None, None, None], r, axis=0)
    w = P.vorticity(u)
    for k in range(3):
        assert np.allclose(w[k], 2 * Wv[k], atol=1e-14)
    assert np.allclose(P.q_criterion(u), Wv @ Wv, atol=1e-14)
    assert np.allclose(P.vorticity_magnitude(u), 2 * np.linalg.norm(Wv), atol=1e-14)


def test_gradient_brute_force_edges():
    """jnp.gradient's rule written as loops: (u[k+1] - u[k-1]) / 2 inside,
    u[1] - u[0] and u[n-1] - u[n-2] at the ends, along each axis."""
    u = rng.standard_normal((2, 5, 4))
    G = P.velocity_gradient(u)
    for i in range(2):
        for j in range(2):
            n = u.shape[1 + j]
            for x in range(5):
                for y in range(4):
                    idx = [x, y]
                    k = idx[j]

                    def at(m):
                        t = list(idx)
                        t[j] = m
                        return u[i, t[0], t[1]]
                    ref = at(1) - at(0) if k == 0 else (at(n - 1) - at(n - 2) if k == n - 1 else (at(k + 1) - at(k - 1)) / 2)
                    assert np.isclose(G[i, j, x, y], ref, rtol=0, atol=1e-15)


def test_quadratic_profile():
    """u_x = x^2: derivative 2x inside (exact), 1 and 2n - 3 at the ends."""
    x, y = _grid2(6, 3)
    u = np.stack([x ** 2, 0 * y])
    G = P.velocity_gradient(u)
    assert np.allclose(G[0, 0, 1:-1], 2 * x[1:-1]) and np.allclose(G[0, 0, 0], 1) and np.allclose(G[0, 0, -1], 9)


def test_energies_pressure_coefficients():
    u = rng.standard_normal((3, 3, 4, 2))
    assert np.allclose(P.kinetic_energy(u), 0.5 * (u ** 2).sum(0))
    assert np.allclose(P.velocity_magnitude(u), np.sqrt((u ** 2).sum(0)))
    assert np.isclose(P.mean_kinetic_energy(u), 0.5 * (u ** 2).sum() / u[0].size)
    assert np.allclose(P.pressure(np.array([3.0, 0.3])), [1.0, 0.1])
    cd, cl = P.force_coefficients(np.array([[2.0, -1.0], [0.5, 0.25]]), D=4.0, U0=0.5)
    assert np.allclose(cd, [4.0, 1.0]) and np.allclose(cl, [-2.0, 0.5])


@pytest.mark.parametrize("dir", ["left", "right", "up", "down"])
def test_fine_to_coarse_brute_force(dir):
    """Index arithmetic written out: the coarse node j of the buffer line gets
    the mean of the fine cells (2j, 2j+1) on the two fine lines next to the
    interface, for the three populations crossing it; nothing else changes."""
    nxc, nyc = 5, 4
    if dir in ("left", "right"):
        nxf, nyf = 6, 2 * nyc
    else:
        nxf, nyf = 2 * nxc, 7
    ff = rng.standard_normal((9, nxf, nyf))
    fc = rng.standard_normal((9, nxc, nyc))
    out = MG.fine_to_coarse(ff, fc, dir)
    dirs = {"left": [3, 6, 7], "right": [1, 5, 8], "up": [2, 5, 6], "down": [4, 7, 8]}[dir]
    exp = fc.copy()
    for q in dirs:
        if dir == "left":
            for j in range(nyc):
                exp[q, nxc - 1, j] = (ff[q, 0, 2 * j] + ff[q, 0, 2 * j + 1] + ff[q, 1, 2 * j] + ff[q, 1, 2 * j + 1]) / 4
        elif dir == "right":
            for j in range(nyc):
                exp[q, 0, j] = (ff[q, nxf - 1, 2 * j] + ff[q, nxf - 1, 2 * j + 1] + ff[q, nxf - 2, 2 * j]
                                + ff[q, nxf - 2, 2 * j + 1]) / 4
        elif dir == "up":
            for j in range(nxc):
                exp[q, j, 0] = (ff[q, 2 * j, nyf - 1] + ff[q, 2 * j + 1, nyf - 1] + ff[q, 2 * j, nyf - 2]
                                + ff[q, 2 * j + 1, nyf - 2]) / 4
        else:
            for j in range(nxc):
                exp[q, j, nyc - 1] = (ff[q, 2 * j, 0] + ff[q, 2 * j + 1, 0] + ff[q, 2 * j, 1] + ff[q, 2 * j + 1, 1]) / 4
    assert np.allclose(out, exp, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dir", ["left", "right", "up", "down"])
def test_coarse_to_fine_brute_force(dir):
    nxc, nyc = 5, 4
    if dir in ("left", "right"):
        nxf, nyf = 6, 2 * nyc
    else:
        nxf, nyf = 2 * nxc, 7
    ff = rng.standard_normal((9, nxf, nyf))
    fc = rng.standard_normal((9, nxc, nyc))
    out = MG.coarse_to_fine(fc, ff, dir)
    dirs = {"left": [3, 6, 7], "right": [1, 5, 8], "up": [2, 5, 6], "down": [4, 7, 8]}[dir]
    exp = ff.copy()
    for q in dirs:
        if dir == "left":
            for j in range(nyf):
                exp[q, nxf - 1, j] = fc[q, 0, j // 2]
        elif dir == "right":
            for j in range(nyf):
                exp[q, 0, j] = fc[q, nxc - 1, j // 2]
        elif dir == "up":
            for j in range(nxf):
                exp[q, j, 0] = fc[q, j // 2, nyc - 1]
        else:
            for j in range(nxf):
                exp[q, j, nyf - 1] = fc[q, j // 2, 0]
    assert np.array_equal(out, exp)


def test_refinement_omega_and_indices():
    """omega at level l equals 1 / (3 nu 2^l + 0.5): the viscosity in units of
    the level's lattice scales with 2^l when the time step scales with the
    cell size."""
    for nu in (0.01, 0.05, 0.2):
        for lev in (-2, -1, 0, 1, 2, 3):
            assert np.isclose(MG.get_omega(nu, lev), 1 / (3 * nu * 2 ** lev + 0.5), rtol=1e-13)
    assert MG.coord_to_indices(10.75, 4.5, 8, 2, level=2) == (11, 10)
    assert MG.coord_to_indices(10.75, 4.5, 8, 2, level=-1) == (1, 1)
    f, r, u = MG.init_grid(16, 8, level=1, buffer_x=2, buffer_y=1)
    assert f.shape == (9, 34, 17) and r.shape == (34, 17) and u.shape == (2, 34, 17) and np.all(r == 1)
