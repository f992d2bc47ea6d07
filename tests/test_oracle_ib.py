"""Pins the IB and structural oracle (oracle/ib.py, oracle/dyn.py)."""

import math

import numpy as np
import pytest

from oracle import ib as IB
from oracle import dyn as DYN

rng = np.random.default_rng(7)


@pytest.mark.parametrize("name,support", [("peskin_2pt", 1), ("peskin_3pt", 1.5),
                                          ("peskin_4pt", 2), ("cosine_4pt", 2)])
def test_kernel_partition_of_unity(name, support):
    k = IB.KERNELS[name]
    for r in rng.random(50):
        js = np.arange(-4, 6)
        assert np.isclose(k(r - js).sum(), 1.0, atol=1e-12)
    assert k(np.array([support + 1e-9]))[0] == 0


def test_kernel_point_values_and_moments():
    k4, k3 = IB.kernel_peskin_4pt, IB.kernel_peskin_3pt
    assert np.allclose(k4(np.array([0.0, 1.0, 2.0, -1.0])), [0.5, 0.25, 0.0, 0.25])
    assert np.allclose(k3(np.array([0.0, 1.0, 1.5])), [2 / 3, 1 / 6, 0.0])
    for r in rng.random(20):
        js = np.arange(-3, 5)
        assert np.isclose(((r - js) * k4(r - js)).sum(), 0, atol=1e-12)      # first moment
        assert np.isclose((k4(r - js) ** 2).sum(), 3 / 8, atol=1e-12)        # sum of squares
        assert np.isclose(((r - js) * k3(r - js)).sum(), 0, atol=1e-12)


def test_stencil_bruteforce():
    nx, ny = 20, 17
    mx = 5 + 10 * rng.random(9)
    my = 5 + 7 * rng.random(9)
    W, I = IB.get_ib_stencil(mx, my, ny)
    assert I.dtype == np.int32 and W.shape == (9, 16)
    for m in range(9):
        for i in range(4):
            for j in range(4):
                sx = math.floor(mx[m]) - 1 + i
                sy = math.floor(my[m]) - 1 + j
                assert I[m, 4 * i + j] == sx * ny + sy
                assert np.isclose(W[m, 4 * i + j], IB.kernel_peskin_4pt(np.array(sx - mx[m]))
                                  * IB.kernel_peskin_4pt(np.array(sy - my[m])))
    assert np.allclose(W.sum(1), 1)
    # 3-point stencil with radius 2 still spans offsets -1..2 (stencil.py:28-30)
    W3, I3 = IB.get_ib_stencil(mx, my, ny, IB.kernel_peskin_3pt, 2)
    assert np.array_equal(I3, I) and np.allclose(W3.sum(1), 1)


def test_stencil3d_bruteforce_and_errors():
    shape = (9, 10, 11)
    mc = 3 + 4 * rng.random((5, 3))
    W, I = IB.get_ib_stencil3d(mc, shape)
    for m in range(5):
        for k in range(64):
            i, j, l = k // 16, (k // 4) % 4, k % 4
            sx, sy, sz = (math.floor(mc[m, 0]) - 1 + i, math.floor(mc[m, 1]) - 1 + j, math.floor(mc[m, 2]) - 1 + l)
            assert I[m, k] == sx * 110 + sy * 11 + sz
    assert np.allclose(W.sum(1), 1)
    with pytest.raises(ValueError):
        IB.get_ib_stencil3d(mc, (9, 10))
    with pytest.raises(ValueError):
        IB.get_ib_stencil3d(mc[:, :2], shape)


def test_interp_spread_identities():
    nx, ny = 24, 20
    mx = 6 + 10 * rng.random(15)
    my = 6 + 8 * rng.random(15)
    W, I = IB.get_ib_stencil(mx, my, ny)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    lin = np.stack([0.3 + 0.01 * X - 0.02 * Y, np.full((nx, ny), 0.7)])
    got = IB.interpolate(lin, W, I)
    assert np.allclose(got[:, 0], 0.3 + 0.01 * mx - 0.02 * my) and np.allclose(got[:, 1], 0.7)
    F = rng.standard_normal((15, 2))
    g = IB.spread(F, np.zeros((2, nx, ny)), W, I)
    assert np.allclose(g.reshape(2, -1).sum(1), F.sum(0))
    u = rng.standard_normal((2, nx, ny))
    assert np.isclose((g * u).sum(), (F * IB.interpolate(u, W, I)).sum())     # adjoint identity
    base = rng.standard_normal((2, nx, ny))
    assert np.allclose(IB.spread(F, base, W, I), base + g)


def test_out_of_range_semantics():
    grid = np.arange(12, dtype=float).reshape(1, 3, 4)
    W = np.ones((1, 3))
    I = np.array([[-1, 12, 5]], dtype=np.int32)
    assert np.allclose(IB.interpolate(grid, W, I), [[11 + 11 + 5]])       # wrap, clamp, in-range
    out = IB.spread(np.ones((1, 1)), np.zeros((1, 3, 4)), W, I)
    expect = np.zeros(12)
    expect[11] += 1
    expect[5] += 1
    assert np.allclose(out.ravel(), expect)                                 # wrap, drop, in-range


def test_mdf_isolated_marker_closed_form():
    """One marker: interp(spread(dF)) = dF * sum(w^2) = dF * (3/8)^2 for Peskin-4, so
    ub_n = U - (U - ub_0) (1 - 2 ds * 0.5 * 9/64)^n."""
    nx = ny = 16
    u = np.zeros((2, nx, ny))
    u[0] = 0.04
    u[1] = -0.01
    W, I = IB.get_ib_stencil(np.array([7.3]), np.array([8.6]), ny)
    U = np.array([[0.0, 0.02]])
    ds = 0.7
    for n in range(5):
        g, h = IB.multi_direct_forcing(u, W, I, U, ds, n)
        q = 1 - ds * 9 / 64
        ub_n = U - (U - np.array([[0.04, -0.01]])) * q ** n
        F = -h
        ub_from_F = np.array([[0.04, -0.01]]) + 0.5 * F * 9 / 64
        assert np.allclose(ub_from_F, ub_n)
        assert np.allclose(g.reshape(2, -1).sum(1), F[0])
    g, h = IB.multi_direct_forcing(u, W, I, U, ds, 0)
    assert np.all(g == 0) and np.all(h == 0)


def test_mdf_one_iter_bruteforce():
    nx, ny = 18, 16
    u = 0.05 * rng.standard_normal((2, nx, ny))
    mx, my = 5 + 8 * rng.random(6), 5 + 6 * rng.random(6)
    W, I = IB.get_ib_stencil(mx, my, ny)
    U = 0.01 * rng.standard_normal((6, 2))
    ds = 0.3 + rng.random(6)
    g, h = IB.multi_direct_forcing(u, W, I, U, ds, 1)
    flat = u.reshape(2, -1)
    ub = np.array([[sum(W[m, s] * flat[c, I[m, s]] for s in range(16)) for c in range(2)] for m in range(6)])
    F = (U - ub) * 2 * ds[:, None]
    gg = np.zeros((2, nx * ny))
    for m in range(6):
        for s in range(16):
            gg[:, I[m, s]] += F[m] * W[m, s]
    assert np.allclose(h, -F) and np.allclose(g.reshape(2, -1), gg)


def test_geometry_2d():
    N, R = 64, 10.0
    th = 2 * np.pi * np.arange(N) / N
    p = np.stack([3 + R * np.cos(th), 4 + R * np.sin(th)], 1)
    assert np.isclose(IB.get_area(p), 0.5 * N * R ** 2 * np.sin(2 * np.pi / N))
    assert np.allclose(IB.get_ds(p), 2 * R * np.sin(np.pi / N))
    line = np.stack([np.arange(5.0), np.zeros(5)], 1)
    assert np.allclose(IB.get_ds(line, closed=False), [0.5, 1, 1, 1, 0.5])


def test_geometry_3d():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    f = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    assert np.isclose(IB.get_volume(v, f), 1 / 6)
    area = 1.5 + np.sqrt(3) / 2
    assert np.isclose(IB.get_surface_area(v, f), area)
    assert np.isclose(IB.get_ds3d(v, f).sum(), area)


def test_newmark_closed_form():
    """Undamped free SDOF: the average-acceleration rule is a rotation by
    Omega = 2 atan(w dt / 2): d_n = d0 cos(n Omega)."""
    m, k, d0 = 3.0, 0.7, 0.4
    w = math.sqrt(k / m)
    Om = 2 * math.atan(w / 2)
    a, v, d = -k * d0 / m, 0.0, d0
    for n in range(1, 60):
        a, v, d = DYN.newmark(a, v, d, 0.0, m, k, 0.0)
        assert np.isclose(d, d0 * math.cos(n * Om), atol=1e-12)
        assert np.isclose(0.5 * m * v * v + 0.5 * k * d * d, 0.5 * k * d0 * d0)
    # matrix form equals per-DOF scalars for diagonal systems
    A, V, D = np.array([0.1, -0.2]), np.array([0.3, 0.0]), np.array([0.0, 0.5])
    H = np.array([0.05, 0.01])
    mm, kk, cc = np.diag([2.0, 3.0]), np.diag([0.5, 0.7]), np.diag([0.1, 0.0])
    r1 = DYN.newmark(A, V, D, H, mm, kk, cc)
    for i in range(2):
        r2 = DYN.newmark(A[i], V[i], D[i], H[i], mm[i, i], kk[i, i], cc[i, i])
        for x, y in zip(r1, r2):
            assert np.isclose(x[i], y)


def test_kinematics():
    th = 2 * np.pi * np.arange(8) / 8
    x, y = 5 + np.cos(th), 6 + np.sin(th)
    d = np.array([0.1, -0.2, 0.3])
    xr, yr = DYN.get_markers_coords_3dof(x, y, 5.0, 6.0, d)
    assert np.allclose(np.hypot(xr - 5.1, yr - 5.8), 1)
    assert np.allclose(np.arctan2(yr - 5.8, xr - 5.1), np.arctan2(np.sin(th + 0.3), np.cos(th + 0.3)))
    v = np.array([0.0, 0.0, 2.0])
    vm = DYN.get_markers_velocity_3dof(xr, yr, 5.0, 6.0, d, v)
    assert np.allclose(vm[:, 0] * (xr - 5.1) + vm[:, 1] * (yr - 5.8), 0)
    ht = np.stack([-np.sin(th), np.cos(th)], 1)           # unit tangential forces on a unit circle
    assert np.isclose(DYN.get_torque_to_obj(x, y, 5.0, 6.0, np.zeros(3), ht), 8.0)
    assert np.allclose(DYN.get_force_to_obj(ht), 0)


def test_markers_velocity_3dof_is_time_derivative_of_coords():
    """Pins dyn.py:103-120 against the kinematics it must satisfy: along the
    rigid motion d(t) = d0 + v t, the marker coordinates of
    get_markers_coords_3dof (dyn.py:84-100) move with d/dt X = v + omega x r.
    A central difference in float64 checks sign, scale and the v0 / v1 terms;
    the closed-form part checks |v_m - v_c| = |omega| r, counter-clockwise for
    omega > 0."""
    th = 2 * np.pi * np.arange(12) / 12 + 0.1
    x0, y0 = 3.0 + 1.7 * np.cos(th), -2.0 + 1.7 * np.sin(th)
    d0 = np.array([0.4, -0.3, 0.25])
    v = np.array([0.07, -0.05, 0.9])
    h = 1e-6
    xp, yp = DYN.get_markers_coords_3dof(x0, y0, 3.0, -2.0, d0 + h * v)
    xm, ym = DYN.get_markers_coords_3dof(x0, y0, 3.0, -2.0, d0 - h * v)
    fd = np.stack([(xp - xm) / (2 * h), (yp - ym) / (2 * h)], axis=-1)
    xs, ys = DYN.get_markers_coords_3dof(x0, y0, 3.0, -2.0, d0)
    vm = DYN.get_markers_velocity_3dof(xs, ys, 3.0, -2.0, d0, v)
    assert np.abs(vm - fd).max() < 1e-8
    rel = vm - v[:2]
    rx, ry = xs - (3.0 + d0[0]), ys - (-2.0 + d0[1])
    assert np.allclose(np.hypot(rel[:, 0], rel[:, 1]), abs(v[2]) * np.hypot(rx, ry), rtol=1e-12)
    assert np.all(rx * rel[:, 1] - ry * rel[:, 0] > 0)      # omega > 0: counter-clockwise
    # pure translation: every marker moves with (v0, v1)
    vt = DYN.get_markers_velocity_3dof(xs, ys, 3.0, -2.0, d0, np.array([0.07, -0.05, 0.0]))
    assert np.allclose(vt, [0.07, -0.05], rtol=0, atol=0)
