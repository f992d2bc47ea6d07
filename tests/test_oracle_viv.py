"""The oracle's VIV step composition (oracle/steps.py) pinned by properties that
do not restate its formulas: windowed IB equals full-grid IB, spread conserves
the marker force, the window start follows lax.dynamic_slice clamping, n_iter=0
reduces to the plain step, inactive DOFs stay at rest."""

import numpy as np
import pytest

from oracle import ib as IB
from oracle import lbm2d as L2
from oracle import steps as S


def small(**kw):
    base = dict(nx=96, ny=80, D=16.0, cx=40.0, cy=40.0, n_markers=64, U0=0.05, Re=60.0, n_iter=3, pad=6)
    base.update(kw)
    return S.VIV2D(**base)


def test_window_equals_full_grid_ib():
    """The IB on the window (relative marker coordinates) equals the IB on the
    whole grid with absolute coordinates, when the window lies inside the grid."""
    c = small()
    f = c.initial_f(np.float64)
    _, u = L2.get_macroscopic(f)
    d, v = np.array([0.3, -0.7]), np.array([0.0, 2e-3])
    xc, yc, g_win, h_win = S.viv_ib(c, u, d, v, np.float64)
    mx, my = c.markers()
    W, I = IB.get_ib_stencil(mx + d[0], my + d[1], c.ny, IB.KERNELS[c.kernel], 2)
    target = np.repeat(v[None, :], c.n_markers, axis=0)
    g_full, h_full = IB.multi_direct_forcing(u, W, I, target, c.ds, c.n_iter)
    S_ = c.ib_size
    assert np.allclose(h_win, h_full, rtol=1e-12, atol=1e-15)
    assert np.allclose(g_win, g_full[:, xc:xc + S_, yc:yc + S_], rtol=1e-12, atol=1e-15)
    outside = g_full.copy()
    outside[:, xc:xc + S_, yc:yc + S_] = 0
    assert np.all(outside == 0)


def test_spread_conserves_marker_force():
    """Each marker's kernel weights sum to 1 inside the window, so the spread
    force sums to the total marker force (= minus the fluid force on the body)."""
    c = small()
    f = c.initial_f(np.float64)
    _, u = L2.get_macroscopic(f)
    _, _, g_win, h_m = S.viv_ib(c, u, np.zeros(2), np.array([0.0, 1e-3]), np.float64)
    assert np.allclose(g_win.sum(axis=(1, 2)), -h_m.sum(axis=0), rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("x0,y0,expect", [(10, 20, (10, 20)), (-4, 20, (0, 20)), (90, -3, (68, 0)),
                                          (70, 60, (68, 52))])
def test_window_clamp(x0, y0, expect):
    c = small()   # window 28, grid 96 x 80
    assert S.viv_window_clamp(c, x0, y0) == expect


def test_window_origin_truncates_toward_zero():
    """int32 conversion truncates: floor for positive sums, ceil for negative ones."""
    c = small()
    assert S.viv_window_origin(c, np.array([0.9, -0.9])) == (c.ib_x0, c.ib_y0 - 1)
    assert S.viv_window_origin(c, np.array([-1.5, 2.5])) == (c.ib_x0 - 2, c.ib_y0 + 2)
    assert S.viv_window_origin(c, np.array([-c.ib_x0 - 0.5, -c.ib_y0 - 1.7])) == (0, -1)


def test_clamped_window_step_runs():
    """A body near the inlet and across the periodic y edge: the clamped window
    keeps every slice inside the grid and the step stays finite."""
    c = small(cx=10.0, cy=6.0)
    f = c.initial_f(np.float64)
    f2, d, v, a, dh, hh = S.viv_run(c, f, 3, v=np.array([0.0, 1e-3]))
    assert np.all(np.isfinite(f2)) and f2.shape == f.shape
    assert np.all(np.isfinite(hh))


def test_zero_iterations_is_plain_step():
    c = small(n_iter=0)
    f = c.initial_f(np.float64)
    f_viv = S.viv_run(c, f, 3)[0]
    op, fop = L2.get_mrt_collision_operator(c.omega), L2.get_mrt_forcing_operator(c.omega)
    f_plain = f
    for _ in range(3):
        f_plain = S.lbm_step2d(f_plain, c.omega, c.walls(), c.collision, c.forcing, None, op, fop)
    assert np.allclose(f_viv, f_plain, rtol=0, atol=1e-15)


def test_inactive_dof_stays_at_rest():
    c = small(dof=(False, True))
    f = c.initial_f(np.float64)
    _, d, v, a, dh, hh = S.viv_run(c, f, 4, v=np.array([0.0, 1e-3]))
    assert d[0] == 0 and v[0] == 0 and a[0] == 0 and np.all(dh[:, 0] == 0)
    assert np.any(dh[:, 1] != 0)
    assert np.any(hh[:, 0] != 0)   # the drag is still reported
