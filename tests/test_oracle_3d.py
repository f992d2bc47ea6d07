"""Pins the D3Q19 oracle (oracle/lbm3d.py).  Literal rows / entries quoted from
the reference text serve as known answers where the reference states a table."""

import numpy as np
import pytest

from oracle import lattice as LT
from oracle import lbm3d as L

rng = np.random.default_rng(99)


def rand_f(shape=(4, 5, 3)):
    rho = 1 + 0.05 * rng.standard_normal(shape)
    u = 0.04 * rng.standard_normal((3,) + shape)
    feq = L.get_equilibrium(rho, u)
    return feq + 1e-3 * rng.standard_normal(feq.shape)


def moments(f):
    c = LT.C3.astype(float)
    return f.sum(0), np.tensordot(c.T, f, axes=([1], [0])), np.einsum("qa,qb,q...->ab...", c, c, f)


def test_tables_and_isotropy():
    assert LT.check_tables()
    w, c = LT.W3, LT.C3.astype(float)
    assert np.isclose(w.sum(), 1) and np.allclose(w @ c, 0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(3) / 3)
    d = np.eye(3)
    iso4 = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(np.einsum("q,qa,qb,qc,qd->abcd", w, c, c, c, c), iso4)
    assert list(LT.OPP3) == [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]


def test_streaming_bruteforce():
    f = rand_f((4, 3, 5))
    s = L.streaming(f)
    nx, ny, nz = f.shape[1:]
    for i in range(19):
        cx, cy, cz = LT.C3[i]
        for x in range(nx):
            for y in range(ny):
                for z in range(nz):
                    assert s[i, x, y, z] == f[i, (x - cx) % nx, (y - cy) % ny, (z - cz) % nz]


def test_equilibrium_moments():
    rho = 1 + 0.1 * rng.standard_normal((3, 2, 2))
    u = 0.1 * rng.standard_normal((3, 3, 2, 2))
    m0, m1, m2 = moments(L.get_equilibrium(rho, u))
    assert np.allclose(m0, rho) and np.allclose(m1, rho * u)
    assert np.allclose(m2, rho * (np.eye(3)[:, :, None, None, None] / 3 + u[:, None] * u[None, :]))


def test_mrt_basis_rows_match_reference_text():
    M = L.M_MRT
    # rows quoted from lbm3d/collision/mrt.py:9-28
    assert list(M[5]) == [0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2]
    assert list(M[10]) == [0] * 7 + [1, 1, -1, -1] + [0] * 8
    assert list(M[12]) == [0] * 7 + [1, -1, 1, -1] + [0] * 8
    assert list(M[13]) == [0] * 15 + [1, 1, -1, -1]
    assert list(M[14]) == [0] * 11 + [1, -1, 1, -1] + [0] * 4
    assert list(M[17]) == [0] * 11 + [1, 1, 1, 1] + [0] * 4
    assert abs(np.linalg.det(M)) > 0


def test_mrt_operator():
    om = 1.9
    op = L.get_mrt_collision_operator(om)
    assert np.allclose(L.M_MRT[:4] @ op, 0, atol=1e-12)
    rates = L.mrt_relaxation_rates(om)
    Minv = np.linalg.inv(L.M_MRT)
    for k in range(19):
        assert np.allclose(op @ Minv[:, k], rates[k] * Minv[:, k], atol=1e-12)
    assert (rates[5:10] == om).all()
    fop = L.get_mrt_forcing_operator(om)
    for k in range(19):
        assert np.allclose(fop @ Minv[:, k], (1 - rates[k] / 2) * Minv[:, k], atol=1e-12)
    assert np.allclose(L.get_mrt_collision_operator(om, np.float32), op, atol=2e-5)


def test_projection_matches_reference_entries_and_moments():
    P = L.PROJ3
    # entries quoted from lbm3d/collision/reg.py:12-31
    assert np.allclose(P[0], [0] + [-0.5] * 6 + [-1] * 12)
    assert np.allclose(P[1], [0, 1 / 6, 1 / 6] + [-1 / 12] * 4 + [1 / 12] * 8 + [-1 / 6] * 4)
    assert np.allclose(P[7, 7:11], [5 / 12, -1 / 12, -1 / 12, 5 / 12])
    assert np.allclose(P[11, 11:15], [5 / 12, -1 / 12, -1 / 12, 5 / 12])
    assert np.allclose(P[16, 15:19], [-1 / 12, 5 / 12, 5 / 12, -1 / 12])
    assert np.allclose(P[15, [1, 3, 5, 7]], [-1 / 24, 1 / 12, 1 / 12, 1 / 24])
    fneq = 1e-3 * rng.standard_normal((19, 4))
    pr = P @ fneq
    m0, m1, m2 = moments(pr)
    _, _, p2 = moments(fneq)
    assert np.allclose(m0, 0) and np.allclose(m1, 0) and np.allclose(m2, p2)
    assert np.allclose(P @ pr, pr)


def test_reg_kbc_conserve():
    f = rand_f()
    rho, u = L.get_macroscopic(f)
    feq = L.get_equilibrium(rho, u)
    for fn in (L.collision_reg, L.collision_kbc, L.collision_bgk):
        m0, m1, _ = moments(fn(f, feq, 1.85))
        assert np.allclose(m0, rho) and np.allclose(m1, rho * u)
    # KBC: post-collision fneq is 1/feq-orthogonal to the high-order part
    post = L.collision_kbc(f, feq, 1.85)
    s = L.get_second_order_projection(f - feq)
    h = (f - feq) - s
    assert np.allclose((h * (post - feq) / feq).sum(0), 0, atol=1e-15)


def test_guo_moments():
    u = 0.05 * rng.standard_normal((3, 2, 3, 2))
    g = 1e-3 * rng.standard_normal((3, 2, 3, 2))
    F = L.get_guo_forcing_term(g, u)
    m0, m1, m2 = moments(F)
    assert np.allclose(m0, 0, atol=1e-16) and np.allclose(m1, g, atol=1e-16)
    assert np.allclose(m2, u[:, None] * g[None] + g[:, None] * u[None], atol=1e-16)
    rho = 1 + 0.1 * rng.standard_normal((2, 3, 2))
    edm = L.get_equilibrium(rho, u + g / (2 * rho)) - L.get_equilibrium(rho, u - g / (2 * rho))
    assert np.allclose(F, edm, atol=1e-16)


FACES = ["left", "right", "bottom", "top", "back", "front"]


def face(f, loc, key="idx"):
    s = LT.WALLS3[loc]
    sl = [slice(None)] * 4
    sl[1 + s["axis"]] = s[key]
    return f[tuple(sl)]


@pytest.mark.parametrize("loc", FACES)
def test_nebb_velocity_3d(loc):
    fs = L.streaming(rand_f())
    out = L.boundary_velocity_nebb(fs, loc, ux_wall=0.03, uy_wall=0.01, uz_wall=-0.02)
    r, u = L.get_macroscopic(face(out, loc))
    rho_w = L.get_rho_wall_from_velocity(fs, loc, 0.03, 0.01, -0.02)
    s = LT.WALLS3[loc]
    assert np.allclose(r, rho_w)
    assert np.allclose(u[s["axis"]], [0.03, 0.01, -0.02][s["axis"]])


@pytest.mark.parametrize("loc", FACES)
def test_bounce_back_pairs_by_position(loc):
    pre = rand_f()
    fs = L.streaming(pre)
    s = LT.WALLS3[loc]
    out = L.boundary_bounce_back(pre, fs, loc)
    fo, fp = face(out, loc), face(pre, loc)
    for i, o in zip(s["in_dirs"], s["out_dirs"]):
        assert np.allclose(fo[i], fp[o])
    assert np.allclose(L.boundary_specular_reflection(pre, fs, loc, 0.01, 0.02, 0.03),
                       L.boundary_bounce_back(pre, fs, loc, 0.01, 0.02, 0.03))
    # moving wall: correction 2 w rho_pre (c.u)/cs2
    out = L.boundary_bounce_back(pre, fs, loc, ux_wall=0.02)
    rho = fp.sum(0)
    for i, o in zip(s["in_dirs"], s["out_dirs"]):
        assert np.allclose(face(out, loc)[i], fp[o] + 6 * LT.W3[i] * rho * LT.C3[i, 0] * 0.02)


@pytest.mark.parametrize("loc", FACES)
def test_nee_eq_3d(loc):
    f = rand_f()
    out = L.boundary_equilibrium(f, loc, rho_wall=1.01, ux_wall=0.02)
    r, u = L.get_macroscopic(face(out, loc))
    assert np.allclose(r, 1.01) and np.allclose(u[0], 0.02) and np.allclose(u[1:], 0)
    out = L.boundary_pressure_nee(f, loc, rho_wall=1.0)
    nb = face(f, loc, "nb")
    rn, un = L.get_macroscopic(nb)
    uw = list(L.get_wall_velocity_from_pressure(f, loc, 1.0))
    shp = rn.shape
    expect = L.get_equilibrium(np.ones(shp), np.stack([np.broadcast_to(v, shp) for v in uw])) \
        + nb - L.get_equilibrium(rn, un)
    assert np.allclose(face(out, loc), expect)


def test_characteristic_3d():
    rho = np.full((4, 5, 3), 1.02)
    u = np.zeros((3, 4, 5, 3)) + np.array([0.01, 0.02, 0.03])[:, None, None, None]
    for loc in FACES:
        r, uu = L.boundary_characteristic(rho, u, loc)
        assert np.allclose(r, 1.02) and np.allclose(uu[0], 0.01) and np.allclose(uu[2], 0.03)
    with pytest.raises(ValueError):
        L.boundary_characteristic(rho, u, "north")


# ------------------------------------------------------------------ sphere step composition

def _small_sphere(**kw):
    from oracle import steps as S
    base = dict(nx=30, ny=24, nz=22, D=8.0, cx=12.0, cy=12.0, cz=11.0, n_markers=200, U0=0.05, Re=40.0, n_iter=2,
                pad=3)
    base.update(kw)
    return S.Sphere3D(**base)


def test_sphere_window_equals_full_grid_ib():
    """IB on the window with relative coordinates == IB on the whole grid with
    absolute coordinates (the window contains every stencil point)."""
    from oracle import ib as IB
    c = _small_sphere()
    rng = np.random.default_rng(7)
    u = 0.05 * rng.standard_normal((3, c.nx, c.ny, c.nz))
    W, I = c.stencil(np.float64)
    S_ = c.ib_size
    x0, y0, z0 = c.ib_origin
    tgt = np.zeros((c.n_markers, 3))
    gw, hw = IB.multi_direct_forcing(u[:, x0:x0 + S_, y0:y0 + S_, z0:z0 + S_], W, I, tgt, c.ds, c.n_iter)
    Wf, If = IB.get_ib_stencil3d(c.markers(), (c.nx, c.ny, c.nz), IB.KERNELS[c.kernel], 2)
    gf, hf = IB.multi_direct_forcing(u, Wf, If, tgt, c.ds, c.n_iter)
    assert np.allclose(hw, hf, rtol=1e-12, atol=1e-16)
    assert np.allclose(gw, gf[:, x0:x0 + S_, y0:y0 + S_, z0:z0 + S_], rtol=1e-12, atol=1e-16)
    assert np.allclose(gw.sum(axis=(1, 2, 3)), -hw.sum(axis=0), rtol=1e-10, atol=1e-15)


def test_sphere_zero_iterations_is_plain_step():
    from oracle import steps as S
    c = _small_sphere(n_iter=0)
    f = c.initial_f(np.float64)
    f1, h = S.sphere_step(c, f)
    f2 = S.lbm_step3d(f, c.omega, c.walls(), c.collision, c.forcing, None)
    assert np.allclose(f1, f2, rtol=0, atol=1e-15) and np.all(h == 0)


def test_sphere_drag_points_downstream():
    """After a few steps of uniform inflow the fluid force on the markers
    opposes the flow (h = -F has the sign of the drag on the body, +x)."""
    from oracle import steps as S
    c = _small_sphere()
    f = c.initial_f(np.float64)
    for _ in range(3):
        f, h = S.sphere_step(c, f)
    assert h[0] > 0 and abs(h[1]) < 1e-2 * h[0] and abs(h[2]) < 1e-2 * h[0]


FACES3 = ("left", "right", "bottom", "top", "back", "front")


@pytest.mark.parametrize("loc", FACES3)
def test_characteristic_invariants_3d(loc):
    """Pins lbm3d/boundary/cbc.py:16-51 the way the 2D test pins cbc.py: with
    lambda = u_n - s cs, du_n - (cs / rho) drho = 0 and
    du_n + (cs / rho) drho = -lambda (u_n' - s cs / rho rho') with the exact
    derivatives of quadratic profiles; both tangential components are copied
    from the face."""
    from tests.test_oracle_2d import _cbc_profiles
    s = LT.WALLS3[loc]
    na, sg, ta = s["axis"], s["sign"], s["tan_axes"]
    shape = (6, 5, 4)
    x, t, rho, un, drho, dun = _cbc_profiles(shape, na)
    u = np.zeros((3,) + shape)
    u[na] = un
    u[ta[0]] = 0.01 + 0.003 * t + 0.005 * x
    u[ta[1]] = -0.02 + 0.001 * t - 0.004 * x
    r_next, u_next = L.boundary_characteristic(rho, u, loc)
    w = 0 if sg > 0 else -1
    take = lambda a: np.take(a, w, axis=na)
    r1, n1, dr, dn = take(rho), take(un), take(drho), take(dun)
    cs = np.sqrt(1 / 3)
    d_rho, d_un = r_next - r1, u_next[na] - n1
    assert np.abs(d_un - cs / r1 * d_rho).max() < 1e-13
    assert np.abs((d_un + cs / r1 * d_rho) + (n1 - sg * cs) * (dn - sg * cs / r1 * dr)).max() < 1e-13
    assert np.abs(d_un).min() > 1e-5
    for a in ta:
        assert np.array_equal(u_next[a], take(u[a]))


@pytest.mark.parametrize("loc", FACES3)
@pytest.mark.parametrize("kind", ["nee", "nebb", "equilibrium"])
def test_pressure_wrappers_3d_mass_balance(loc, kind):
    """Pins the pressure wrappers (lbm3d/boundary/_helpers.py:66-77, 105-116)
    by the mass balance they impose, computed here from the lattice moments
    alone.  For a uniform equilibrium feq(rho0, u0) the face sums satisfy
    sum_zero + 2 sum_out = rho0 (1 - u0 . n_in) (n_in the inward normal), so a
    prescribed face density rho1 fixes the inward normal velocity
    1 - (rho0 / rho1)(1 - u0 . n_in).  After each pressure wrapper the face has
    density rho1 and that normal velocity; EQ and NEE also carry the
    neighbour's tangential velocity (checked on a state whose tangential
    velocity changes along the normal); with rho1 = rho0 a uniform
    equilibrium is unchanged."""
    s = LT.WALLS3[loc]
    na, sg, ta = s["axis"], s["sign"], s["tan_axes"]
    shape = (5, 4, 6)
    rho0, u0 = 1.01, np.array([0.03, -0.02, 0.015])
    uf = np.broadcast_to(u0[:, None, None, None], (3,) + shape).copy()
    f = L.get_equilibrium(np.full(shape, rho0), uf)
    # the same with tangential velocities that change along the normal (face != neighbour)
    xn = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")[na]
    ug = uf.copy()
    for k, a in enumerate(ta):
        ug[a] += (0.004 - 0.007 * k) * xn
    g = L.get_equilibrium(np.full(shape, rho0), ug)
    fn = getattr(L, f"boundary_pressure_{kind}")
    un_in = sg * u0[na]
    w, nb = (0, 1) if sg > 0 else (-1, -2)
    for state, uu in ((f, uf), (g, ug)):
        for rho1 in (rho0, 0.997):
            out = fn(state, loc, rho_wall=rho1)
            r, u = L.get_macroscopic(np.take(out, w, axis=1 + na))
            assert np.abs(r - rho1).max() < 1e-13
            assert np.abs(sg * u[na] - (1 - rho0 / rho1 * (1 - un_in))).max() < 1e-13
            if kind != "nebb":
                for a in ta:
                    assert np.abs(u[a] - np.take(uu[a], nb, axis=na)).max() < 1e-13
            if rho1 == rho0 and state is f:
                assert np.abs(out - f).max() < 1e-15
