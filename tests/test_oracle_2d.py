"""Pins the D2Q9 oracle (oracle/lbm2d.py) with closed forms, invariants and
brute force.  The reference has no tests or golden vectors for this path, so
each check is chosen so that a plausible slip (a wrong sign, index, constant or
transposed operand) fails it."""

import numpy as np
import pytest

from oracle import lattice as LT
from oracle import lbm2d as L

rng = np.random.default_rng(1234)


def rand_f(nx=7, ny=5, dtype=np.float64):
    rho = 1 + 0.05 * rng.standard_normal((nx, ny))
    u = 0.05 * rng.standard_normal((2, nx, ny))
    feq = L.get_equilibrium(rho, u)
    return (feq + 1e-3 * rng.standard_normal(feq.shape)).astype(dtype)


def moments(f):
    c = LT.C2.astype(f.dtype)
    m0 = f.sum(0)
    m1 = np.tensordot(c.T, f, axes=([1], [0]))
    m2 = np.einsum("qa,qb,q...->ab...", c, c, f)
    return m0, m1, m2


def test_tables():
    assert LT.check_tables()
    w, c = LT.W2, LT.C2
    assert np.isclose(w.sum(), 1)
    assert np.allclose(w @ c, 0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(2) / 3)
    # reference numbering (lbm/lattice.py:44-57, basic.py:6-10 diagram): 5 = (+1,+1), 6 = (-1,+1)
    assert tuple(c[5]) == (1, 1) and tuple(c[6]) == (-1, 1) and tuple(c[8]) == (1, -1)
    assert list(LT.OPP2) == [0, 3, 4, 1, 2, 7, 8, 5, 6]


def test_streaming_bruteforce():
    f = rand_f(6, 5)
    s = L.streaming(f)
    nx, ny = f.shape[1:]
    for i in range(9):
        for x in range(nx):
            for y in range(ny):
                assert s[i, x, y] == f[i, (x - LT.C2[i, 0]) % nx, (y - LT.C2[i, 1]) % ny]


def test_macroscopic_matches_moments():
    f = rand_f()
    rho, u = L.get_macroscopic(f)
    m0, m1, _ = moments(f)
    assert np.allclose(rho, m0, rtol=1e-14)
    assert np.allclose(u, m1 / m0, rtol=1e-12, atol=1e-15)


def test_equilibrium_moments():
    rho = 1 + 0.1 * rng.standard_normal((4, 3))
    u = 0.1 * rng.standard_normal((2, 4, 3))
    feq = L.get_equilibrium(rho, u)
    m0, m1, m2 = moments(feq)
    assert np.allclose(m0, rho, atol=1e-14)
    assert np.allclose(m1, rho * u, atol=1e-14)
    expect = rho * (np.eye(2)[:, :, None, None] / 3 + u[:, None] * u[None, :])
    assert np.allclose(m2, expect, atol=1e-14)
    # rest state
    assert np.allclose(L.get_equilibrium(np.ones(3), np.zeros((2, 3))), LT.W2[:, None])


def test_bgk_limits():
    f = rand_f()
    rho, u = L.get_macroscopic(f)
    feq = L.get_equilibrium(rho, u)
    assert np.allclose(L.collision_bgk(f, feq, 1.0), feq)
    assert np.allclose(L.collision_bgk(f, feq, 0.0), f)
    post = L.collision_bgk(f, feq, 1.7)
    m0, m1, _ = moments(post)
    assert np.allclose(m0, rho) and np.allclose(m1, rho * u)
    assert np.isclose(L.get_omega(0.01), 1 / 0.53)


def test_mrt_basis_is_physical():
    M = L.M_MRT
    c = LT.C2
    cx, cy = c[:, 0], c[:, 1]
    assert (M[0] == 1).all()
    assert (M[3] == cx).all() and (M[5] == cy).all()
    assert (M[7] == cx ** 2 - cy ** 2).all() and (M[8] == cx * cy).all()
    assert (M[1] == 3 * (cx ** 2 + cy ** 2) - 4).all()
    G = M @ M.T
    assert np.allclose(G, np.diag(np.diag(G)))           # orthogonal rows
    assert np.allclose(L.M_MRT_INV @ M, np.eye(9))


def test_mrt_operator():
    om = 1.6
    op = L.get_mrt_collision_operator(om)
    M = L.M_MRT
    # conserved moments untouched
    assert np.allclose(M[[0, 3, 5]] @ op, 0, atol=1e-13)
    # each moment is an eigenvector with its rate; shear moments relax with omega
    rates = L.mrt_relaxation_rates(om)
    for k in range(9):
        v = L.M_MRT_INV[:, k]
        assert np.allclose(op @ v, rates[k] * v, atol=1e-13)
    assert rates[7] == om and rates[8] == om
    # collision_mrt equals BGK on a field whose non-equilibrium part is pure shear stress
    f = rand_f()
    rho, u = L.get_macroscopic(f)
    feq = L.get_equilibrium(rho, u)
    fneq = np.tensordot(L.M_MRT_INV[:, 7:9], rng.standard_normal((2,) + f.shape[1:]) * 1e-3, axes=([1], [0]))
    out = L.collision_mrt(feq + fneq, feq, op)
    assert np.allclose(out, L.collision_bgk(feq + fneq, feq, om), atol=1e-15)
    # float32 mirror is close to the float64 operator
    assert np.allclose(L.get_mrt_collision_operator(om, np.float32), op, atol=1e-6)


def test_kbc():
    om = 1.8
    f = rand_f()
    rho, u = L.get_macroscopic(f)
    feq = L.get_equilibrium(rho, u)
    post = L.collision_kbc(f, feq, om)
    m0, m1, _ = moments(post)
    assert np.allclose(m0, rho, atol=1e-14) and np.allclose(m1, rho * u, atol=1e-14)
    assert np.allclose(L.collision_kbc(feq, feq, om), feq)
    # pure-shear fneq (h = 0) -> BGK
    M = L.M_MRT
    sh = np.tensordot(L.M_MRT_INV[:, 7:9], 1e-3 * rng.standard_normal((2,) + f.shape[1:]), axes=([1], [0]))
    assert np.allclose(L.collision_kbc(feq + sh, feq, om), feq + (1 - om) * sh, atol=1e-15)
    # fneq with zero shear content (s = 0) -> full relaxation of h: result = feq
    gh = np.tensordot(L.M_MRT_INV[:, [1, 2, 4, 6]], 1e-3 * rng.standard_normal((4,) + f.shape[1:]), axes=([1], [0]))
    assert np.allclose(L.collision_kbc(feq + gh, feq, om), feq, atol=1e-15)
    # defining property of the stabiliser: post-collision fneq is 1/feq-orthogonal to h
    fneq = f - feq
    nsd = fneq[1] - fneq[2] + fneq[3] - fneq[4]
    ss = fneq[5] - fneq[6] + fneq[7] - fneq[8]
    s = np.zeros_like(fneq)
    s[1:5] = np.array([1, -1, 1, -1])[:, None, None] * nsd / 4
    s[5:9] = np.array([1, -1, 1, -1])[:, None, None] * ss / 4
    h = fneq - s
    assert np.allclose(M[[7, 8]] @ s.reshape(9, -1), M[[7, 8]] @ fneq.reshape(9, -1))  # s carries the stress
    assert np.allclose((h * (post - feq) / feq).sum(0), 0, atol=1e-15)


def test_reg():
    om = 1.7
    f = rand_f()
    rho, u = L.get_macroscopic(f)
    feq = L.get_equilibrium(rho, u)
    post = L.collision_reg(f, feq, om)
    reg = (post - feq) / (1 - om)
    m0, m1, m2 = moments(reg)
    _, _, p2 = moments(f - feq)
    assert np.allclose(m0, 0, atol=1e-15) and np.allclose(m1, 0, atol=1e-15)
    assert np.allclose(m2, p2, atol=1e-15)
    # idempotent projection
    again = (L.collision_reg(feq + reg, feq, om) - feq) / (1 - om)
    assert np.allclose(again, reg, atol=1e-15)


def test_guo_moments_and_edm_identity():
    rho = 1 + 0.1 * rng.standard_normal((5, 4))
    u = 0.05 * rng.standard_normal((2, 5, 4))
    g = 1e-3 * rng.standard_normal((2, 5, 4))
    F = L.get_guo_forcing_term(g, u)
    m0, m1, m2 = moments(F)
    assert np.allclose(m0, 0, atol=1e-16)
    assert np.allclose(m1, g, atol=1e-16)
    assert np.allclose(m2, u[:, None] * g[None, :] + g[:, None] * u[None, :], atol=1e-16)
    edm = L.get_equilibrium(rho, u + g / (2 * rho)) - L.get_equilibrium(rho, u - g / (2 * rho))
    assert np.allclose(F, edm, atol=1e-16)
    f = rand_f(5, 4)
    assert np.allclose(L.forcing_edm(f, g, u), f + F)
    assert np.allclose(L.forcing_guo_bgk(f, g, u, 1.5), f + 0.25 * F)


def test_guo_mrt_operator():
    om = 1.6
    fop = L.get_mrt_forcing_operator(om)
    rates = L.mrt_relaxation_rates(om)
    for k in range(9):
        v = L.M_MRT_INV[:, k]
        assert np.allclose(fop @ v, (1 - rates[k] / 2) * v, atol=1e-13)
    # with S = 0 it is the identity: equals EDM
    u = 0.05 * rng.standard_normal((2, 5, 4))
    g = 1e-3 * rng.standard_normal((2, 5, 4))
    f = rand_f(5, 4)
    out = L.forcing_guo_mrt(f, g, u, fop)
    m0, m1, _ = moments(out - f)
    assert np.allclose(m0, 0, atol=1e-16) and np.allclose(m1, g, atol=1e-16)


# ------------------------------------------------------------------ boundaries

LOCS = ["left", "right", "top", "bottom"]


def wall_line(f, loc):
    s = LT.WALLS2[loc]
    return f[:, s["idx"], :] if s["axis"] == 0 else f[:, :, s["idx"]]


def nb_line(f, loc):
    s = LT.WALLS2[loc]
    return f[:, s["nb"], :] if s["axis"] == 0 else f[:, :, s["nb"]]


def others_unchanged(a, b, loc):
    s = LT.WALLS2[loc]
    mask = np.ones(a.shape[1:], dtype=bool)
    if s["axis"] == 0:
        mask[s["idx"], :] = False
    else:
        mask[:, s["idx"]] = False
    return np.array_equal(a[:, mask], b[:, mask])


@pytest.mark.parametrize("loc", LOCS)
def test_equilibrium_and_nee(loc):
    f = rand_f(6, 5)
    out = L.boundary_equilibrium(f, loc, rho_wall=1.02, ux_wall=0.03, uy_wall=-0.01)
    n = wall_line(f, loc).shape[1]
    assert np.allclose(wall_line(out, loc), L.get_equilibrium(np.full(n, 1.02), np.array([[0.03] * n, [-0.01] * n])))
    assert others_unchanged(f, out, loc)
    # NEE: wall = feq(wall values) + fneq(neighbour)
    out = L.boundary_nee(f, loc, rho_wall=1.02, ux_wall=0.03, uy_wall=-0.01)
    nb = nb_line(f, loc)
    r, uu = L.get_macroscopic(nb)
    expect = L.get_equilibrium(np.full(n, 1.02), np.array([[0.03] * n, [-0.01] * n])) + nb - L.get_equilibrium(r, uu)
    assert np.allclose(wall_line(out, loc), expect)
    # neighbour at equilibrium -> wall exactly at the prescribed equilibrium
    g = f.copy()
    s = LT.WALLS2[loc]
    sl = (slice(None), s["nb"], slice(None)) if s["axis"] == 0 else (slice(None), slice(None), s["nb"])
    r, uu = L.get_macroscopic(g[sl])
    g[sl] = L.get_equilibrium(r, uu)
    out = L.boundary_nee(g, loc, rho_wall=0.99, ux_wall=0.02, uy_wall=0.01)
    rw, uw = L.get_macroscopic(wall_line(out, loc))
    assert np.allclose(rw, 0.99) and np.allclose(uw[0], 0.02) and np.allclose(uw[1], 0.01)


@pytest.mark.parametrize("loc", LOCS)
def test_nebb_velocity_imposes_wall_state(loc):
    f = rand_f(6, 5)
    fs = L.streaming(f)
    out = L.boundary_velocity_nebb(fs, loc, ux_wall=0.04, uy_wall=-0.02)
    rho_expect = L.get_rho_wall_from_velocity(fs, loc, 0.04, -0.02)
    r, u = L.get_macroscopic(wall_line(out, loc))
    assert np.allclose(r, rho_expect, rtol=1e-13)
    assert np.allclose(u[0], 0.04, atol=1e-13) and np.allclose(u[1], -0.02, atol=1e-13)
    assert others_unchanged(fs, out, loc)
    # known dirs untouched
    s = LT.WALLS2[loc]
    keep = [i for i in range(9) if i not in s["in_dirs"]]
    assert np.array_equal(wall_line(out, loc)[keep], wall_line(fs, loc)[keep])


@pytest.mark.parametrize("loc", LOCS)
def test_nebb_pressure(loc):
    fs = L.streaming(rand_f(6, 5))
    out = L.boundary_pressure_nebb(fs, loc, rho_wall=1.01)
    r, u = L.get_macroscopic(wall_line(out, loc))
    assert np.allclose(r, 1.01, rtol=1e-13)
    s = LT.WALLS2[loc]
    _, unb = L.get_macroscopic(nb_line(fs, loc))
    ta = 1 - s["axis"]
    assert np.allclose(u[ta], unb[ta], atol=1e-13)   # tangential copied from the neighbour


@pytest.mark.parametrize("loc", LOCS)
def test_force_corrected(loc):
    fs = L.streaming(rand_f(6, 5))
    a = L.boundary_force_corrected_nee(fs, loc, rho_wall=1.0, ux_wall=0.03, uy_wall=0.0, gx_wall=2e-3, gy_wall=-4e-3)
    b = L.boundary_nee(fs, loc, rho_wall=1.0, ux_wall=0.03 - 1e-3, uy_wall=2e-3)
    assert np.allclose(a, b)


@pytest.mark.parametrize("loc", LOCS)
def test_bounce_back_and_specular(loc):
    pre = rand_f(6, 5)
    fs = L.streaming(pre)
    s = LT.WALLS2[loc]
    uw = (0.03, -0.02)
    out = L.boundary_bounce_back(pre, fs, loc, ux_wall=uw[0], uy_wall=uw[1])
    wp = wall_line(pre, loc)
    wo = wall_line(out, loc)
    for i in s["in_dirs"]:
        corr = 6 * LT.W2[i] * (LT.C2[i] @ np.array(uw))        # 2 w rho (c.u)/cs2, rho = 1
        assert np.allclose(wo[i], wp[LT.OPP2[i]] + corr)
    keep = [i for i in range(9) if i not in s["in_dirs"]]
    assert np.array_equal(wo[keep], wall_line(fs, loc)[keep])
    # specular: incoming direction = outgoing with its normal component flipped
    out = L.boundary_specular_reflection(pre, fs, loc, ux_wall=0.0, uy_wall=0.0)
    wo = wall_line(out, loc)
    for i in s["in_dirs"]:
        cm = LT.C2[i].copy()
        cm[s["axis"]] *= -1
        j = int(np.flatnonzero((LT.C2 == cm).all(1))[0])
        assert np.allclose(wo[i], wp[j])


def test_obstacle_bounce_back():
    f = rand_f(6, 5)
    mask = rng.random((6, 5)) < 0.3
    out = L.obstacle_bounce_back(f, mask)
    assert np.array_equal(out[:, ~mask], f[:, ~mask])
    for i in range(9):
        assert np.array_equal(out[i][mask], f[LT.OPP2[i]][mask])


def test_bad_loc_raises_keyerror():
    with pytest.raises(KeyError):
        L.boundary_nee(rand_f(), "front")


def test_characteristic():
    rho = np.ones((6, 5)) * 1.01
    u = np.zeros((2, 6, 5))
    u[0] = 0.04
    u[1] = 0.01
    for loc in LOCS:
        r, uu = L.boundary_characteristic(rho, u, loc)
        assert np.allclose(r, 1.01) and np.allclose(uu[0], 0.04) and np.allclose(uu[1], 0.01)
    # mirror test: "right" on the x-mirrored field (u_x negated) mirrors the velocity
    # update of "left"; the density correction flips sign because the reference
    # writes rho_next = rho - 0.5 rho / cs * L_out on both sides (cbc.py:46).
    rho = 1 + 0.01 * rng.standard_normal((6, 5))
    u = 0.02 * rng.standard_normal((2, 6, 5))
    rl, ul = L.boundary_characteristic(rho, u, "left")
    um = u[:, ::-1].copy()
    um[0] *= -1
    rr, ur = L.boundary_characteristic(rho[::-1], um, "right")
    assert np.allclose(ul[0], -ur[0]) and np.allclose(ul[1], ur[1])
    assert np.allclose(rl - rho[0], -(rr - rho[0]))
    # 2nd-order one-sided derivative is exact on a quadratic profile: with u = 0 and
    # rho(x) = 1 + a x + b x^2 the left update is rho0 - 0.5 rho0 / cs * (cs^2 a / rho0)
    a, b = 0.01, 0.003
    x = np.arange(6)[:, None] * np.ones((1, 5))
    rho = 1 + a * x + b * x ** 2
    r, _ = L.boundary_characteristic(rho, np.zeros((2, 6, 5)), "left")
    cs = 1 / np.sqrt(3)
    assert np.allclose(r, 1 - 0.5 / cs * cs ** 2 * a)
    with pytest.raises(ValueError):
        L.boundary_characteristic(rho, u, "front")


def test_mrt_moment_space_equilibrium():
    """The unforced moment-space MRT of the fused kernels (cell.hpp
    mrt_d2q9_unforced) uses the equilibrium moments of the Lallemand-Luo
    basis in closed form; M feq(rho, u) must equal them, and f + M^-1 S (meq - M f)
    must equal the population-space collision_mrt."""
    rng = np.random.default_rng(11)
    rho = 1 + 0.05 * rng.standard_normal((6, 5))
    u = 0.08 * rng.standard_normal((2, 6, 5))
    feq = L.get_equilibrium(rho, u)
    ux, uy = u
    u2 = ux * ux + uy * uy
    meq = np.stack([rho, rho * (3 * u2 - 2), rho * (1 - 3 * u2), rho * ux, -rho * ux, rho * uy, -rho * uy,
                    rho * (ux * ux - uy * uy), rho * ux * uy])
    M = L.M_MRT.astype(np.float64)
    assert np.abs(np.tensordot(M, feq, axes=([1], [0])) - meq).max() < 1e-15
    f = feq + 1e-3 * rng.standard_normal(feq.shape)
    omega = 1.7
    S = L.mrt_relaxation_rates(omega)
    m = np.tensordot(M, f, axes=([1], [0]))
    moment_form = f + np.tensordot(L.M_MRT_INV, S[:, None, None] * (meq - m), axes=([1], [0]))
    ref = L.collision_mrt(f, feq, L.get_mrt_collision_operator(omega))
    assert np.abs(moment_form - ref).max() < 1e-15


def _cbc_profiles(shape, na, dt=np.float64):
    """Quadratic profiles along the normal axis na (coefficients varying across
    the wall) and their exact derivatives d/dx_na, as functions of the index grid."""
    idx = np.meshgrid(*[np.arange(n, dtype=dt) for n in shape], indexing="ij")
    x = idx[na]
    t = sum(idx[a] for a in range(len(shape)) if a != na)
    a_r, b_r = 0.01 + 0.001 * t, 0.002
    a_u, b_u = 0.004 + 0.0005 * t, -0.001
    rho = 1 + a_r * x + b_r * x * x
    un = 0.03 + 0.002 * t + a_u * x + b_u * x * x
    return x, t, rho, un, a_r + 2 * b_r * x, a_u + 2 * b_u * x


@pytest.mark.parametrize("loc", LOCS)
def test_characteristic_invariants(loc):
    """Pins boundary/cbc.py:14-53 by the characteristic form of its update with
    exact derivatives.  With J+- = u_n +- cs ln(rho) linearised as
    dJ+- = du_n +- (cs / rho) drho and lambda = u_n - s cs (s = inward sign):
      du_n - (cs / rho) drho = 0 and du_n + (cs / rho) drho = -lambda (u_n' - s cs / rho rho'),
    where ' is the exact derivative along the axis of the quadratic profiles
    (the 3-point one-sided difference of the reference is exact on them).  At
    s = -1 this is the outgoing right-running characteristic advected with
    u + cs while the incoming one is held; at s = +1 the reference applies the
    left-running advection term to the same combination.  The tangential
    velocity is copied from the wall line."""
    s = LT.WALLS2[loc]
    na, sg = s["axis"], s["sign"]
    shape = (7, 6)
    x, t, rho, un, drho, dun = _cbc_profiles(shape, na)
    ut = 0.01 + 0.003 * t + 0.005 * x
    u = np.zeros((2,) + shape)
    u[na], u[1 - na] = un, ut
    r_next, u_next = L.boundary_characteristic(rho, u, loc)
    w = 0 if sg > 0 else -1
    take = lambda a: np.take(a, w, axis=na)
    r1, n1, dr, dn = take(rho), take(un), take(drho), take(dun)
    cs = 1 / np.sqrt(3)
    lam = n1 - sg * cs
    d_rho, d_un = r_next - r1, u_next[na] - n1
    assert np.abs(d_un - cs / r1 * d_rho).max() < 1e-13
    assert np.abs((d_un + cs / r1 * d_rho) + lam * (dn - sg * cs / r1 * dr)).max() < 1e-13
    assert np.abs(d_un).min() > 1e-5                       # the case moves the state
    assert np.array_equal(u_next[1 - na], take(ut))
