"""Oracle restatement of ``vivsim.lbm`` (D2Q9): core ops, collisions, forcing, boundaries.

Arrays follow the reference layout: ``f[9, NX, NY]``, ``rho[NX, NY]``,
``u[2, NX, NY]``; every function also accepts wall lines ``(9, N)``. Each
function computes in the floating dtype of its inputs (float32 mirrors the
reference's default JAX arithmetic, float64 is the truth).
"""

import numpy as np

from .lattice import C2, W2, OPP2, WALLS2

Q = 9


def _ft(*xs):
    """Working float dtype of the inputs (at least float32)."""
    dt = np.result_type(*[np.asarray(x).dtype for x in xs])
    return dt if np.issubdtype(dt, np.floating) else np.dtype(np.float32)


def _col(v, ndim):
    return np.asarray(v).reshape((-1,) + (1,) * ndim)


# ------------------------------------------------------------------ basic.py

def streaming(f):
    """Periodic pull streaming ``f'_i(x) = f_i(x - c_i)`` — lbm/basic.py:60-85 (shifts 44-57)."""
    out = np.empty_like(f)
    for i in range(Q):
        out[i] = np.roll(f[i], shift=(int(C2[i, 0]), int(C2[i, 1])), axis=(0, 1))
    return out


def get_macroscopic(f):
    """rho = sum f; u = (sum_right - sum_left, sum_up - sum_down) / rho — lbm/basic.py:88-111."""
    rho = f.sum(axis=0)
    ux = (f[1] + f[5] + f[8]) - (f[3] + f[7] + f[6])
    uy = (f[2] + f[5] + f[6]) - (f[4] + f[7] + f[8])
    u = np.stack([ux / rho, uy / rho])
    return rho, u


def get_equilibrium(rho, u):
    """feq_i = rho w_i (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 |u|^2) — lbm/basic.py:114-135."""
    rho = np.asarray(rho)
    u = np.asarray(u)
    dt = _ft(rho, u)
    nd = rho.ndim
    cx = _col(C2[:, 0].astype(dt), nd)
    cy = _col(C2[:, 1].astype(dt), nd)
    uc = u[0][None] * cx + u[1][None] * cy
    usq = (u * u).sum(axis=0)
    w = _col(W2.astype(dt), nd)
    return (rho[None] * w * (1 + 3 * uc + 4.5 * uc ** 2 - 1.5 * usq[None])).astype(dt)


def collision_bgk(f, feq, omega):
    """(1 - omega) f + omega feq — lbm/basic.py:138-156."""
    return (1 - omega) * f + omega * feq


def get_omega(nu):
    """omega = 1 / (3 nu + 0.5) — lbm/basic.py:159-172."""
    return 1 / (3 * nu + 0.5)


def get_velocity_correction(g, rho=1):
    """0.5 g / rho — lbm/basic.py:175-194."""
    return g * 0.5 / rho


# ------------------------------------------------------------------ collision/mrt.py

# Lallemand-Luo moment basis, rows: rho, e, eps, jx, qx, jy, qy, pxx, pxy (collision/mrt.py:10-22).
M_MRT = np.array([
    [1, 1, 1, 1, 1, 1, 1, 1, 1],
    [-4, -1, -1, -1, -1, 2, 2, 2, 2],
    [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1],
    [0, -2, 0, 2, 0, 1, -1, -1, 1],
    [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1],
    [0, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 1, -1, 1, -1],
], dtype=np.int64)
M_MRT_INV = np.linalg.inv(M_MRT.astype(np.float64))   # collision/mrt.py:25


def mrt_relaxation_rates(omega):
    """diag(S) = (0, 1.4, 1.4, 0, 1.2, 0, 1.2, omega, omega) — collision/mrt.py:28-44."""
    return np.array([0, 1.4, 1.4, 0, 1.2, 0, 1.2, omega, omega], dtype=np.float64)


def get_mrt_collision_operator(omega, dtype=np.float64):
    """M^-1 S M — collision/mrt.py:47-63.

    ``dtype=np.float32`` mirrors the reference under default JAX (x32): M^-1 and
    S are rounded to float32 and the two products are float32 matmuls.
    """
    S = np.diag(mrt_relaxation_rates(omega)).astype(dtype)
    return (M_MRT_INV.astype(dtype) @ S) @ M_MRT.astype(dtype)


def collision_mrt(f, feq, mrt_collision_matrix):
    """f + op (feq - f) — collision/mrt.py:66-88."""
    op = np.asarray(mrt_collision_matrix).astype(_ft(f))
    return f + np.tensordot(op, feq - f, axes=([1], [0]))


# ------------------------------------------------------------------ collision/kbc.py

def collision_kbc(f, feq, omega):
    """Entropic KBC collision — collision/kbc.py:10-59.

    shear part from N = Pxx - Pyy and Pxy, h = fneq - s,
    gamma/2 = 1/omega - (1 - 1/omega) <h,s>/<h,h> (inner products weighted by 1/(feq+1e-20)).
    """
    fneq = f - feq
    nsd = fneq[1] - fneq[2] + fneq[3] - fneq[4]
    ss = fneq[5] - fneq[6] + fneq[7] - fneq[8]
    s = np.zeros_like(fneq)
    sg = np.array([1, -1, 1, -1], dtype=f.dtype).reshape((4,) + (1,) * (f.ndim - 1))
    s[1:5] = sg * (nsd / 4.0)[None]
    s[5:9] = sg * (ss / 4.0)[None]
    h = fneq - s
    denom = feq + 1e-20
    ish = (h * s / denom).sum(axis=0)
    ihh = (h * h / denom).sum(axis=0)
    half_gamma = 1.0 / omega - (1.0 - 1.0 / omega) * ish / (ihh + 1e-20)
    return f - omega * (s + half_gamma[None] * h)


# ------------------------------------------------------------------ collision/reg.py

def collision_reg(f, feq, omega):
    """Regularized BGK: feq + (1 - omega) * Hermite-2 projection of fneq — collision/reg.py:4-49."""
    fneq = f - feq
    dsum = fneq[5] + fneq[6] + fneq[7] + fneq[8]
    pxx = fneq[1] + fneq[3] + dsum
    pyy = fneq[2] + fneq[4] + dsum
    pxy = fneq[5] - fneq[6] + fneq[7] - fneq[8]
    tr = (pxx + pyy) / 12.0
    ax = (2.0 * pxx - pyy) / 6.0
    ay = (2.0 * pyy - pxx) / 6.0
    dp = tr + pxy / 4.0
    dn = tr - pxy / 4.0
    reg = np.stack([-8.0 * tr, ax, ay, ax, ay, dp, dn, dp, dn])
    return feq + (1 - omega) * reg


# ------------------------------------------------------------------ forcing/guo.py, edm.py

def get_guo_forcing_term(g, u):
    """F_i = w_i [gx (3 (cx - ux) + 9 (c.u) cx) + gy (3 (cy - uy) + 9 (c.u) cy)] — forcing/guo.py:6-33."""
    g = np.asarray(g)
    u = np.asarray(u)
    dt = _ft(g, u)
    nd = u.ndim - 1
    cx = _col(C2[:, 0].astype(dt), nd)
    cy = _col(C2[:, 1].astype(dt), nd)
    uc = u[0][None] * cx + u[1][None] * cy
    w = _col(W2.astype(dt), nd)
    return w * (g[0][None] * (3 * (cx - u[0][None]) + 9 * (uc * cx))
                + g[1][None] * (3 * (cy - u[1][None]) + 9 * (uc * cy)))


def forcing_guo_bgk(f, g, u, omega):
    """f + F (1 - omega/2) — forcing/guo.py:38-57."""
    return f + get_guo_forcing_term(g, u) * (1 - 0.5 * omega)


def get_mrt_forcing_operator(omega, dtype=np.float64):
    """M^-1 (I - S/2) M — forcing/guo.py:62-79 (dtype as in get_mrt_collision_operator)."""
    S = np.diag(mrt_relaxation_rates(omega)).astype(dtype)
    return (M_MRT_INV.astype(dtype) @ (np.eye(9, dtype=dtype) - dtype(0.5) * S)) @ M_MRT.astype(dtype)


def forcing_guo_mrt(f, g, u, mrt_forcing_operator):
    """f + opF F — forcing/guo.py:82-107."""
    op = np.asarray(mrt_forcing_operator).astype(_ft(f))
    return f + np.tensordot(op, get_guo_forcing_term(g, u), axes=([1], [0]))


def forcing_edm(f, g, u):
    """f + F (EDM written in its Guo-equivalent form) — forcing/edm.py:4-31."""
    return f + get_guo_forcing_term(g, u)


# ------------------------------------------------------------------ boundary/_helpers.py

def _wall(f, loc, which="idx"):
    s = WALLS2[loc]
    i = s[which]
    return f[:, i, :] if s["axis"] == 0 else f[:, :, i]


def _set_wall(f, loc, vals):
    s = WALLS2[loc]
    out = f.copy()
    if s["axis"] == 0:
        out[:, s["idx"], :] = vals
    else:
        out[:, :, s["idx"]] = vals
    return out


def get_boundary_shape(f, loc):
    """Nodes on the wall — boundary/_helpers.py:47-50 (KeyError on a bad loc)."""
    s = WALLS2[loc]
    return f.shape[2] if s["axis"] == 0 else f.shape[1]


def broadcast_wall_values(f, loc, rho_wall=1, ux_wall=0, uy_wall=0):
    """Scalar -> (N,) broadcast of wall data — boundary/_helpers.py:53-77."""
    n = get_boundary_shape(f, loc)
    dt = f.dtype

    def b(v):
        return np.full(n, v, dtype=dt) if np.isscalar(v) else np.asarray(v)

    return b(rho_wall), np.stack([b(ux_wall), b(uy_wall)])


def get_signed_wall_velocity(loc, ux_wall=0, uy_wall=0):
    """(inward normal, tangential) components times normal_sign — boundary/_helpers.py:148-153."""
    s = WALLS2[loc]
    un = ux_wall if s["axis"] == 0 else uy_wall
    ut = uy_wall if s["axis"] == 0 else ux_wall
    return s["sign"] * un, s["sign"] * ut


def _rho_wall_numerator(f, loc):
    """f0 + f_t0 + f_t1 + 2 (sum of out dirs) at the wall — boundary/_helpers.py:135-145."""
    s = WALLS2[loc]
    w = _wall(f, loc)
    t0, t1 = s["tan"]
    o0, o1, o2 = s["out_dirs"]
    return w[0] + w[t0] + w[t1] + 2 * (w[o0] + w[o1] + w[o2])


def get_rho_wall_from_velocity(f, loc, ux_wall=0, uy_wall=0):
    """numerator / (1 - u_n) — boundary/_helpers.py:80-95."""
    un, _ = get_signed_wall_velocity(loc, ux_wall, uy_wall)
    return _rho_wall_numerator(f, loc) / (1 - un)


def get_wall_velocity_from_pressure(f, loc, rho_wall=1):
    """Normal velocity from the prescribed density; tangential one copied from the
    neighbour line's macroscopic velocity — boundary/_helpers.py:98-132."""
    s = WALLS2[loc]
    nb = _wall(f, loc, "nb")
    rho_nb = nb.sum(axis=0)
    un = s["sign"] * (1 - _rho_wall_numerator(f, loc) / rho_wall)
    p, n = s["pos"], s["neg"]
    ut = (nb[p[0]] - nb[n[0]] + nb[p[1]] - nb[n[1]] + nb[p[2]] - nb[n[2]]) / rho_nb
    vel = [None, None]
    vel[s["axis"]] = un
    vel[1 - s["axis"]] = ut
    return tuple(vel)


def get_corrected_wall_velocity(ux_wall, uy_wall, rho_wall=1, gx_wall=0, gy_wall=0):
    """u - 0.5 g / rho — boundary/_helpers.py:156-177."""
    return (ux_wall - get_velocity_correction(gx_wall, rho_wall),
            uy_wall - get_velocity_correction(gy_wall, rho_wall))


def _wrap_velocity(fn):
    # boundary/_helpers.py:180-197
    def wrapped(f, loc, ux_wall=0, uy_wall=0):
        rho_wall = get_rho_wall_from_velocity(f, loc, ux_wall=ux_wall, uy_wall=uy_wall)
        return fn(f, loc, rho_wall=rho_wall, ux_wall=ux_wall, uy_wall=uy_wall)
    return wrapped


def _wrap_pressure(fn):
    # boundary/_helpers.py:200-217
    def wrapped(f, loc, rho_wall=1):
        ux_wall, uy_wall = get_wall_velocity_from_pressure(f, loc, rho_wall=rho_wall)
        return fn(f, loc, rho_wall=rho_wall, ux_wall=ux_wall, uy_wall=uy_wall)
    return wrapped


def _wrap_force_corrected(fn):
    # boundary/_helpers.py:220-239
    def wrapped(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, gx_wall=0, gy_wall=0):
        ux_wall, uy_wall = get_corrected_wall_velocity(ux_wall, uy_wall, rho_wall, gx_wall, gy_wall)
        return fn(f, loc, rho_wall=rho_wall, ux_wall=ux_wall, uy_wall=uy_wall)
    return wrapped


# ------------------------------------------------------------------ boundary/nee.py, eq.py, nebb.py

def boundary_nee(f, loc, rho_wall=1, ux_wall=0, uy_wall=0):
    """wall := feq(rho_w, u_w) + (f_nb - feq(macro(f_nb))) — boundary/nee.py:19-60."""
    rw, uw = broadcast_wall_values(f, loc, rho_wall, ux_wall, uy_wall)
    feq_w = get_equilibrium(rw, uw)
    nb = _wall(f, loc, "nb")
    r_nb, u_nb = get_macroscopic(nb)
    return _set_wall(f, loc, feq_w + (nb - get_equilibrium(r_nb, u_nb)))


def boundary_equilibrium(f, loc, rho_wall=1, ux_wall=0, uy_wall=0):
    """wall := feq(rho_w, u_w) — boundary/eq.py:25-56."""
    rw, uw = broadcast_wall_values(f, loc, rho_wall, ux_wall, uy_wall)
    return _set_wall(f, loc, get_equilibrium(rw, uw))


def boundary_nebb(f, loc, rho_wall=1, ux_wall=0, uy_wall=0):
    """Zou/He: the 3 unknown in-dirs from the out-dirs — boundary/nebb.py:20-58."""
    s = WALLS2[loc]
    un, ut = get_signed_wall_velocity(loc, ux_wall, uy_wall)
    w = _wall(f, loc).copy()
    shear = 0.5 * (w[s["tan"][0]] - w[s["tan"][1]]) * s["sign"]
    nterm = 1 / 6 * un * rho_wall
    tterm = 0.5 * ut * rho_wall
    o = s["out_dirs"]
    new = [w[o[0]] + 2 / 3 * un * rho_wall,
           w[o[1]] - shear + nterm + tterm,
           w[o[2]] + shear + nterm - tterm]
    for k, i in enumerate(s["in_dirs"]):
        w[i] = new[k]
    return _set_wall(f, loc, w)


boundary_velocity_nee = _wrap_velocity(boundary_nee)
boundary_pressure_nee = _wrap_pressure(boundary_nee)
boundary_force_corrected_nee = _wrap_force_corrected(boundary_nee)
boundary_velocity_equilibrium = _wrap_velocity(boundary_equilibrium)
boundary_pressure_equilibrium = _wrap_pressure(boundary_equilibrium)
boundary_force_corrected_equilibrium = _wrap_force_corrected(boundary_equilibrium)
boundary_velocity_nebb = _wrap_velocity(boundary_nebb)
boundary_pressure_nebb = _wrap_pressure(boundary_nebb)
boundary_force_corrected_nebb = _wrap_force_corrected(boundary_nebb)


# ------------------------------------------------------------------ boundary/bb.py

def _bb_values(f_before_stream, loc, ux_wall, uy_wall):
    s = WALLS2[loc]
    un, ut = get_signed_wall_velocity(loc, ux_wall, uy_wall)
    pre = _wall(f_before_stream, loc)
    o = s["out_dirs"]
    return [pre[o[0]] + 2 / 3 * un, pre[o[1]] + 1 / 6 * (un + ut), pre[o[2]] + 1 / 6 * (un - ut)]


def boundary_bounce_back(f_before_stream, f, loc, ux_wall=0, uy_wall=0):
    """in_k := f_pre[out_k] + {2/3 u_n, (u_n+u_t)/6, (u_n-u_t)/6} — boundary/bb.py:15-59."""
    s = WALLS2[loc]
    new = _bb_values(f_before_stream, loc, ux_wall, uy_wall)
    w = _wall(f, loc).copy()
    for k, i in enumerate(s["in_dirs"]):
        w[i] = new[k]
    return _set_wall(f, loc, w)


def boundary_specular_reflection(f_before_stream, f, loc, ux_wall=0, uy_wall=0):
    """As bounce-back with the two diagonal targets swapped — boundary/bb.py:62-95."""
    s = WALLS2[loc]
    new = _bb_values(f_before_stream, loc, ux_wall, uy_wall)
    w = _wall(f, loc).copy()
    i0, i1, i2 = s["in_dirs"]
    w[i0], w[i2], w[i1] = new[0], new[1], new[2]
    return _set_wall(f, loc, w)


def obstacle_bounce_back(f, mask):
    """f_i := f_opp(i) on masked cells — boundary/bb.py:98-110."""
    out = f.copy()
    m = np.asarray(mask, dtype=bool)
    out[:, m] = f[:, m][OPP2]
    return out


# ------------------------------------------------------------------ boundary/cbc.py

def _take3(field, axis, inward_sign):
    idx = [0, 1, 2] if inward_sign > 0 else [-1, -2, -3]
    return np.moveaxis(np.take(field, idx, axis=axis), axis, 0)


def boundary_characteristic(rho, u, loc="right"):
    """Non-reflective 1-D characteristic update -> (rho_next (N,), u_next (2, N)) — boundary/cbc.py:14-53."""
    if loc not in WALLS2:
        raise ValueError("loc must be one of: 'left', 'right', 'top', 'bottom'")
    s = WALLS2[loc]
    na, ta, sg = s["axis"], 1 - s["axis"], s["sign"]
    dt = _ft(rho, u)
    cs = dt.type(1 / np.sqrt(3.0))
    r1, r2, r3 = _take3(rho, na, sg)
    n1, n2, n3 = _take3(u[na], na, sg)
    t1 = np.take(u[ta], 0 if sg > 0 else -1, axis=na)
    k = -0.5 * sg
    drho = k * (3 * r1 - 4 * r2 + r3)
    dun = k * (3 * n1 - 4 * n2 + n3)
    ch = dun - sg * cs / r1 * drho
    lo = (n1 - sg * cs) * ch
    rho_next = r1 - 0.5 * r1 / cs * lo
    un_next = n1 - 0.5 * lo
    vel = [None, None]
    vel[na] = un_next
    vel[ta] = t1
    return rho_next, np.stack(vel)
