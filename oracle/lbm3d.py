"""Oracle restatement of ``vivsim.lbm3d`` (D3Q19).

Layout ``f[19, NX, NY, NZ]``, ``u[3, NX, NY, NZ]``; faces are ``(19, N1, N2)``.
"""

import numpy as np

from .lattice import C3, W3, OPP3, WALLS3

Q = 19
CS2 = 1.0 / 3.0


def _ft(*xs):
    dt = np.result_type(*[np.asarray(x).dtype for x in xs])
    return dt if np.issubdtype(dt, np.floating) else np.dtype(np.float32)


def _col(v, ndim):
    return np.asarray(v).reshape((-1,) + (1,) * ndim)


def streaming(f):
    """Periodic pull streaming f'_i(x) = f_i(x - c_i) — lbm3d/basic.py:52-87 (shifts 28-49)."""
    out = np.empty_like(f)
    for i in range(Q):
        out[i] = np.roll(f[i], shift=tuple(int(v) for v in C3[i]), axis=(0, 1, 2))
    return out


def get_macroscopic(f):
    """rho = sum f, u = sum c f / rho — lbm3d/basic.py:90-106."""
    rho = f.sum(axis=0)
    c = C3.astype(f.dtype)
    mom = np.tensordot(c.T, f, axes=([1], [0]))
    return rho, mom / rho[None]


def get_equilibrium(rho, u):
    """rho w (1 + cu/cs2 + (cu)^2/(2 cs2^2) - |u|^2/(2 cs2)) (division form) — lbm3d/basic.py:109-130."""
    rho = np.asarray(rho)
    u = np.asarray(u)
    dt = _ft(rho, u)
    nd = rho.ndim
    uc = np.tensordot(C3.astype(dt), u, axes=([1], [0]))
    usq = (u * u).sum(axis=0)
    cs2 = dt.type(CS2)
    cs4 = dt.type(CS2 ** 2)
    w = _col(W3.astype(dt), nd)
    return (rho[None] * w * (1 + uc / cs2 + 0.5 * (uc ** 2) / cs4 - 0.5 * usq[None] / cs2)).astype(dt)


def collision_bgk(f, feq, omega):
    """lbm3d/basic.py:133-146."""
    return (1 - omega) * f + omega * feq


def get_omega(nu):
    """lbm3d/basic.py:149-152."""
    return 1 / (3 * nu + 0.5)


def get_velocity_correction(g, rho=1):
    """lbm3d/basic.py:155-167."""
    return g * 0.5 / rho


# ------------------------------------------------------------------ collision/mrt.py

# Raw-moment basis (lbm3d/collision/mrt.py:7-30), stored float32 in the reference.
M_MRT = np.zeros((19, 19), dtype=np.float64)


def _build_m3():
    c = C3
    cx, cy, cz = c[:, 0], c[:, 1], c[:, 2]
    c2 = cx * cx + cy * cy + cz * cz
    rows = [
        np.ones(19), cx, cy, cz,                    # rho, jx, jy, jz
        c2, 3 * cx * cx - c2, cy * cy - cz * cz,    # energy, two normal-stress differences
        cx * cy, cx * cz, cy * cz,                  # shear stresses
        cx * cx * cy, cx * cx * cz, cx * cy * cy,   # third order
        cy * cy * cz, cx * cz * cz, cy * cz * cz,
        cx * cx * cy * cy, cx * cx * cz * cz, cy * cy * cz * cz,   # fourth order
    ]
    return np.array(rows, dtype=np.float64)


M_MRT[:] = _build_m3()
M_MRT_INV32 = np.linalg.inv(M_MRT.astype(np.float32))   # mrt.py:31 keeps float32


def mrt_relaxation_rates(omega):
    """diag(S) — lbm3d/collision/mrt.py:34-73."""
    return np.array([0, 0, 0, 0, 1.1] + [omega] * 5 + [1.2] * 6 + [1.4] * 3, dtype=np.float64)


def get_mrt_collision_operator(omega, dtype=np.float64):
    """M^-1 S M — lbm3d/collision/mrt.py:76-80."""
    minv = np.linalg.inv(M_MRT) if dtype == np.float64 else M_MRT_INV32.astype(dtype)
    S = np.diag(mrt_relaxation_rates(omega)).astype(dtype)
    return (minv @ S) @ M_MRT.astype(dtype)


def collision_mrt(f, feq, mrt_collision_matrix):
    """f + op (feq - f) — lbm3d/collision/mrt.py:83-98."""
    op = np.asarray(mrt_collision_matrix).astype(_ft(f))
    return f + np.tensordot(op, feq - f, axes=([1], [0]))


# ------------------------------------------------------------------ collision/reg.py, kbc.py

def _projection():
    """Second-order Hermite projection of fneq (lbm3d/collision/reg.py:10-33).

    Built from H_ab(c) = c_a c_b - cs2 delta_ab as
    P[q, r] = w_q / (2 cs2^2) * sum_ab H_ab(c_q) c_ra c_rb,
    i.e. fneq_reg_q = w_q / (2 cs2^2) H(c_q) : Pi_neq with Pi_neq = sum_r c_r c_r fneq_r.
    The tests check it against entries of the reference's literal table and
    against the moment identities it must satisfy.
    """
    c = C3.astype(np.float64)
    P = np.zeros((19, 19))
    for q in range(19):
        Hq = np.outer(c[q], c[q]) - CS2 * np.eye(3)
        for r in range(19):
            P[q, r] = W3[q] / (2 * CS2 ** 2) * np.sum(Hq * (np.outer(c[r], c[r])))
    return P


PROJ3 = _projection()


def get_second_order_projection(fneq):
    """P fneq — lbm3d/collision/reg.py:36-42."""
    P = PROJ3.astype(fneq.dtype)
    return np.tensordot(P, fneq, axes=([1], [0]))


def collision_reg(f, feq, omega):
    """feq + (1 - omega) P fneq — lbm3d/collision/reg.py:45-62."""
    return feq + (1 - omega) * get_second_order_projection(f - feq)


def collision_kbc(f, feq, omega):
    """KBC with s = P fneq, h = fneq - s — lbm3d/collision/kbc.py:8-42."""
    fneq = f - feq
    s = get_second_order_projection(fneq)
    h = fneq - s
    inv = 1.0 / (feq + 1e-20)
    ish = (h * s * inv).sum(axis=0)
    ihh = (h * h * inv).sum(axis=0)
    half_gamma = 1.0 / omega - (1.0 - 1.0 / omega) * ish / (ihh + 1e-20)
    return f - omega * (s + half_gamma[None] * h)


# ------------------------------------------------------------------ forcing

def get_guo_forcing_term(g, u):
    """w (3 (c.g - g.u) + 9 (c.g)(c.u)) — lbm3d/forcing/guo.py:13-38."""
    g = np.asarray(g)
    u = np.asarray(u)
    dt = _ft(g, u)
    c = C3.astype(dt)
    cu = np.tensordot(c, u, axes=([1], [0]))
    cg = np.tensordot(c, g, axes=([1], [0]))
    gu = (g * u).sum(axis=0)
    w = _col(W3.astype(dt), u.ndim - 1)
    return w * (3.0 * (cg - gu[None]) + 9.0 * cg * cu)


def forcing_guo_bgk(f, g, u, omega):
    """lbm3d/forcing/guo.py:41-56."""
    return f + get_guo_forcing_term(g, u) * (1 - 0.5 * omega)


def get_mrt_forcing_operator(omega, dtype=np.float64):
    """M^-1 (I - S/2) M — lbm3d/forcing/guo.py:59-68."""
    minv = np.linalg.inv(M_MRT) if dtype == np.float64 else M_MRT_INV32.astype(dtype)
    S = np.diag(mrt_relaxation_rates(omega)).astype(dtype)
    return (minv @ (np.eye(19, dtype=dtype) - dtype(0.5) * S)) @ M_MRT.astype(dtype)


def forcing_guo_mrt(f, g, u, mrt_forcing_operator):
    """lbm3d/forcing/guo.py:71-87."""
    op = np.asarray(mrt_forcing_operator).astype(_ft(f))
    return f + np.tensordot(op, get_guo_forcing_term(g, u), axes=([1], [0]))


def forcing_edm(f, g, u):
    """lbm3d/forcing/edm.py:6-26."""
    return f + get_guo_forcing_term(g, u)


# ------------------------------------------------------------------ boundaries

def _face(f, loc, which="idx"):
    s = WALLS3[loc]
    sl = [slice(None)] * 4
    sl[1 + s["axis"]] = s[which]
    return f[tuple(sl)]


def _set_face(f, loc, vals):
    s = WALLS3[loc]
    out = f.copy()
    sl = [slice(None)] * 4
    sl[1 + s["axis"]] = s["idx"]
    out[tuple(sl)] = vals
    return out


def get_boundary_shape(f, loc):
    """lbm3d/boundary/_helpers.py:9-15."""
    s = WALLS3[loc]
    return tuple(f.shape[1 + a] for a in range(3) if a != s["axis"])


def broadcast_wall_values(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/_helpers.py:18-32."""
    shp = get_boundary_shape(f, loc)

    def b(v):
        return np.full(shp, v, dtype=f.dtype) if np.isscalar(v) else np.asarray(v)

    return b(rho_wall), np.stack([b(ux_wall), b(uy_wall), b(uz_wall)])


def _rho_wall_numerator(f, loc):
    """sum(zero dirs) + 2 sum(out dirs) — lbm3d/boundary/_helpers.py:35-42."""
    s = WALLS3[loc]
    w = _face(f, loc)
    return w[s["zero"]].sum(axis=0) + 2.0 * w[s["out_dirs"]].sum(axis=0)


def get_signed_wall_velocity(loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/_helpers.py:45-54."""
    s = WALLS3[loc]
    v = [np.asarray(ux_wall), np.asarray(uy_wall), np.asarray(uz_wall)]
    return s["sign"] * v[s["axis"]], np.stack([s["sign"] * v[a] for a in s["tan_axes"]])


def get_rho_wall_from_velocity(f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/_helpers.py:57-63."""
    un, _ = get_signed_wall_velocity(loc, ux_wall, uy_wall, uz_wall)
    return _rho_wall_numerator(f, loc) / (1 - un)


def get_wall_velocity_from_pressure(f, loc, rho_wall=1):
    """Normal from density, tangential = neighbour macroscopic u — lbm3d/boundary/_helpers.py:66-77."""
    s = WALLS3[loc]
    rw = np.full(get_boundary_shape(f, loc), rho_wall, dtype=f.dtype) if np.isscalar(rho_wall) \
        else np.broadcast_to(rho_wall, get_boundary_shape(f, loc))
    uns = 1 - _rho_wall_numerator(f, loc) / rw
    _, unb = get_macroscopic(_face(f, loc, "nb"))
    vel = [unb[0], unb[1], unb[2]]
    vel[s["axis"]] = s["sign"] * uns
    return tuple(vel)


def get_corrected_wall_velocity(ux_wall, uy_wall, uz_wall, rho_wall=1, gx_wall=0, gy_wall=0, gz_wall=0):
    """lbm3d/boundary/_helpers.py:80-88."""
    return (ux_wall - get_velocity_correction(gx_wall, rho_wall),
            uy_wall - get_velocity_correction(gy_wall, rho_wall),
            uz_wall - get_velocity_correction(gz_wall, rho_wall))


def _wrap_velocity(fn):
    # lbm3d/boundary/_helpers.py:91-102
    def wrapped(f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
        rw = get_rho_wall_from_velocity(f, loc, ux_wall, uy_wall, uz_wall)
        return fn(f, loc, rho_wall=rw, ux_wall=ux_wall, uy_wall=uy_wall, uz_wall=uz_wall)
    return wrapped


def _wrap_pressure(fn):
    # lbm3d/boundary/_helpers.py:105-116
    def wrapped(f, loc, rho_wall=1):
        ux, uy, uz = get_wall_velocity_from_pressure(f, loc, rho_wall)
        return fn(f, loc, rho_wall=rho_wall, ux_wall=ux, uy_wall=uy, uz_wall=uz)
    return wrapped


def _wrap_force_corrected(fn):
    # lbm3d/boundary/_helpers.py:119-133
    def wrapped(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0, gx_wall=0, gy_wall=0, gz_wall=0):
        ux, uy, uz = get_corrected_wall_velocity(ux_wall, uy_wall, uz_wall, rho_wall, gx_wall, gy_wall, gz_wall)
        return fn(f, loc, rho_wall=rho_wall, ux_wall=ux, uy_wall=uy, uz_wall=uz)
    return wrapped


def boundary_nee(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/nee.py:13-33."""
    rw, uw = broadcast_wall_values(f, loc, rho_wall, ux_wall, uy_wall, uz_wall)
    nb = _face(f, loc, "nb")
    r_nb, u_nb = get_macroscopic(nb)
    return _set_face(f, loc, get_equilibrium(rw, uw) + (nb - get_equilibrium(r_nb, u_nb)))


def boundary_equilibrium(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/eq.py:13-26."""
    rw, uw = broadcast_wall_values(f, loc, rho_wall, ux_wall, uy_wall, uz_wall)
    return _set_face(f, loc, get_equilibrium(rw, uw))


def boundary_nebb(f, loc, rho_wall=1, ux_wall=0, uy_wall=0, uz_wall=0):
    """f_in := f_opp(in) + feq_in - feq_opp(in) — lbm3d/boundary/nebb.py:13-30."""
    s = WALLS3[loc]
    ind = s["in_dirs"]
    oud = list(OPP3[ind])
    rw, uw = broadcast_wall_values(f, loc, rho_wall, ux_wall, uy_wall, uz_wall)
    feq = get_equilibrium(rw, uw)
    w = _face(f, loc).copy()
    w[ind] = w[oud] + feq[ind] - feq[oud]
    return _set_face(f, loc, w)


boundary_velocity_nee = _wrap_velocity(boundary_nee)
boundary_pressure_nee = _wrap_pressure(boundary_nee)
boundary_force_corrected_nee = _wrap_force_corrected(boundary_nee)
boundary_velocity_equilibrium = _wrap_velocity(boundary_equilibrium)
boundary_pressure_equilibrium = _wrap_pressure(boundary_equilibrium)
boundary_force_corrected_equilibrium = _wrap_force_corrected(boundary_equilibrium)
boundary_velocity_nebb = _wrap_velocity(boundary_nebb)
boundary_pressure_nebb = _wrap_pressure(boundary_nebb)
boundary_force_corrected_nebb = _wrap_force_corrected(boundary_nebb)


def _moving_wall_bb(f_before_stream, f, loc, ux_wall, uy_wall, uz_wall):
    """in_k := pre[out_k] + 2 w_in rho_pre (c_in . u_w)/cs2 — lbm3d/boundary/bb.py:9-34."""
    s = WALLS3[loc]
    pre = _face(f_before_stream, loc)
    rho = pre.sum(axis=0).astype(f.dtype)
    _, uw = broadcast_wall_values(f, loc, 1, ux_wall, uy_wall, uz_wall)
    ind = s["in_dirs"]
    cin = C3[ind].astype(f.dtype)
    cu = np.tensordot(cin, uw.astype(f.dtype), axes=([1], [0]))
    wi = W3[ind].astype(f.dtype).reshape((len(ind),) + (1,) * (uw.ndim - 1))
    corr = 2.0 * wi * rho[None] * cu / f.dtype.type(CS2)
    w = _face(f, loc).copy()
    w[ind] = pre[s["out_dirs"]] + corr
    return _set_face(f, loc, w)


def boundary_bounce_back(f_before_stream, f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """lbm3d/boundary/bb.py:37-53."""
    return _moving_wall_bb(f_before_stream, f, loc, ux_wall, uy_wall, uz_wall)


def boundary_specular_reflection(f_before_stream, f, loc, ux_wall=0, uy_wall=0, uz_wall=0):
    """Identical body to bounce-back in the reference — lbm3d/boundary/bb.py:37-53."""
    return _moving_wall_bb(f_before_stream, f, loc, ux_wall, uy_wall, uz_wall)


def obstacle_bounce_back(f, mask):
    """lbm3d/boundary/bb.py:56-59."""
    out = f.copy()
    m = np.asarray(mask, dtype=bool)
    out[:, m] = f[:, m][OPP3]
    return out


def boundary_characteristic(rho, u, loc="right"):
    """lbm3d/boundary/cbc.py:16-51."""
    if loc not in WALLS3:
        raise ValueError("loc must be one of: 'left', 'right', 'bottom', 'top', 'back', 'front'")
    s = WALLS3[loc]
    na, ta, sg = s["axis"], s["tan_axes"], s["sign"]
    dt = _ft(rho, u)
    cs = dt.type(np.sqrt(CS2))

    def take3(field):
        idx = [0, 1, 2] if sg > 0 else [-1, -2, -3]
        return np.moveaxis(np.take(field, idx, axis=na), na, 0)

    r1, r2, r3 = take3(rho)
    n1, n2, n3 = take3(u[na])
    t0 = take3(u[ta[0]])[0]
    t1 = take3(u[ta[1]])[0]
    k = -0.5 * sg
    drho = k * (3 * r1 - 4 * r2 + r3)
    dun = k * (3 * n1 - 4 * n2 + n3)
    ch = dun - sg * cs / r1 * drho
    lo = (n1 - sg * cs) * ch
    rho_next = r1 - 0.5 * r1 / cs * lo
    un_next = n1 - 0.5 * lo
    vel = [None, None, None]
    vel[na] = un_next
    vel[ta[0]] = t0
    vel[ta[1]] = t1
    return rho_next, np.stack(vel)
