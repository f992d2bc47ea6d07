"""Oracle step compositions: the user-side ``update_step`` functions of the
reference examples, composed from the restated pure functions.

* ``lbm_step2d``  — macroscopic -> equilibrium -> collision -> forcing -> streaming
  -> ordered walls (README.md:105-121; examples/2d/lid_driven_cavity.py:55-64;
  examples/2d/poiseuille_channel.py:80-128 for the forcing recipes).
* ``VIV2D``       — the IB-LBM vortex-induced-vibration step
  (examples/2d/vortex_induced_vibration.py:96-148) with the collision / forcing
  recipe selectable (MRT + Guo per poiseuille_channel.py:113-128, or EDM).
* ``lbm_step3d`` / ``Sphere3D`` — examples/3d/flow_past_sphere.py:139-171.
"""

import math
from dataclasses import dataclass, field

import numpy as np

from . import lbm2d as L2
from . import lbm3d as L3
from . import ib as IB
from . import dyn as DYN

_PLAIN = {"nee": "boundary_nee", "nebb": "boundary_nebb", "eq": "boundary_equilibrium"}
_PREFIX = {"velocity": "boundary_velocity_", "pressure": "boundary_pressure_",
           "force_corrected": "boundary_force_corrected_"}
_SUFFIX = {"nee": "nee", "nebb": "nebb", "eq": "equilibrium"}


def wall_fn(mod, kind, mode):
    """The reference function for a (kind, mode) wall, e.g. ('nee', 'pressure') ->
    boundary_pressure_nee (lbm/boundary/nee.py:62-65)."""
    if mode == "plain":
        return getattr(mod, _PLAIN[kind])
    return getattr(mod, _PREFIX[mode] + _SUFFIX[kind])


def apply_walls(mod, f, f_pre, walls):
    """Apply an ordered wall list to the streamed field (each call rewrites its whole line)."""
    for w in walls:
        kind, mode, loc = w["kind"], w.get("mode", "plain"), w["loc"]
        if kind in ("bb", "specular"):
            fn = mod.boundary_bounce_back if kind == "bb" else mod.boundary_specular_reflection
            vel = {k: w[k] for k in ("ux_wall", "uy_wall", "uz_wall") if k in w}
            f = fn(f_pre, f, loc, **vel)
            continue
        kw = {k: w[k] for k in ("rho_wall", "ux_wall", "uy_wall", "uz_wall",
                                 "gx_wall", "gy_wall", "gz_wall") if k in w}
        if mode == "velocity":
            kw.pop("rho_wall", None)
        if mode == "pressure":
            kw = {"rho_wall": kw["rho_wall"]} if "rho_wall" in kw else {}
        if mode != "force_corrected":
            for k in ("gx_wall", "gy_wall", "gz_wall"):
                kw.pop(k, None)
        f = wall_fn(mod, kind, mode)(f, loc, **kw)
    return f


def collide(mod, f, feq, collision, omega, mrt_op=None):
    if collision == "bgk":
        return mod.collision_bgk(f, feq, omega)
    if collision == "mrt":
        return mod.collision_mrt(f, feq, mrt_op)
    if collision == "kbc":
        return mod.collision_kbc(f, feq, omega)
    if collision == "reg":
        return mod.collision_reg(f, feq, omega)
    raise ValueError(collision)


def collide_force(mod, f, rho, u, g, collision, forcing, omega, mrt_op=None, mrt_fop=None):
    """Collision + forcing for one step.  Guo uses the corrected velocity u + g/(2 rho)
    in both the equilibrium and the forcing term; EDM uses the uncorrected u
    (poiseuille_channel.py:80-128)."""
    if forcing is None or g is None:
        return collide(mod, f, mod.get_equilibrium(rho, u), collision, omega, mrt_op)
    if forcing == "guo":
        uc = u + mod.get_velocity_correction(g, rho)
        f = collide(mod, f, mod.get_equilibrium(rho, uc), collision, omega, mrt_op)
        if collision == "mrt":
            return mod.forcing_guo_mrt(f, g, uc, mrt_fop)
        return mod.forcing_guo_bgk(f, g, uc, omega)
    if forcing == "edm":
        f = collide(mod, f, mod.get_equilibrium(rho, u), collision, omega, mrt_op)
        return mod.forcing_edm(f, g, u)
    raise ValueError(forcing)


def lbm_step2d(f, omega, walls=(), collision="bgk", forcing=None, g=None, mrt_op=None, mrt_fop=None):
    """One D2Q9 step (no IB)."""
    rho, u = L2.get_macroscopic(f)
    f = collide_force(L2, f, rho, u, g, collision, forcing, omega, mrt_op, mrt_fop)
    fs = L2.streaming(f)
    return apply_walls(L2, fs, f, walls)


def lbm_step3d(f, omega, walls=(), collision="bgk", forcing=None, g=None, mrt_op=None, mrt_fop=None):
    rho, u = L3.get_macroscopic(f)
    f = collide_force(L3, f, rho, u, g, collision, forcing, omega, mrt_op, mrt_fop)
    fs = L3.streaming(f)
    return apply_walls(L3, fs, f, walls)


def cavity_walls(U0, order=("left", "right", "bottom", "top")):
    """Lid-driven cavity: NEE on all four walls, lid velocity on top (README.md:117-120 order)."""
    return [dict(kind="nee", mode="plain", loc=loc, **({"ux_wall": U0} if loc == "top" else {}))
            for loc in order]


# ====================================================================== VIV (2D IB-LBM)

@dataclass
class VIV2D:
    """Configuration of the 2D vortex-induced-vibration IB-LBM case
    (examples/2d/vortex_induced_vibration.py:31-74, defaults: the 1024 x 1024 benchmark case).

    ``dof``: which rigid-body DOFs move (x, y); inactive DOFs stay at zero.
    """
    nx: int = 1024
    ny: int = 1024
    D: float = 40.0
    cx: float = 256.0
    cy: float = 512.0
    n_markers: int = 512
    U0: float = 0.05
    Re: float = 100.0
    Ur: float = 5.0
    mass_ratio: float = 2.0
    damping_ratio: float = 0.0
    n_iter: int = 3
    pad: int = 10
    collision: str = "mrt"
    forcing: str = "guo"
    kernel: str = "peskin_4pt"
    dof: tuple = (False, True)
    inlet: str = "eq"            # left: boundary_equilibrium(ux_wall=U0)
    outlet: str = "pressure_nee"  # right: boundary_pressure_nee(rho_wall=1)
    seed: int = 0
    perturb: float = 0.01

    # derived -------------------------------------------------------------
    @property
    def nu(self):
        return self.U0 * self.D / self.Re

    @property
    def omega(self):
        return 1 / (3 * self.nu + 0.5)

    @property
    def area(self):
        return math.pi * (self.D / 2) ** 2

    @property
    def mass(self):
        return self.area * self.mass_ratio

    @property
    def fn(self):
        return self.U0 / (self.Ur * self.D)

    @property
    def stiffness(self):
        return (2 * math.pi * self.fn) ** 2 * self.mass * (1 + 1 / self.mass_ratio)

    @property
    def damping(self):
        return 2 * math.sqrt(self.stiffness * self.mass) * self.damping_ratio

    @property
    def ib_x0(self):
        return int(self.cx - 0.5 * self.D - self.pad)

    @property
    def ib_y0(self):
        return int(self.cy - 0.5 * self.D - self.pad)

    @property
    def ib_size(self):
        return int(self.D + 2 * self.pad)

    @property
    def ds(self):
        return math.pi * self.D / self.n_markers

    def markers(self):
        """Marker coordinates (float64), theta_k = 2 pi k / N (vortex_induced_vibration.py:40-44)."""
        th = 2 * np.pi * np.arange(self.n_markers) / self.n_markers
        return self.cx + 0.5 * self.D * np.cos(th), self.cy + 0.5 * self.D * np.sin(th)

    def walls(self):
        w = []
        if self.inlet == "eq":
            w.append(dict(kind="eq", mode="plain", loc="left", ux_wall=self.U0))
        elif self.inlet == "fc_nebb":
            w.append(dict(kind="nebb", mode="force_corrected", loc="left", ux_wall=self.U0))
        if self.outlet == "pressure_nee":
            w.append(dict(kind="nee", mode="pressure", loc="right", rho_wall=1.0))
        elif self.outlet == "eq":
            w.append(dict(kind="eq", mode="plain", loc="right", ux_wall=self.U0))
        return w

    def initial_f(self, dtype=np.float64):
        """rho = 1, u = U0 (1 + eps_x, eps_y), eps ~ U(-p, p) seeded; f = feq."""
        rng = np.random.default_rng(self.seed)
        eps = rng.uniform(-self.perturb, self.perturb, size=(2, self.nx, self.ny))
        u = np.empty((2, self.nx, self.ny))
        u[0] = self.U0 * (1 + eps[0])
        u[1] = self.U0 * eps[1]
        rho = np.ones((self.nx, self.ny))
        return L2.get_equilibrium(rho.astype(dtype), u.astype(dtype))


def viv_window_origin(case, d):
    """int32(IB_X0 + d) truncation (vortex_induced_vibration.py:104-105), evaluated in FP64."""
    return int(math.trunc(case.ib_x0 + float(d[0]))), int(math.trunc(case.ib_y0 + float(d[1])))


def viv_window_clamp(case, x0, y0):
    """Start of the window slice as lax.dynamic_slice / dynamic_update_slice place it:
    the start is clamped so that the whole window lies inside the grid
    (vortex_induced_vibration.py:108-109, 141).  The marker coordinates stay
    relative to the unclamped origin (vortex_induced_vibration.py:116-117)."""
    S = case.ib_size
    return min(max(x0, 0), case.nx - S), min(max(y0, 0), case.ny - S)


def viv_ib(case, u, d, v, dtype):
    """IB part of the VIV step: window, stencil, MDF.  Returns (xc, yc, g_win, h_markers)
    with (xc, yc) the clamped window start."""
    S = case.ib_size
    x0, y0 = viv_window_origin(case, d)
    xc, yc = viv_window_clamp(case, x0, y0)
    mx, my = case.markers()
    rx = (mx + d[0] - x0).astype(dtype)
    ry = (my + d[1] - y0).astype(dtype)
    W, I = IB.get_ib_stencil(rx, ry, S, IB.KERNELS[case.kernel], 2)
    ib_u = u[:, xc:xc + S, yc:yc + S]
    target = np.repeat(np.asarray(v, dtype=np.float64)[None, :], case.n_markers, axis=0).astype(dtype)
    g_win, h_m = IB.multi_direct_forcing(ib_u, W, I, target, np.asarray(case.ds, dtype=dtype), case.n_iter)
    return xc, yc, g_win, h_m


def viv_newmark(case, a, v, d, h_markers):
    """h = sum h_m + a * area; Newmark on the active DOFs (FP64) — vortex_induced_vibration.py:135-137."""
    h = np.sum(h_markers.astype(np.float64), axis=0) + a * case.area
    an, vn, dn = DYN.newmark(a, v, d, h, case.mass, case.stiffness, case.damping)
    mask = np.array(case.dof, dtype=bool)
    return np.where(mask, an, 0.0), np.where(mask, vn, 0.0), np.where(mask, dn, 0.0), h


def viv_step(case, f, d, v, a, mrt_op=None, mrt_fop=None):
    """One VIV IB-LBM step.  f in the working dtype, (d, v, a) FP64.  Returns (f, d, v, a, h)."""
    dt = f.dtype
    rho, u = L2.get_macroscopic(f)
    x0, y0, g_win, h_m = viv_ib(case, u, d, v, dt)
    a, v, d, h = viv_newmark(case, a, v, d, h_m)
    S = case.ib_size
    g = np.zeros((2, case.nx, case.ny), dtype=dt)
    g[:, x0:x0 + S, y0:y0 + S] = g_win
    omega = case.omega
    if case.collision == "mrt":
        mrt_op = L2.get_mrt_collision_operator(omega, dt.type) if mrt_op is None else mrt_op
        mrt_fop = L2.get_mrt_forcing_operator(omega, dt.type) if mrt_fop is None else mrt_fop
    f = collide_force(L2, f, rho, u, g, case.collision, case.forcing, omega, mrt_op, mrt_fop)
    fs = L2.streaming(f)
    return apply_walls(L2, fs, f, case.walls()), d, v, a, h


def viv_run(case, f, n_steps, dtype=None, d=None, v=None, a=None):
    """n_steps VIV steps; returns (f, d, v, a, d_hist (n,2), h_hist (n,2))."""
    f = np.asarray(f) if dtype is None else np.asarray(f, dtype=dtype)
    d = np.zeros(2) if d is None else np.asarray(d, dtype=np.float64)
    v = np.zeros(2) if v is None else np.asarray(v, dtype=np.float64)
    a = np.zeros(2) if a is None else np.asarray(a, dtype=np.float64)
    dh, hh = [], []
    for _ in range(n_steps):
        f, d, v, a, h = viv_step(case, f, d, v, a)
        dh.append(d.copy())
        hh.append(h.copy())
    return f, d, v, a, np.array(dh).reshape(-1, 2), np.array(hh).reshape(-1, 2)


# ====================================================================== sphere (3D IB-LBM)

@dataclass
class Sphere3D:
    """Flow past a fixed sphere (examples/3d/flow_past_sphere.py:29-124), defaults: the
    256 x 128 x 128 benchmark case; markers on a Fibonacci lattice, uniform ds = pi D^2 / N."""
    nx: int = 256
    ny: int = 128
    nz: int = 128
    D: float = 24.0
    cx: float = 64.0
    cy: float = 64.0
    cz: float = 64.0
    n_markers: int = 4096
    U0: float = 0.05
    Re: float = 200.0
    n_iter: int = 2
    pad: int = 4
    collision: str = "bgk"
    forcing: str = "edm"
    kernel: str = "peskin_4pt"
    inlet: str = "eq"
    outlet: str = "pressure_nee"

    @property
    def nu(self):
        return self.U0 * self.D / self.Re

    @property
    def omega(self):
        return 1 / (3 * self.nu + 0.5)

    @property
    def ib_origin(self):
        h = 0.5 * self.D + self.pad
        return int(self.cx - h), int(self.cy - h), int(self.cz - h)

    @property
    def ib_size(self):
        return int(self.D + 2 * self.pad)

    @property
    def ds(self):
        return math.pi * self.D ** 2 / self.n_markers

    def markers(self):
        """Fibonacci lattice: z_k = 1 - (2k+1)/N, phi_k = k pi (3 - sqrt 5)."""
        k = np.arange(self.n_markers)
        z = 1 - (2 * k + 1) / self.n_markers
        r = np.sqrt(1 - z * z)
        ph = k * np.pi * (3 - np.sqrt(5.0))
        R = 0.5 * self.D
        return np.stack([self.cx + R * r * np.cos(ph), self.cy + R * r * np.sin(ph), self.cz + R * z], axis=1)

    def walls(self):
        w = []
        if self.inlet == "eq":
            w.append(dict(kind="eq", mode="plain", loc="left", ux_wall=self.U0))
        elif self.inlet == "nebb":
            w.append(dict(kind="nebb", mode="plain", loc="left", ux_wall=self.U0))
        if self.outlet == "pressure_nee":
            w.append(dict(kind="nee", mode="pressure", loc="right", rho_wall=1.0))
        elif self.outlet == "eq":
            w.append(dict(kind="eq", mode="plain", loc="right", ux_wall=self.U0))
        return w

    def stencil(self, dtype=np.float64):
        S = self.ib_size
        o = np.array(self.ib_origin, dtype=np.float64)
        rel = (self.markers() - o).astype(dtype)
        return IB.get_ib_stencil3d(rel, (S, S, S), IB.KERNELS[self.kernel], 2)

    def initial_f(self, dtype=np.float64):
        u = np.zeros((3, self.nx, self.ny, self.nz), dtype=dtype)
        u[0] = self.U0
        return L3.get_equilibrium(np.ones((self.nx, self.ny, self.nz), dtype=dtype), u)


def sphere_step(case, f, W=None, I=None, mrt_op=None, mrt_fop=None):
    """One sphere IB-LBM step; returns (f, total marker force (3,) FP64)."""
    dt = f.dtype
    if W is None:
        W, I = case.stencil(dt)
    rho, u = L3.get_macroscopic(f)
    S = case.ib_size
    x0, y0, z0 = case.ib_origin
    ib_u = u[:, x0:x0 + S, y0:y0 + S, z0:z0 + S]
    target = np.zeros((case.n_markers, 3), dtype=dt)
    g_win, h_m = IB.multi_direct_forcing(ib_u, W, I, target, np.asarray(case.ds, dtype=dt), case.n_iter)
    g = np.zeros((3, case.nx, case.ny, case.nz), dtype=dt)
    g[:, x0:x0 + S, y0:y0 + S, z0:z0 + S] = g_win
    omega = case.omega
    if case.collision == "mrt":
        mrt_op = L3.get_mrt_collision_operator(omega, dt.type) if mrt_op is None else mrt_op
        mrt_fop = L3.get_mrt_forcing_operator(omega, dt.type) if mrt_fop is None else mrt_fop
    f = collide_force(L3, f, rho, u, g, case.collision, case.forcing, omega, mrt_op, mrt_fop)
    fs = L3.streaming(f)
    return apply_walls(L3, fs, f, case.walls()), np.sum(h_m.astype(np.float64), axis=0)
