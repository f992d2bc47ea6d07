"""Fused time-step drivers.

The reference leaves the time step to the user, who composes the pure
functions and compiles a chunk of steps with ``jax.jit`` + ``lax.scan``
(README.md:105-121; examples/2d/vortex_induced_vibration.py:96-157).  Here the
same compositions run as fused HIP passes over device-resident buffers:

* ``LBM`` — macroscopic -> equilibrium -> collision -> forcing -> streaming ->
  ordered walls, for D2Q9 or D3Q19 (``vsx_d2q9_run`` / ``vsx_d3q19_run``).
* ``VIV2D`` — the 2D IB-LBM vortex-induced-vibration step with multi-direct
  forcing, FP64 body-force reduction and FP64 Newmark (``vsx_viv2d_run``).
* ``Sphere3D`` — the 3D IB-LBM step past a fixed body (``vsx_sphere3d_run``).

Wall lists use the dictionaries ``{"kind": "nee"|"nebb"|"eq"|"bb"|"specular",
"mode": "plain"|"velocity"|"pressure"|"force_corrected", "loc": ...,
"rho_wall": .., "ux_wall": .., ...}``, applied in list order after streaming
exactly like consecutive ``boundary_*`` calls.
"""

import ctypes
from ctypes import c_void_p
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from ._util import WallBuilder, f32, numel, require, wall_nodes_shape

_STORE_DTYPE = {"f32": torch.float32, "f16s": torch.float16, "bf16s": torch.bfloat16, "e3m12s": torch.int16}

MRT_RATES_2D = lambda om: [0, 1.4, 1.4, 0, 1.2, 0, 1.2, om, om]                          # lbm/collision/mrt.py:44
MRT_RATES_3D = lambda om: [0, 0, 0, 0, 1.1] + [om] * 5 + [1.2] * 6 + [1.4] * 3          # lbm3d/collision/mrt.py:92-113


class LBM:
    """A fused LBM step configuration for a grid of ``shape`` (2 or 3 axes)."""

    def __init__(self, shape, omega, collision="bgk", walls=(), forcing=None, g=None, g_uniform=None,
                 storage="f32", mrt_rates=None, store_u=None):
        """store_u: reference velocity of the 16-bit storages, which keep
        f_i - feq_i(1, store_u) (default 0: f_i - w_i); a velocity near the
        flow's keeps the stored numbers small and their rounding error too."""
        shape = tuple(int(s) for s in shape)
        require(len(shape) in (2, 3), "shape must have 2 or 3 axes")
        self.shape, self.dim = shape, len(shape)
        self.q = 9 if self.dim == 2 else 19
        self.pfx = "d2q9" if self.dim == 2 else "d3q19"
        self.storage = storage
        p = L.VsxLbm()
        p.nx, p.ny = shape[0], shape[1]
        p.nz = shape[2] if self.dim == 3 else 1
        p.collision = L.COLLISION[collision]
        p.forcing = L.FORCING[forcing]
        p.storage = L.STORE[storage]
        p.omega = float(omega)
        rates = mrt_rates if mrt_rates is not None else (MRT_RATES_2D if self.dim == 2 else MRT_RATES_3D)(float(omega))
        for i, r in enumerate(rates):
            p.mrt_rates[i] = float(r)
        self._wb = WallBuilder()
        require(len(walls) <= (4 if self.dim == 2 else 6), "too many walls")
        locs = ("left", "right", "top", "bottom") if self.dim == 2 else ("left", "right", "bottom", "top", "back", "front")
        p.n_walls = len(walls)
        for k, w in enumerate(walls):
            if w["loc"] not in locs:
                raise KeyError(w["loc"])
            vals = {"rho": w.get("rho_wall", 1.0), "ux": w.get("ux_wall", 0.0), "uy": w.get("uy_wall", 0.0),
                    "uz": w.get("uz_wall", 0.0), "gx": w.get("gx_wall", 0.0), "gy": w.get("gy_wall", 0.0),
                    "gz": w.get("gz_wall", 0.0)}
            p.walls[k] = self._wb(w["kind"], w.get("mode", "plain"), w["loc"], wall_nodes_shape(shape, w["loc"]), **vals)
        self._g = None
        if g is not None:
            self._g = f32(g)
            require(tuple(self._g.shape) == (self.dim,) + shape, "g must have shape (D, *grid)")
            p.g = self._g.data_ptr()
        for a, v in enumerate(g_uniform or ()):
            p.g_uniform[a] = float(v)
        for a, v in enumerate(store_u or ()):
            p.store_u[a] = float(v)
        self.p = p
        self._work = None

    @property
    def cells(self):
        return numel(self.shape)

    def work(self):
        if self._work is None:
            dt = _STORE_DTYPE[self.storage]
            self._work = [torch.empty(self.q * self.cells, dtype=dt, device=L.device()) for _ in range(2)]
        return self._work

    def run(self, f, n_steps, out=None):
        """f_out = step^n_steps(f) for a post-streaming field f[Q, *grid] (float32)."""
        f = f32(f)
        require(tuple(f.shape) == (self.q,) + self.shape, f"f must have shape {(self.q,) + self.shape}")
        out = torch.empty_like(f) if out is None else out
        wa, wb = self.work()
        L.call(f"vsx_{self.pfx}_run", ctypes.byref(self.p), L.ptr(f), L.ptr(out), L.ptr(wa), L.ptr(wb), int(n_steps))
        return out

    def step(self, src, dst, phase, src_f32=False, dst_f32=False):
        """One pass of the pull scheme (phase 'collide' | 'full' | 'stream')."""
        L.call(f"vsx_{self.pfx}_step", ctypes.byref(self.p), L.ptr(src), int(src_f32), L.ptr(dst), int(dst_f32),
               L.PHASE[phase])


# ====================================================================== VIV 2D

@dataclass
class VIVCase:
    """2D vortex-induced vibration of a cylinder on springs
    (examples/2d/vortex_induced_vibration.py:31-74 for the geometry / structure
    definitions).  ``dof`` selects the moving DOFs (x, y)."""
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
    inlet: str = "eq"
    outlet: str = "pressure_nee"
    seed: int = 0
    perturb: float = 0.01

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
    def stiffness(self):
        fn = self.U0 / (self.Ur * self.D)
        return (2 * math.pi * fn) ** 2 * self.mass * (1 + 1 / self.mass_ratio)

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

    def initial_u(self):
        """u = U0 (1 + eps_x, eps_y), eps ~ U(-p, p) from default_rng(seed) (float64, host)."""
        rng = np.random.default_rng(self.seed)
        eps = rng.uniform(-self.perturb, self.perturb, size=(2, self.nx, self.ny))
        u = np.empty((2, self.nx, self.ny))
        u[0] = self.U0 * (1 + eps[0])
        u[1] = self.U0 * eps[1]
        return u


class _TimeoutWatch:
    """Sticky check of the bounded-wait time-out flag of fused steps without a
    host sync: after each run the flag is copied asynchronously into pinned
    host memory behind an event; the next run (or an explicit ``check()``)
    raises if a completed copy shows a time-out."""

    def __init__(self, flags):
        self.flags = list(flags)            # int32 device scalars (views)
        self.host = torch.zeros(len(self.flags), dtype=torch.int32, pin_memory=True)
        self.event = None

    def after_run(self):
        for k, fl in enumerate(self.flags):
            self.host[k:k + 1].copy_(fl, non_blocking=True)
        self.event = torch.cuda.Event()
        self.event.record()

    def poll(self, wait=False):
        if self.event is None:
            return
        if wait:
            self.event.synchronize()
        elif not self.event.query():
            return
        _raise_ib_codes([int(self.host[k]) for k in range(len(self.flags))])


def _raise_ib_codes(codes):
    """The error codes of fused steps (1: a bounded wait ran out, 2: the
    pipelined window left the hand-off buffer), one per case."""
    if any(c == 2 for c in codes):
        bad = [k for k, c in enumerate(codes) if c == 2]
        raise RuntimeError(f"pipelined VIV step, case(s) {bad}: the IB window came closer than 2 cells to the "
                           "grid edge or moved by more than a cell in a step; run with pipelined=False")
    bad = [k for k, c in enumerate(codes) if c != 0]
    if bad:
        raise RuntimeError(f"fused VIV step: a tile of case(s) {bad} timed out waiting for the IB workgroup "
                           "(the force window of that step was not applied)")


class _Body2D:
    """Device buffers of one 2D immersed body and the VsxIB2D that points at them."""

    def __init__(self, case, dev):
        mx, my = case.markers()
        self.marker_x = torch.as_tensor(mx, dtype=torch.float64, device=dev)
        self.marker_y = torch.as_tensor(my, dtype=torch.float64, device=dev)
        S = case.ib_size
        self.g_win = torch.zeros(2 * S * S, dtype=torch.float32, device=dev)
        # hand-off buffer of the pipelined step: two halves of [9][S + 4][S + 4]
        self.halo = torch.zeros(2 * 9 * (S + 4) ** 2, dtype=torch.float32, device=dev)
        # origin x, y of the last step's window, ready flag, spin-timeout flag, window-tile counters
        self.win_origin = torch.zeros(8, dtype=torch.int32, device=dev)
        # two slots of [a0 a1 v0 v1 d0 d1]: slot (t & 1) holds the state before step t
        self.body = torch.zeros((2, 6), dtype=torch.float64, device=dev)
        ib = L.VsxIB2D()
        ib.n_markers = case.n_markers
        ib.kernel = L.KERNEL[case.kernel]
        ib.n_iter = case.n_iter
        ib.win = S
        ib.marker_x, ib.marker_y = self.marker_x.data_ptr(), self.marker_y.data_ptr()
        ib.ds_arr = None
        ib.ds = case.ds
        ib.dof_mask = int(bool(case.dof[0])) | (int(bool(case.dof[1])) << 1)
        ib.win_x0, ib.win_y0 = float(case.ib_x0), float(case.ib_y0)
        ib.mass, ib.stiffness, ib.damping, ib.area = case.mass, case.stiffness, case.damping, case.area
        ib.dt, ib.gamma, ib.beta = 1.0, 0.5, 0.25
        ib.body = self.body.data_ptr()
        ib.g_win = self.g_win.data_ptr()
        ib.win_origin = self.win_origin.data_ptr()
        ib.halo = self.halo.data_ptr()
        # error word the kernels store into directly (pinned host memory): read
        # without a copy or a sync
        self.status = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self.status_np = self.status.numpy()
        ib.status = self.status.data_ptr()
        self.ib = ib

    def set(self, a, v, d, slot=0):
        self.body.zero_()
        self.body[slot].copy_(torch.tensor(list(a) + list(v) + list(d), dtype=torch.float64))
        self.win_origin.zero_()
        torch.cuda.current_stream().synchronize()
        self.status_np[0] = 0


class VIV2D:
    """Device state and fused run of a ``VIVCase``.  16-bit storages keep the
    populations relative to the equilibrium of the inflow (U0, 0).

    fused=False: separate IB and bulk launches per step.  fused=True with
    pipelined=True: the IB of step t + 1 runs as its own kernel on a second
    stream while the bulk pass of step t streams (the IB window must stay at
    least 2 cells inside the grid; a run that breaks this raises on the next
    call); pipelined=False: one launch per step, the IB workgroup beside the
    bulk tiles (any window position).  pipelined=None picks the pipelined form
    when the initial window is at least 3 cells inside the grid and the IB
    kernel's LDS (27 S^2 bytes for window side S) fits the device."""

    def __init__(self, case: VIVCase, storage="f32", fused=True, pipelined=None):
        self.case = case
        dev = L.device()
        self.lbm = LBM((case.nx, case.ny), case.omega, collision=case.collision, walls=case.walls(),
                       forcing=case.forcing, storage=storage, store_u=(case.U0, 0.0))
        self._b = _Body2D(case, dev)
        self.body, self.win_origin, self.g_win = self._b.body, self._b.win_origin, self._b.g_win
        self.t_total = 0
        self.ib = self._b.ib
        if pipelined is None:
            S = case.ib_size
            lds = 4 * (6 * S * S + (S * S + 3) // 4 + (S * S + 1) // 2) + 2048
            pipelined = min(case.ib_x0, case.ib_y0) >= 3 and case.ib_x0 + S + 3 <= case.nx and \
                case.ib_y0 + S + 3 <= case.ny and lds <= torch.cuda.get_device_properties(dev).shared_memory_per_block
        self.ib.fused = (2 if pipelined else 1) if fused else 0
        self._run_args = None   # bound entry point and the pointers that stay fixed between runs

    def set_body(self, a=(0.0, 0.0), v=(0.0, 0.0), d=(0.0, 0.0)):
        """Reset the structure state (and the step counter of the body slots)."""
        self._b.set(a, v, d)
        self.t_total = 0

    def state(self):
        b = self.body[self.t_total & 1].cpu().numpy()
        self._poll()
        return b[0:2].copy(), b[2:4].copy(), b[4:6].copy()   # a, v, d

    def _poll(self):
        code = int(self._b.status_np[0])
        if code:
            _raise_ib_codes([code])

    def check(self):
        """Raise if a tile of a fused step timed out waiting for the IB workgroup
        (waits for the runs queued so far).  ``run`` raises the same error, without
        waiting, once the kernels' error word of an earlier run has reached the host."""
        torch.cuda.current_stream().synchronize()
        self._poll()

    def run(self, f, n_steps, out=None, history=True):
        """Advance n_steps; returns (f_out, d_hist [n,2], h_hist [n,2]) (histories FP64 on device).
        ``history`` may also be a pair of caller-owned FP64 [>= n_steps, 2] device
        tensors (d, h) to fill, which spares the allocations in a loop of runs."""
        self._poll()
        if not (isinstance(f, torch.Tensor) and f.is_cuda and f.dtype == torch.float32 and f.is_contiguous()):
            f = f32(f)
        out = torch.empty_like(f) if out is None else out
        hd = hh = None
        if isinstance(history, tuple):
            hd, hh = history
            require(hd.dtype == hh.dtype == torch.float64 and hd.shape[0] >= n_steps and hh.shape[0] >= n_steps
                    and hd.is_contiguous() and hh.is_contiguous(), "history buffers must be contiguous FP64 [n, 2]")
        elif history and n_steps > 0:
            hd = torch.empty((n_steps, 2), dtype=torch.float64, device=f.device)
            hh = torch.empty((n_steps, 2), dtype=torch.float64, device=f.device)
        self.ib.hist_d = None if hd is None else hd.data_ptr()
        self.ib.hist_h = None if hh is None else hh.data_ptr()
        if self._run_args is None:
            wa, wb = self.lbm.work()
            self._run_args = (L.lib().vsx_viv2d_run, ctypes.byref(self.lbm.p), ctypes.byref(self.ib),
                              L.ptr(wa), L.ptr(wb))
        fn, p, ib, wa, wb = self._run_args
        self.ib.step0 = self.t_total
        L.check(fn(p, ib, c_void_p(f.data_ptr()), c_void_p(out.data_ptr()), wa, wb, int(n_steps), L.stream()))
        self.t_total += int(n_steps)
        return out, hd, hh


# ====================================================================== sphere 3D

@dataclass
class SphereCase:
    """Flow past a fixed sphere (examples/3d/flow_past_sphere.py:29-124 for the
    geometry and window definitions): a D3Q19 step with an IB window of side
    D + 2 pad around the body; markers on a Fibonacci lattice with uniform
    ds = pi D^2 / N."""
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
        """Fibonacci lattice: z_k = 1 - (2k+1)/N, phi_k = k pi (3 - sqrt 5) (float64, host)."""
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

    def drag_coefficient(self, fx):
        """C_d = F_x / (0.5 rho U0^2 A), A = pi D^2 / 4, rho = 1."""
        return fx / (0.5 * self.U0 ** 2 * math.pi * self.D ** 2 / 4)


class Sphere3D:
    """Device state and fused run of a ``SphereCase`` (``vsx_sphere3d_run``).
    pipelined=True (the window 2+ cells inside the grid): the IB of step t + 1
    runs on a second stream while the bulk pass of step t streams.  The
    default is the sequential form: the sphere's IB kernels are multi-
    workgroup kernels that, run beside the bulk pass, get too few compute
    units and make the pipelined step slower (0.24 against 0.16 ms at
    256 x 128 x 128)."""

    def __init__(self, case: SphereCase, storage="f32", pipelined=False):
        from . import ib3d
        self.case = case
        dev = L.device()
        self.lbm = LBM((case.nx, case.ny, case.nz), case.omega, collision=case.collision, walls=case.walls(),
                       forcing=case.forcing, storage=storage, store_u=(case.U0, 0.0, 0.0))
        S = case.ib_size
        org = case.ib_origin
        # stencil of the fixed markers relative to the window origin (flow_past_sphere.py:114-121)
        rel = torch.as_tensor(case.markers() - np.array(org, dtype=np.float64), dtype=torch.float32, device=dev)
        self.W, self.I = ib3d.get_ib_stencil(rel, (S, S, S), kernel=case.kernel)
        cells = S ** 3
        m, s = self.W.shape
        self.cell_ptr = torch.empty(cells + 1, dtype=torch.int32, device=dev)
        self.cell_slot = torch.empty(m * s, dtype=torch.int32, device=dev)
        tmp = torch.empty(cells + 1, dtype=torch.int32, device=dev)
        L.call("vsx_ib3d_transpose_stencil", L.ptr(self.I), m, s, cells, L.ptr(self.cell_ptr), L.ptr(self.cell_slot),
               L.ptr(tmp))
        self.u_target = torch.zeros((m, 3), dtype=torch.float32, device=dev)
        self.g_win = torch.zeros(3 * cells, dtype=torch.float32, device=dev)
        self.work = torch.zeros(int(L.lib().vsx_ib3d_work_floats(m, cells)), dtype=torch.float32, device=dev)
        ib = L.VsxIB3D()
        ib.n_markers, ib.n_iter = m, case.n_iter
        for a in range(3):
            ib.origin[a] = org[a]
            ib.win[a] = S
        ib.W, ib.I = self.W.data_ptr(), self.I.data_ptr()
        ib.cell_ptr, ib.cell_slot = self.cell_ptr.data_ptr(), self.cell_slot.data_ptr()
        ib.u_target = self.u_target.data_ptr()
        ib.ds_arr = None
        ib.ds = case.ds
        ib.g_win, ib.work = self.g_win.data_ptr(), self.work.data_ptr()
        self.halo = torch.zeros(2 * 19 * (S + 2) ** 3 if pipelined else 1, dtype=torch.float32, device=dev)
        ib.pipelined = int(bool(pipelined))
        ib.halo = self.halo.data_ptr()
        self.ib = ib

    def initial_f(self):
        """Equilibrium at rho = 1, u = (U0, 0, 0) (flow_past_sphere.py:127-131)."""
        from . import lbm3d
        c = self.case
        u = torch.zeros((3, c.nx, c.ny, c.nz), dtype=torch.float32, device=L.device())
        u[0] = c.U0
        return lbm3d.get_equilibrium(torch.ones((c.nx, c.ny, c.nz), dtype=torch.float32, device=L.device()), u)

    def run(self, f, n_steps, out=None, history=True):
        """Advance n_steps; returns (f_out, h_hist [n, 3] FP64 total marker force on the fluid's markers)."""
        f = f32(f)
        out = torch.empty_like(f) if out is None else out
        hh = None
        if history and n_steps > 0:
            hh = torch.empty((n_steps, 3), dtype=torch.float64, device=f.device)
        self.ib.hist_h = None if hh is None else hh.data_ptr()
        wa, wb = self.lbm.work()
        L.call("vsx_sphere3d_run", ctypes.byref(self.lbm.p), ctypes.byref(self.ib), L.ptr(f), L.ptr(out), L.ptr(wa),
               L.ptr(wb), int(n_steps))
        self._check_flags()
        return out, hh

    def _check_flags(self):
        """Raise if a bounded wait of an earlier pipelined run ran out (no host sync: the
        flag word is copied asynchronously and read once its copy has landed)."""
        if not self.ib.pipelined:
            return
        if getattr(self, "_watch", None) is None:
            flag = self.work[7:8].view(torch.int32)
            self._watch = _TimeoutWatch([flag])
        self._watch.poll()
        self._watch.after_run()


# ====================================================================== batched VIV sweep

def default_tile_y(ny):
    """Bulk tile width along y: the contiguous axis rounded up to a wavefront, at most 256."""
    return min(256, ((ny + 63) // 64) * 64)


class BatchGraph:
    """A captured HIP graph of a batched run (``vsx_viv2d_batch_graph``); ``launch()``
    replays it on the current stream, each replay continuing from the batch clock."""

    def __init__(self, handle, refs):
        self._h = handle
        self._refs = refs     # keeps the captured buffers alive

    def launch(self):
        batch = self._refs[2]
        batch._watch.poll()
        L.call("vsx_graph_launch", self._h)
        batch._watch.after_run()

    def __del__(self):
        if getattr(self, "_h", None) and L._LIB is not None:
            L._LIB.vsx_graph_destroy(self._h)
            self._h = None


class VIVBatch:
    """Many ``VIVCase``s (grids of different sizes, different structures) advanced
    together, one launch per step for the whole batch (``vsx_viv2d_batch_*``).

    The populations of all cases live in one buffer: case k occupies
    ``[offset[k], offset[k] + 9 nx ny)`` as its ``f[9, nx, ny]``.
    ``tile_y``: None (default rule), "auto" (the autotuner in ``vivsim_amd.tune``)
    or an int applied to every case."""

    def __init__(self, cases, storage="f32", tile_y=None, hist_cap=0):
        from . import tune
        cases = list(cases)
        require(len(cases) > 0, "empty batch")
        dev = L.device()
        self.cases, self.storage, self.hist_cap = cases, storage, int(hist_cap)
        self.lbms = [LBM((c.nx, c.ny), c.omega, collision=c.collision, walls=c.walls(), forcing=c.forcing,
                         storage=storage, store_u=(c.U0, 0.0)) for c in cases]
        self.bodies = [_Body2D(c, dev) for c in cases]
        self.offsets, off = [], 0
        for c in cases:
            self.offsets.append(off)
            off += ((9 * c.nx * c.ny + 63) // 64) * 64
        self.total = off
        if tile_y == "auto":
            self.tile_y = [tune.best_tile_y(c, storage) for c in cases]
        elif tile_y is None:
            self.tile_y = [default_tile_y(c.ny) for c in cases]
        else:
            self.tile_y = [int(tile_y)] * len(cases)
        self.hist_d = [torch.zeros((max(1, self.hist_cap), 2), dtype=torch.float64, device=dev) for _ in cases]
        self.hist_h = [torch.zeros((max(1, self.hist_cap), 2), dtype=torch.float64, device=dev) for _ in cases]
        vc = (L.VsxVivCase * len(cases))()
        for k, c in enumerate(cases):
            vc[k].lbm = self.lbms[k].p
            ib = self.bodies[k].ib
            ib.hist_d, ib.hist_h = self.hist_d[k].data_ptr(), self.hist_h[k].data_ptr()
            vc[k].ib = ib
            vc[k].ib.status = None   # the batch reads its error flags through its own watch
            vc[k].f_offset = self.offsets[k]
            vc[k].tile_y = self.tile_y[k]
            vc[k].hist_cap = self.hist_cap
        self._vc = vc
        nbytes = int(L.lib().vsx_viv2d_batch_table_bytes(vc, len(cases)))
        require(nbytes > 0, "bad batch description")
        self.table = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
        self.clock_t = torch.zeros(1, dtype=torch.int32, device=dev)
        b = L.VsxVivBatch()
        L.check(L.lib().vsx_viv2d_batch_prepare(vc, len(cases), L.ptr(self.table), L.ptr(self.clock_t),
                                                ctypes.byref(b)))
        self.batch = b
        self._watch = _TimeoutWatch([bd.win_origin[3:4] for bd in self.bodies])
        dt = _STORE_DTYPE[storage]
        self._work = [torch.empty(self.total, dtype=dt, device=dev) for _ in range(2)]

    @property
    def cells(self):
        return sum(c.nx * c.ny for c in self.cases)

    def view(self, f, k):
        c = self.cases[k]
        return f[self.offsets[k]:self.offsets[k] + 9 * c.nx * c.ny].view(9, c.nx, c.ny)

    def initial_f(self):
        """Every case's seeded perturbed inflow at equilibrium, packed into one buffer."""
        from . import lbm
        f = torch.zeros(self.total, dtype=torch.float32, device=L.device())
        for k, c in enumerate(self.cases):
            u = torch.as_tensor(c.initial_u(), dtype=torch.float32, device=f.device)
            self.view(f, k).copy_(lbm.get_equilibrium(torch.ones((c.nx, c.ny), device=f.device), u))
        return f

    def set_bodies(self, v=(0.0, 0.0)):
        """Every body at rest at its origin with velocity v; the batch clock restarts at 0."""
        for b in self.bodies:
            b.set((0.0, 0.0), v, (0.0, 0.0))
        self.clock_t.zero_()

    def clock(self):
        return int(self.clock_t.item())

    def state(self, k):
        b = self.bodies[k].body[self.clock() & 1].cpu().numpy()
        return b[0:2].copy(), b[2:4].copy(), b[4:6].copy()

    def check(self):
        """Raise if a tile of any case timed out waiting for its IB workgroup (waits for the queued work)."""
        self._watch.after_run()
        self._watch.poll(wait=True)

    def run(self, f, n_steps, out=None):
        f = f32(f)
        require(f.numel() == self.total, "f must be the packed batch buffer")
        self._watch.poll()
        out = torch.empty_like(f) if out is None else out
        wa, wb = self._work
        L.call("vsx_viv2d_batch_run", ctypes.byref(self.batch), L.ptr(f), L.ptr(out), L.ptr(wa), L.ptr(wb),
               int(n_steps))
        self._watch.after_run()
        return out

    def graph(self, f_in, f_out, n_steps):
        """Capture a run of n_steps from f_in to f_out into a HIP graph."""
        wa, wb = self._work
        h = ctypes.c_void_p()
        L.check(L.lib().vsx_viv2d_batch_graph(ctypes.byref(self.batch), L.ptr(f_in), L.ptr(f_out), L.ptr(wa),
                                              L.ptr(wb), int(n_steps), ctypes.byref(h)))
        return BatchGraph(h, (f_in, f_out, self))


# ====================================================================== adjoint (2D VIV)

class VIVAdjoint:
    """Gradient of J = (1/T) sum_{t=1..T} d_y(t)^2 over a ``VIVCase`` run with
    respect to the marker radial offsets delta [M], the spring stiffness and the
    inlet velocity (``vsx_viv2d_adjoint``: hand-written reverse pass with
    sqrt(T) checkpointing).  The case must be MRT + Guo with the equilibrium
    inlet and the pressure-NEE outlet; the initial field is the case's seeded
    inflow at the case's U0, the body starts at rest with velocity v0."""

    def __init__(self, case: VIVCase, n_steps, seg=None, v0=(0.0, 0.0)):
        require(case.collision == "mrt" and case.forcing == "guo", "the adjoint needs MRT collision with Guo forcing")
        require(case.inlet == "eq" and case.outlet == "pressure_nee", "the adjoint needs the eq inlet and NEE outlet")
        from . import lbm
        dev = L.device()
        self.case, self.T = case, int(n_steps)
        self.seg = int(seg) if seg else max(1, int(round(math.sqrt(self.T))))
        self.v0 = tuple(float(x) for x in v0)
        nseg = (self.T + self.seg - 1) // self.seg
        n = case.nx * case.ny
        S = case.ib_size
        M = case.n_markers
        u = torch.as_tensor(case.initial_u(), dtype=torch.float32, device=dev)
        self.f0 = lbm.get_equilibrium(torch.ones((case.nx, case.ny), device=dev), u).contiguous()
        th = 2 * np.pi * np.arange(M) / M
        self.dir_x = torch.as_tensor(np.cos(th), dtype=torch.float64, device=dev)
        self.dir_y = torch.as_tensor(np.sin(th), dtype=torch.float64, device=dev)
        z32 = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)
        z64 = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        self.buf = dict(ckpt=z32(nseg, 9 * n), ckpt_body=z64(nseg, 6), tape_f=z32(2, self.seg + 1, 9 * n),
                        tape_g=z32(2, self.seg, 2 * S * S), tape_body=z64(2, self.seg + 1, 6), adj_a=z32(9 * n),
                        adj_b=z32(9 * n), grid=z32(2 * S * S), adj_body=z64(8), hist_d=z64(self.T, 2),
                        u_in_part=z64(self.T), grad_delta=z64(M), out=z64(3),
                        sync=torch.zeros(4, dtype=torch.int32, device=dev))
        self.f_tmp = z32(9 * n)
        a = L.VsxAdj2D()
        a.n_steps, a.seg = self.T, self.seg
        a.dir_x, a.dir_y = self.dir_x.data_ptr(), self.dir_y.data_ptr()
        for k, t in self.buf.items():
            setattr(a, k, t.data_ptr())
        self.adj = a

    def _setup(self, delta, stiffness, u_in):
        c = self.case
        lb = LBM((c.nx, c.ny), c.omega, collision="mrt", forcing="guo",
                 walls=[dict(kind="eq", mode="plain", loc="left", ux_wall=float(u_in)),
                        dict(kind="nee", mode="pressure", loc="right", rho_wall=1.0)])
        body = _Body2D(c, L.device())
        d = torch.as_tensor(delta, dtype=torch.float64, device=L.device()).reshape(-1)
        require(d.numel() == c.n_markers, "delta must have one entry per marker")
        body.marker_x.add_(d * self.dir_x)
        body.marker_y.add_(d * self.dir_y)
        body.ib.stiffness = float(stiffness)
        body.set((0.0, 0.0), self.v0, (0.0, 0.0))
        return lb, body

    def evaluate(self, delta, stiffness, u_in):
        """(J, dJ/ddelta [M] float64, dJ/dK, dJ/du_in)."""
        lb, body = self._setup(delta, stiffness, u_in)
        wa, wb = lb.work()
        L.call("vsx_viv2d_adjoint", ctypes.byref(lb.p), ctypes.byref(body.ib), ctypes.byref(self.adj),
               L.ptr(self.f0), L.ptr(wa), L.ptr(wb), L.ptr(self.f_tmp))
        out = self.buf["out"].cpu().numpy()
        if int(self.buf["sync"][3].item()) != 0:
            raise RuntimeError("vsx_viv2d_adjoint: a wait between the reverse kernels timed out; the gradient is invalid")
        return float(out[0]), self.buf["grad_delta"].clone(), float(out[1]), float(out[2])

    def objective(self, delta, stiffness, u_in):
        """J alone, from a plain fused run (no reverse pass)."""
        lb, body = self._setup(delta, stiffness, u_in)
        hd = torch.empty((self.T, 2), dtype=torch.float64, device=L.device())
        body.ib.hist_d, body.ib.hist_h = hd.data_ptr(), None
        body.ib.fused = 2
        body.ib.step0 = 0
        wa, wb = lb.work()
        out = torch.empty_like(self.f0)
        L.call("vsx_viv2d_run", ctypes.byref(lb.p), ctypes.byref(body.ib), L.ptr(self.f0), L.ptr(out), L.ptr(wa),
               L.ptr(wb), self.T)
        return float((hd[:, 1] ** 2).mean().item())


class VIVObjective(torch.autograd.Function):
    """J(delta, K, u_in) of a ``VIVAdjoint`` as an autograd function:
    ``J = VIVObjective.apply(delta, K, u_in, adjoint)``."""

    @staticmethod
    def forward(ctx, delta, stiffness, u_in, adjoint):
        J, gd, gk, gu = adjoint.evaluate(delta.detach(), float(stiffness), float(u_in))
        ctx.save_for_backward(gd.to(delta.device, delta.dtype), torch.tensor(gk, dtype=torch.float64),
                              torch.tensor(gu, dtype=torch.float64))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        gd, gk, gu = ctx.saved_tensors
        return gJ * gd, (gJ * gk).to(gJ.dtype), (gJ * gu).to(gJ.dtype), None
