"""Loader of the C restatement (oracle/c/viv_cpu.c, OpenMP) of the reference's
D2Q9 cavity and VIV steps: the CPU baseline timed beside the GPU and a
host-side cross-check of the NumPy oracle.  Test infrastructure only.

``build()`` compiles it with gcc into oracle/_c/libvivcpu.so."""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "viv_cpu.c")
LIB = os.path.join(HERE, "_c", "libvivcpu.so")
_LIB = None


def build():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c11", SRC, "-o", LIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


class VcViv(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("n_markers", ctypes.c_int), ("win", ctypes.c_int),
                ("n_iter", ctypes.c_int), ("dof_mask", ctypes.c_int), ("u_in", ctypes.c_float), ("ds", ctypes.c_float),
                ("win_x0", ctypes.c_double), ("win_y0", ctypes.c_double), ("mass", ctypes.c_double),
                ("stiffness", ctypes.c_double), ("damping", ctypes.c_double), ("area", ctypes.c_double),
                ("op", ctypes.c_void_p), ("opf", ctypes.c_void_p), ("marker_x", ctypes.c_void_p),
                ("marker_y", ctypes.c_void_p)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.vc_cavity_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_int]
        L.vc_viv_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p]
        L.vc_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def cavity_run(f, omega, u_lid, n_steps):
    """n_steps of the cavity (BGK, NEE on left, right, bottom, top) on a float32 copy of f."""
    f = np.ascontiguousarray(f, dtype=np.float32).copy()
    tmp = np.empty_like(f)
    rc = lib().vc_cavity_run(_p(f), _p(tmp), f.shape[1], f.shape[2], float(omega), float(u_lid), int(n_steps))
    assert rc == 0
    return f


class VivRunner:
    """The VIV case of oracle.steps.VIV2D (MRT + Guo, eq inlet, pressure-NEE outlet) in C."""

    def __init__(self, case):
        from . import lbm2d as L2
        assert case.collision == "mrt" and case.forcing == "guo" and case.kernel == "peskin_4pt"
        assert case.inlet == "eq" and case.outlet == "pressure_nee"
        self.case = case
        self.op = np.ascontiguousarray(L2.get_mrt_collision_operator(case.omega, np.float32))
        self.opf = np.ascontiguousarray(L2.get_mrt_forcing_operator(case.omega, np.float32))
        mx, my = case.markers()
        self.mx, self.my = np.ascontiguousarray(mx, np.float64), np.ascontiguousarray(my, np.float64)
        c = VcViv()
        c.nx, c.ny, c.n_markers, c.win, c.n_iter = case.nx, case.ny, case.n_markers, case.ib_size, case.n_iter
        c.dof_mask = int(bool(case.dof[0])) | (int(bool(case.dof[1])) << 1)
        c.u_in, c.ds = case.U0, case.ds
        c.win_x0, c.win_y0 = float(case.ib_x0), float(case.ib_y0)
        c.mass, c.stiffness, c.damping, c.area = case.mass, case.stiffness, case.damping, case.area
        c.op, c.opf, c.marker_x, c.marker_y = _p(self.op), _p(self.opf), _p(self.mx), _p(self.my)
        self.c = c
        self.scratch = np.empty(11 * case.nx * case.ny, np.float32)

    def run(self, f, n_steps, body):
        """Advances the float32 field f (in place) and body [a0 a1 v0 v1 d0 d1] (float64, in place);
        returns the (d, h) histories [n, 2]."""
        assert f.dtype == np.float32 and f.flags.c_contiguous and body.dtype == np.float64
        hd = np.empty((n_steps, 2))
        hh = np.empty((n_steps, 2))
        rc = lib().vc_viv_run(ctypes.byref(self.c), _p(f), _p(self.scratch), _p(body), int(n_steps), _p(hd), _p(hh))
        assert rc == 0
        return hd, hh


def threads():
    return int(lib().vc_threads())
