"""ctypes binding of ``libvivsim_amd.so`` (the C ABI declared in include/vivsim_amd.h).

The library is loaded lazily on first use.  There is no fallback: if the
library is missing, or no HIP device is visible when a kernel is called, the
call raises.
"""

import ctypes
import os
import re

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvivsim_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "vivsim_amd.h")

c_int32, c_int64, c_float, c_double, c_void_p = (ctypes.c_int32, ctypes.c_int64, ctypes.c_float,
                                                 ctypes.c_double, ctypes.c_void_p)

# ---------------------------------------------------------------- enums (mirror of the header)
OK, ERR_VALUE, ERR_KEY, ERR_HIP = 0, 1, 2, 3
LOC = {"left": 0, "right": 1, "bottom": 2, "top": 3, "back": 4, "front": 5}
WALL = {"nee": 0, "nebb": 1, "eq": 2, "bb": 3, "specular": 4}
MODE = {"plain": 0, "velocity": 1, "pressure": 2, "force_corrected": 3}
COLLISION = {"bgk": 0, "mrt": 1, "kbc": 2, "reg": 3}
FORCING = {None: 0, "none": 0, "guo": 1, "edm": 2}
KERNEL = {"peskin_2pt": 0, "peskin_3pt": 1, "peskin_4pt": 2, "cosine_4pt": 3}
STORE = {"f32": 0, "f16s": 1, "bf16s": 2, "e3m12s": 3}
PHASE = {"collide": 0, "full": 1, "stream": 2}


class VsxWall(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("mode", c_int32), ("loc", c_int32), ("reserved0", c_int32),
                ("rho", c_float), ("ux", c_float), ("uy", c_float), ("uz", c_float),
                ("gx", c_float), ("gy", c_float), ("gz", c_float), ("reserved1", c_float),
                ("rho_arr", c_void_p), ("ux_arr", c_void_p), ("uy_arr", c_void_p), ("uz_arr", c_void_p),
                ("gx_arr", c_void_p), ("gy_arr", c_void_p), ("gz_arr", c_void_p)]


class VsxLbm(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
                ("collision", c_int32), ("forcing", c_int32), ("storage", c_int32),
                ("omega", c_float), ("mrt_rates", c_float * 19), ("n_walls", c_int32),
                ("walls", VsxWall * 6), ("g", c_void_p), ("g_uniform", c_float * 3), ("store_u", c_float * 3),
                ("reserved", c_int32)]


class VsxIB2D(ctypes.Structure):
    _fields_ = [("n_markers", c_int32), ("kernel", c_int32), ("n_iter", c_int32), ("win", c_int32),
                ("marker_x", c_void_p), ("marker_y", c_void_p), ("ds_arr", c_void_p), ("ds", c_float),
                ("dof_mask", c_int32), ("win_x0", c_double), ("win_y0", c_double),
                ("mass", c_double), ("stiffness", c_double), ("damping", c_double), ("area", c_double),
                ("dt", c_double), ("gamma", c_double), ("beta", c_double),
                ("body", c_void_p), ("g_win", c_void_p), ("hist_d", c_void_p), ("hist_h", c_void_p),
                ("win_origin", c_void_p), ("step0", c_int32), ("fused", c_int32),
                ("reserved", c_int32), ("halo", c_void_p), ("status", c_void_p)]


class VsxVivCase(ctypes.Structure):
    _fields_ = [("lbm", VsxLbm), ("ib", VsxIB2D), ("f_offset", c_int64), ("tile_y", c_int32), ("hist_cap", c_int32)]


class VsxVivBatch(ctypes.Structure):
    _fields_ = [("n_cases", c_int32), ("collision", c_int32), ("storage", c_int32), ("lds_bytes", c_int32),
                ("n_blocks", c_int64), ("total_pops", c_int64), ("table", c_void_p), ("clock", c_void_p)]


class VsxAdj2D(ctypes.Structure):
    _fields_ = [("n_steps", c_int32), ("seg", c_int32), ("dir_x", c_void_p), ("dir_y", c_void_p),
                ("ckpt", c_void_p), ("ckpt_body", c_void_p), ("tape_f", c_void_p), ("tape_g", c_void_p),
                ("tape_body", c_void_p), ("adj_a", c_void_p), ("adj_b", c_void_p), ("grid", c_void_p),
                ("adj_body", c_void_p), ("hist_d", c_void_p), ("u_in_part", c_void_p), ("grad_delta", c_void_p),
                ("out", c_void_p), ("sync", c_void_p)]


class VsxIB3D(ctypes.Structure):
    _fields_ = [("n_markers", c_int32), ("n_iter", c_int32), ("origin", c_int32 * 3), ("win", c_int32 * 3),
                ("W", c_void_p), ("I", c_void_p), ("cell_ptr", c_void_p), ("cell_slot", c_void_p),
                ("u_target", c_void_p), ("ds_arr", c_void_p), ("ds", c_float), ("pipelined", c_int32),
                ("g_win", c_void_p), ("work", c_void_p), ("hist_h", c_void_p), ("halo", c_void_p)]


_CT = {"p": c_void_p, "i": c_int32, "l": c_int64, "f": c_float, "d": c_double}

# name -> argument codes (all return int except the two string getters)
_SIGS = {
    "vsx_device_sync": "",
    "vsx_timer_start": "i",
    "vsx_timer_stop": "pip",
    "vsx_velocity_correction": "ppfpil p",
    "vsx_ib_kernel": "ippl p",
    "vsx_ib_get_ib_stencil2d": "ppiiiipp p",
    "vsx_ib_get_ib_stencil3d": "piiiiiipp p",
    "vsx_ib_interpolate": "pilppiip p",
    "vsx_ib_spread": "pppilppii p",
    "vsx_ib_multi_direct_forcing": "pilppiipp fippp p",
    "vsx_ib_multi_direct_forcing_noalloc": "pilppiipp fipp p",
    "vsx_ib_get_area": "pip p",
    "vsx_ib_get_ds": "piip p",
    "vsx_ib3d_get_triangle_areas": "pipip p",
    "vsx_ib3d_get_surface_area": "pipip p",
    "vsx_ib3d_get_volume": "pipip p",
    "vsx_ib3d_get_ds": "pipip p",
    "vsx_dyn_newmark": "ppppipppidddppp p",
    "vsx_dyn_get_force_to_obj": "piip p",
    "vsx_dyn_get_torque_to_obj": "ppddppip p",
    "vsx_dyn_markers_coords": "ppiddpipp p",
    "vsx_dyn_markers_velocity_3dof": "ppiddppp p",
    "vsx_viv2d_run": "pppppp i p",
    "vsx_sphere3d_run": "pppppp i p",
    "vsx_ib3d_transpose_stencil": "piilppp p",
    "vsx_viv2d_batch_prepare": "pippp",
    "vsx_viv2d_batch_run": "ppppp i p",
    "vsx_viv2d_batch_graph": "ppppp i p",
    "vsx_graph_launch": "pp",
    "vsx_graph_destroy": "p",
    "vsx_viv2d_adjoint": "ppppppp p",
    "vsx_post_field": "piiiiip p",
    "vsx_post_mean": "piiiipp p",
    "vsx_post_pressure": "plfp p",
    "vsx_force_coefficients": "piddpp p",
    "vsx_mg_fine_to_coarse": "piipiii p",
    "vsx_mg_coarse_to_fine": "piipiii p",
}
for _pfx in ("d2q9", "d3q19"):
    _SIGS.update({
        f"vsx_{_pfx}_get_macroscopic": "pppl p",
        f"vsx_{_pfx}_get_equilibrium": "pppl p",
        f"vsx_{_pfx}_collision_bgk": "ppfpl p",
        f"vsx_{_pfx}_collision_mrt": "ppppl p",
        f"vsx_{_pfx}_collision_kbc": "ppfpl p",
        f"vsx_{_pfx}_collision_reg": "ppfpl p",
        f"vsx_{_pfx}_get_guo_forcing_term": "pppl p",
        f"vsx_{_pfx}_forcing_guo_bgk": "pppfpl p",
        f"vsx_{_pfx}_forcing_guo_mrt": "pppppl p",
        f"vsx_{_pfx}_forcing_edm": "ppppl p",
        f"vsx_{_pfx}_obstacle_bounce_back": "pppl p",
        f"vsx_{_pfx}_step": "ppipii p",
        f"vsx_{_pfx}_run": "pppppi p",
    })
_SIGS.update({
    "vsx_d2q9_streaming": "ppii p",
    "vsx_d3q19_streaming": "ppiii p",
    "vsx_d2q9_boundary": "ppppii p",
    "vsx_d3q19_boundary": "ppppiii p",
    "vsx_d2q9_boundary_characteristic": "ppippii p",
    "vsx_d3q19_boundary_characteristic": "ppippiii p",
})

# declared in _declare with their own return types
NON_STATUS = {"vsx_version", "vsx_last_error", "vsx_ib3d_work_floats", "vsx_viv2d_batch_table_bytes"}

_LIB = None


def _declare(L):
    for name, codes in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = c_int32
        fn.argtypes = [_CT[c] for c in codes.replace(" ", "")]
    # entry points that do not return a status code
    L.vsx_ib3d_work_floats.restype = c_int64
    L.vsx_ib3d_work_floats.argtypes = [c_int32, c_int64]
    L.vsx_viv2d_batch_table_bytes.restype = c_int64
    L.vsx_viv2d_batch_table_bytes.argtypes = [c_void_p, c_int32]
    for name in ("vsx_version", "vsx_last_error"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_char_p
        fn.argtypes = []


def lib():
    """The loaded library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m vivsim_amd._build`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _LIB = L
    return _LIB


def header_symbols():
    """Function names declared in include/vivsim_amd.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vsx_[a-z0-9_]+)\s*\(", txt)))


def check(rc):
    if rc != OK:
        msg = lib().vsx_last_error().decode()
        if rc == ERR_KEY:
            raise KeyError(msg)
        if rc == ERR_VALUE:
            raise ValueError(msg)
        raise RuntimeError(msg)


def call(name, *args):
    """Call a C entry point; the last argument (the HIP stream) is appended here."""
    check(getattr(lib(), name)(*args, stream()))


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("vivsim_amd needs a HIP device (none is visible)")
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    """The current HIP stream of the current device (raw handle, no Stream object)."""
    return c_void_p(torch._C._cuda_getCurrentRawStream(torch.cuda.current_device()))


def ptr(t):
    return None if t is None else c_void_p(t.data_ptr())
