"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, the ctypes struct mirrors match the C layout, and argument
validation maps to the reference's exception types (no GPU needed: every
checked path returns before any HIP call)."""

import ctypes
import os
import subprocess

import pytest

from vivsim_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    lib = L.lib()
    syms = L.header_symbols()
    assert len(syms) > 50
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.vsx_version().startswith(b"vivsim_amd")
    # every declared entry point has a ctypes signature
    assert set(syms) - set(L._SIGS) == L.NON_STATUS


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    fields = {"VsxWall": L.VsxWall, "VsxLbm": L.VsxLbm, "VsxIB2D": L.VsxIB2D, "VsxIB3D": L.VsxIB3D,
              "VsxVivCase": L.VsxVivCase, "VsxVivBatch": L.VsxVivBatch, "VsxAdj2D": L.VsxAdj2D}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vivsim_amd.h"', "int main(void) {"]
    for sname, cls in fields.items():
        lines.append(f'printf("{sname} %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname}.{fname} %zu\\n", offsetof({sname}, {fname}));')
    lines += ["return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for sname, cls in fields.items():
        assert int(got[sname]) == ctypes.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(got[f"{sname}.{fname}"]) == getattr(cls, fname).offset, f"{sname}.{fname}"


def _lbm(**kw):
    p = L.VsxLbm()
    p.nx, p.ny, p.nz = 16, 16, 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_validation_codes():
    lib = L.lib()
    # unknown delta kernel -> VALUE
    assert lib.vsx_ib_kernel(9, None, None, 10, None) == L.ERR_VALUE
    assert b"unknown kernel" in lib.vsx_last_error()
    # bad collision id -> VALUE
    p = _lbm(collision=7)
    assert lib.vsx_d2q9_run(ctypes.byref(p), None, None, None, None, 1, None) == L.ERR_VALUE
    # unknown wall location in 2D -> KEY (the reference raises KeyError)
    p = _lbm(n_walls=1)
    p.walls[0].loc = 4
    assert lib.vsx_d2q9_run(ctypes.byref(p), None, None, None, None, 1, None) == L.ERR_KEY
    w = L.VsxWall()
    w.loc = 9
    assert lib.vsx_d2q9_boundary(ctypes.byref(w), None, None, None, 8, 8, None) == L.ERR_KEY
    # stencil radius out of range
    assert lib.vsx_ib_get_ib_stencil2d(None, None, 4, 8, 2, 7, None, None, None) == L.ERR_VALUE
    # grid too small / 3D nz mismatch
    p = _lbm()
    p.nx = 2
    assert lib.vsx_d2q9_step(ctypes.byref(p), None, 0, None, 0, 1, None) == L.ERR_VALUE
    p = _lbm()
    p.nz = 1
    assert lib.vsx_d3q19_run(ctypes.byref(p), None, None, None, None, 1, None) == L.ERR_VALUE


def test_python_exception_mapping():
    with pytest.raises(ValueError):
        L.check(L.ERR_VALUE)
    with pytest.raises(KeyError):
        L.check(L.ERR_KEY)
    with pytest.raises(RuntimeError):
        L.check(L.ERR_HIP)


def test_product_has_no_oracle_dependency():
    """The product package never imports the oracle (the oracle is test-only)."""
    pkg = os.path.join(ROOT, "vivsim_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".cpp")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in txt.replace("the oracle", ""), os.path.join(dirpath, fn)


def test_api_surface_matches_reference_names():
    """Every public name of the reference modules (lbm/__init__.py:1-42,
    lbm3d/__init__.py:1-44, ib/__init__.py:1-6, ib3d/__init__.py:1-12, dyn.py)
    exists in the drop-in modules."""
    import vivsim_amd as V
    lbm_names = ["streaming", "collision_bgk", "get_equilibrium", "get_macroscopic", "get_velocity_correction",
                 "get_omega", "collision_kbc", "get_mrt_collision_operator", "collision_mrt", "collision_reg",
                 "forcing_edm", "get_guo_forcing_term", "forcing_guo_bgk", "get_mrt_forcing_operator",
                 "forcing_guo_mrt", "boundary_bounce_back", "boundary_specular_reflection", "obstacle_bounce_back",
                 "boundary_characteristic", "boundary_equilibrium", "boundary_force_corrected_equilibrium",
                 "boundary_pressure_equilibrium", "boundary_velocity_equilibrium", "boundary_force_corrected_nebb",
                 "boundary_nebb", "boundary_pressure_nebb", "boundary_velocity_nebb", "boundary_nee",
                 "boundary_force_corrected_nee", "boundary_pressure_nee", "boundary_velocity_nee"]
    for mod in (V.lbm, V.lbm3d):
        for n in lbm_names:
            assert callable(getattr(mod, n)), (mod.__name__, n)
    for n in ["get_area", "get_ds", "kernel_cosine_4pt", "kernel_peskin_3pt", "kernel_peskin_4pt",
              "multi_direct_forcing", "get_ib_stencil", "interpolate", "spread"]:
        assert callable(getattr(V.ib, n))
    for n in ["get_ds", "get_surface_area", "get_triangle_areas", "get_volume", "get_ib_stencil", "kernel_cosine_4pt",
              "kernel_peskin_3pt", "kernel_peskin_4pt", "multi_direct_forcing", "interpolate", "spread"]:
        assert callable(getattr(V.ib3d, n))
    for n in ["newmark", "newmark_2dof", "newmark_3dof", "get_markers_coords_2dof", "get_markers_coords_3dof",
              "get_markers_velocity_3dof", "get_force_to_obj", "get_torque_to_obj"]:
        assert callable(getattr(V.dyn, n))


def test_signatures_match_reference():
    """Every public function of the drop-in modules has the reference's
    parameter names, order, kinds and defaults.  The reference signatures are
    a committed fixture (tests/golden/reference_signatures.json), read from
    the reference source text by tests/golden/make_signatures.py."""
    import inspect
    import json

    import vivsim_amd as V
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_signatures.json")))
    kinds = {"positional": (inspect.Parameter.POSITIONAL_OR_KEYWORD, inspect.Parameter.POSITIONAL_ONLY),
             "var_positional": (inspect.Parameter.VAR_POSITIONAL,),
             "keyword_only": (inspect.Parameter.KEYWORD_ONLY,),
             "var_keyword": (inspect.Parameter.VAR_KEYWORD,)}
    checked = 0
    for modname, funcs in ref.items():
        mod = getattr(V, modname)
        for fname, params in funcs.items():
            ours = inspect.signature(getattr(mod, fname))
            got = list(ours.parameters.values())
            assert [p.name for p in got] == [p["name"] for p in params], (modname, fname, str(ours))
            for p, r in zip(got, params):
                assert p.kind in kinds[r["kind"]], (modname, fname, p.name, p.kind)
                d = r["default"]
                if d is None:
                    assert p.default is inspect.Parameter.empty, (modname, fname, p.name)
                elif "name" in d:
                    assert getattr(p.default, "__name__", None) == d["name"], (modname, fname, p.name, p.default)
                else:
                    assert p.default == d["value"] and type(p.default) is type(d["value"]), \
                        (modname, fname, p.name, p.default, d["value"])
            checked += 1
    assert checked == sum(len(v) for v in ref.values()) >= 90
