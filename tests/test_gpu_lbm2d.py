"""GPU parity of the D2Q9 operators (vivsim_amd.lbm) against the oracle.

Inputs are float32; the oracle runs in float64 on the same float32 values.
Tolerance for single operators: 2e-6 relative L-inf (a few float32 ulps);
streaming, obstacle bounce-back and index work are bit-exact."""

import numpy as np
import pytest
import torch

from oracle import lbm2d as O
from tests.gpu_util import assert_close, dev, say, to_np

pytestmark = pytest.mark.gpu

TOL = 2e-6
rng = np.random.default_rng(2024)


def field(nx=37, ny=29):
    rho = 1 + 0.05 * rng.standard_normal((nx, ny))
    u = 0.06 * rng.standard_normal((2, nx, ny))
    f = O.get_equilibrium(rho, u) + 2e-3 * rng.standard_normal((9, nx, ny))
    return f.astype(np.float32)


@pytest.fixture(scope="module")
def L(gpu_lib):
    from vivsim_amd import lbm
    return lbm


def test_streaming_exact(L):
    f = field()
    say("d2q9 streaming")
    assert np.array_equal(to_np(L.streaming(dev(f))), O.streaming(f))


def test_macro_eq(L):
    f = field()
    rho, u = L.get_macroscopic(dev(f))
    r64, u64 = O.get_macroscopic(f.astype(np.float64))
    assert_close(rho, r64, TOL, "rho")
    assert_close(u, u64, TOL, "u")
    # wall lines (9, N) as in the NEE helper
    r1, u1 = L.get_macroscopic(dev(f[:, 0, :]))
    assert_close(u1, O.get_macroscopic(f[:, 0, :].astype(np.float64))[1], TOL, "line u")
    feq = L.get_equilibrium(rho, u)
    assert_close(feq, O.get_equilibrium(to_np(rho).astype(np.float64), to_np(u).astype(np.float64)), TOL, "feq")


@pytest.mark.parametrize("name", ["bgk", "kbc", "reg", "mrt"])
def test_collisions(L, name):
    f = field()
    f64 = f.astype(np.float64)
    r, u = O.get_macroscopic(f64)
    feq = O.get_equilibrium(r, u).astype(np.float32)
    om = 1.7
    if name == "mrt":
        op = L.get_mrt_collision_operator(om)
        assert np.allclose(to_np(op), O.get_mrt_collision_operator(om), atol=1e-6)
        got = L.collision_mrt(dev(f), dev(feq), op)
        ref = O.collision_mrt(f64, feq.astype(np.float64), to_np(op).astype(np.float64))
    else:
        got = getattr(L, "collision_" + name)(dev(f), dev(feq), om)
        ref = getattr(O, "collision_" + name)(f64, feq.astype(np.float64), om)
    assert_close(got, ref, 4e-6 if name == "kbc" else TOL, name)


def test_forcing(L):
    f = field()
    u = (0.05 * rng.standard_normal((2,) + f.shape[1:])).astype(np.float32)
    g = (1e-3 * rng.standard_normal((2,) + f.shape[1:])).astype(np.float32)
    f64, u64, g64 = f.astype(np.float64), u.astype(np.float64), g.astype(np.float64)
    assert_close(L.get_guo_forcing_term(dev(g), dev(u)), O.get_guo_forcing_term(g64, u64), TOL, "guo term")
    assert_close(L.forcing_guo_bgk(dev(f), dev(g), dev(u), 1.6), O.forcing_guo_bgk(f64, g64, u64, 1.6), TOL, "guo bgk")
    assert_close(L.forcing_edm(dev(f), dev(g), dev(u)), O.forcing_edm(f64, g64, u64), TOL, "edm")
    op = L.get_mrt_forcing_operator(1.6)
    assert_close(L.forcing_guo_mrt(dev(f), dev(g), dev(u), op),
                 O.forcing_guo_mrt(f64, g64, u64, to_np(op).astype(np.float64)), TOL, "guo mrt")
    # velocity correction, scalar and array density
    rho = (1 + 0.1 * rng.random(f.shape[1:])).astype(np.float32)
    assert_close(L.get_velocity_correction(dev(g), dev(rho)), g64 * 0.5 / rho, TOL, "vcorr")
    assert abs(L.get_velocity_correction(2e-3, 2.0) - 5e-4) < 1e-12


WALL_CASES = [
    ("boundary_nee", dict(rho_wall=1.01, ux_wall=0.03, uy_wall=-0.02)),
    ("boundary_velocity_nee", dict(ux_wall=0.03, uy_wall=0.01)),
    ("boundary_pressure_nee", dict(rho_wall=0.99)),
    ("boundary_force_corrected_nee", dict(rho_wall=1.0, ux_wall=0.02, gx_wall=1e-3, gy_wall=-2e-3)),
    ("boundary_nebb", dict(rho_wall=1.02, ux_wall=0.01, uy_wall=0.02)),
    ("boundary_velocity_nebb", dict(ux_wall=0.05, uy_wall=-0.01)),
    ("boundary_pressure_nebb", dict(rho_wall=1.0)),
    ("boundary_force_corrected_nebb", dict(ux_wall=0.05, gx_wall=1e-4)),
    ("boundary_equilibrium", dict(rho_wall=1.0, ux_wall=0.05)),
    ("boundary_velocity_equilibrium", dict(ux_wall=0.02)),
    ("boundary_pressure_equilibrium", dict(rho_wall=1.01)),
    ("boundary_force_corrected_equilibrium", dict(ux_wall=0.05, gy_wall=1e-3)),
]


@pytest.mark.parametrize("loc", ["left", "right", "top", "bottom"])
@pytest.mark.parametrize("name,kw", WALL_CASES, ids=[c[0] for c in WALL_CASES])
def test_walls(L, name, kw, loc):
    fs = O.streaming(field())
    got = getattr(L, name)(dev(fs), loc, **kw)
    ref = getattr(O, name)(fs.astype(np.float64), loc, **kw)
    assert_close(got, ref, 4e-6, f"{name} {loc}")


@pytest.mark.parametrize("loc", ["left", "right", "top", "bottom"])
def test_wall_arrays(L, loc):
    """Per-node wall values (arrays of the wall's length)."""
    fs = O.streaming(field())
    n = fs.shape[2] if loc in ("left", "right") else fs.shape[1]
    ux = (0.05 * rng.random(n)).astype(np.float32)
    rw = (1 + 0.01 * rng.random(n)).astype(np.float32)
    got = L.boundary_nee(dev(fs), loc, rho_wall=dev(rw), ux_wall=dev(ux), uy_wall=0.01)
    ref = O.boundary_nee(fs.astype(np.float64), loc, rho_wall=rw.astype(np.float64), ux_wall=ux.astype(np.float64),
                         uy_wall=0.01)
    assert_close(got, ref, 4e-6, "nee arrays")


@pytest.mark.parametrize("loc", ["left", "right", "top", "bottom"])
def test_bounce_back(L, loc):
    pre = field()
    fs = O.streaming(pre)
    for name in ("boundary_bounce_back", "boundary_specular_reflection"):
        got = getattr(L, name)(dev(pre), dev(fs), loc, ux_wall=0.02, uy_wall=-0.01)
        ref = getattr(O, name)(pre.astype(np.float64), fs.astype(np.float64), loc, ux_wall=0.02, uy_wall=-0.01)
        assert_close(got, ref, TOL, name)


def test_obstacle_and_errors(L):
    f = field()
    mask = rng.random(f.shape[1:]) < 0.2
    assert np.array_equal(to_np(L.obstacle_bounce_back(dev(f), torch.as_tensor(mask).cuda())),
                          O.obstacle_bounce_back(f, mask))
    with pytest.raises(KeyError):
        L.boundary_nee(dev(f), "front")
    with pytest.raises(ValueError):
        L.boundary_characteristic(dev(f[0]), dev(f[1:3]), "front")


@pytest.mark.parametrize("loc", ["left", "right", "top", "bottom"])
def test_characteristic(L, loc):
    rho = (1 + 0.01 * rng.standard_normal((12, 9))).astype(np.float32)
    u = (0.03 * rng.standard_normal((2, 12, 9))).astype(np.float32)
    r, uu = L.boundary_characteristic(dev(rho), dev(u), loc)
    r0, u0 = O.boundary_characteristic(rho.astype(np.float64), u.astype(np.float64), loc)
    assert_close(r, r0, 1e-5, "rho_next")
    assert_close(uu, u0, 1e-5, "u_next")
