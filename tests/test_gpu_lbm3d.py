"""GPU parity of the D3Q19 operators (vivsim_amd.lbm3d) and the fused D3Q19
step against the oracle (float64 on the same float32 inputs)."""

import numpy as np
import torch
import pytest

from oracle import lbm3d as O
from oracle import steps as S
from tests.gpu_util import assert_close, dev, say, to_np

pytestmark = pytest.mark.gpu
TOL = 3e-6
rng = np.random.default_rng(3003)


def field(shape=(9, 8, 7)):
    rho = 1 + 0.05 * rng.standard_normal(shape)
    u = 0.05 * rng.standard_normal((3,) + shape)
    return (O.get_equilibrium(rho, u) + 2e-3 * rng.standard_normal((19,) + shape)).astype(np.float32)


@pytest.fixture(scope="module")
def L(gpu_lib):
    from vivsim_amd import lbm3d
    return lbm3d


def test_streaming_exact(L):
    f = field()
    assert np.array_equal(to_np(L.streaming(dev(f))), O.streaming(f))


def test_macro_eq(L):
    f = field()
    rho, u = L.get_macroscopic(dev(f))
    r64, u64 = O.get_macroscopic(f.astype(np.float64))
    assert_close(rho, r64, TOL, "rho")
    assert_close(u, u64, TOL, "u")
    feq = L.get_equilibrium(rho, u)
    assert_close(feq, O.get_equilibrium(to_np(rho).astype(np.float64), to_np(u).astype(np.float64)), TOL, "feq")


@pytest.mark.parametrize("name", ["bgk", "kbc", "reg", "mrt"])
def test_collisions(L, name):
    f = field()
    f64 = f.astype(np.float64)
    r, u = O.get_macroscopic(f64)
    feq = O.get_equilibrium(r, u).astype(np.float32)
    om = 1.85
    if name == "mrt":
        op = L.get_mrt_collision_operator(om)
        assert np.allclose(to_np(op), O.get_mrt_collision_operator(om), atol=2e-5)
        got = L.collision_mrt(dev(f), dev(feq), op)
        ref = O.collision_mrt(f64, feq.astype(np.float64), to_np(op).astype(np.float64))
    else:
        got = getattr(L, "collision_" + name)(dev(f), dev(feq), om)
        ref = getattr(O, "collision_" + name)(f64, feq.astype(np.float64), om)
    assert_close(got, ref, 1e-5 if name == "kbc" else TOL, name)


def test_forcing(L):
    f = field()
    sh = f.shape[1:]
    u = (0.05 * rng.standard_normal((3,) + sh)).astype(np.float32)
    g = (1e-3 * rng.standard_normal((3,) + sh)).astype(np.float32)
    f64, u64, g64 = f.astype(np.float64), u.astype(np.float64), g.astype(np.float64)
    assert_close(L.get_guo_forcing_term(dev(g), dev(u)), O.get_guo_forcing_term(g64, u64), TOL, "guo term")
    assert_close(L.forcing_guo_bgk(dev(f), dev(g), dev(u), 1.6), O.forcing_guo_bgk(f64, g64, u64, 1.6), TOL, "guo bgk")
    assert_close(L.forcing_edm(dev(f), dev(g), dev(u)), O.forcing_edm(f64, g64, u64), TOL, "edm")
    op = L.get_mrt_forcing_operator(1.6)
    assert_close(L.forcing_guo_mrt(dev(f), dev(g), dev(u), op),
                 O.forcing_guo_mrt(f64, g64, u64, to_np(op).astype(np.float64)), TOL, "guo mrt")


WALLS = [("boundary_nee", dict(rho_wall=1.01, ux_wall=0.03, uy_wall=-0.02, uz_wall=0.01)),
         ("boundary_velocity_nee", dict(ux_wall=0.03, uz_wall=0.01)),
         ("boundary_pressure_nee", dict(rho_wall=0.99)),
         ("boundary_force_corrected_nee", dict(ux_wall=0.02, gx_wall=1e-3, gz_wall=-2e-3)),
         ("boundary_nebb", dict(rho_wall=1.02, ux_wall=0.01, uy_wall=0.02)),
         ("boundary_velocity_nebb", dict(ux_wall=0.05, uy_wall=-0.01, uz_wall=0.02)),
         ("boundary_pressure_nebb", dict(rho_wall=1.0)),
         ("boundary_equilibrium", dict(rho_wall=1.0, ux_wall=0.05)),
         ("boundary_pressure_equilibrium", dict(rho_wall=1.01)),
         ("boundary_force_corrected_equilibrium", dict(ux_wall=0.05, gy_wall=1e-3))]


@pytest.mark.parametrize("loc", ["left", "right", "bottom", "top", "back", "front"])
@pytest.mark.parametrize("name,kw", WALLS, ids=[w[0] for w in WALLS])
def test_walls(L, name, kw, loc):
    fs = O.streaming(field())
    assert_close(getattr(L, name)(dev(fs), loc, **kw), getattr(O, name)(fs.astype(np.float64), loc, **kw), 5e-6,
                 f"{name} {loc}")


@pytest.mark.parametrize("loc", ["left", "right", "bottom", "top", "back", "front"])
def test_bounce_back_and_cbc(L, loc):
    pre = field()
    fs = O.streaming(pre)
    for name in ("boundary_bounce_back", "boundary_specular_reflection"):
        got = getattr(L, name)(dev(pre), dev(fs), loc, ux_wall=0.02, uz_wall=-0.01)
        ref = getattr(O, name)(pre.astype(np.float64), fs.astype(np.float64), loc, ux_wall=0.02, uz_wall=-0.01)
        assert_close(got, ref, TOL, name)
    rho = (1 + 0.01 * rng.standard_normal((7, 6, 5))).astype(np.float32)
    u = (0.03 * rng.standard_normal((3, 7, 6, 5))).astype(np.float32)
    r, uu = L.boundary_characteristic(dev(rho), dev(u), loc)
    r0, u0 = O.boundary_characteristic(rho.astype(np.float64), u.astype(np.float64), loc)
    assert_close(r, r0, 1e-5, "rho_next")
    assert_close(uu, u0, 1e-5, "u_next")


def test_obstacle(L):
    f = field()
    mask = rng.random(f.shape[1:]) < 0.25
    assert np.array_equal(to_np(L.obstacle_bounce_back(dev(f), mask)), O.obstacle_bounce_back(f, mask))


# ------------------------------------------------------------------ fused D3Q19 step

@pytest.fixture(scope="module")
def sim(gpu_lib):
    from vivsim_amd import sim
    return sim


WALLSETS3 = {
    "none": [],
    "inlet_outlet": [dict(kind="eq", mode="plain", loc="left", ux_wall=0.05),
                     dict(kind="nee", mode="pressure", loc="right", rho_wall=1.0)],
    "sphere_example": [dict(kind="nebb", mode="plain", loc="left", ux_wall=0.05),
                       dict(kind="eq", mode="plain", loc="right", ux_wall=0.05)],
    "box": [dict(kind="nebb", mode="velocity", loc="left", ux_wall=0.03),
            dict(kind="nee", mode="pressure", loc="right", rho_wall=1.0),
            dict(kind="bb", loc="bottom"), dict(kind="specular", loc="top"),
            dict(kind="nee", mode="plain", loc="back", ux_wall=0.01), dict(kind="eq", mode="velocity", loc="front")],
}


def run_oracle(f, n, omega, walls, collision, forcing, g):
    f = f.astype(np.float64)
    op = O.get_mrt_collision_operator(omega) if collision == "mrt" else None
    fop = O.get_mrt_forcing_operator(omega) if collision == "mrt" else None
    for _ in range(n):
        f = S.lbm_step3d(f, omega, walls, collision, forcing, g, op, fop)
    return f


@pytest.mark.parametrize("walls", list(WALLSETS3))
@pytest.mark.parametrize("collision", ["bgk", "mrt", "kbc", "reg"])
def test_fused3d(sim, walls, collision):
    sh, n, om = (12, 10, 9), 6, 1.8
    f0 = field(sh)
    say(f"fused d3q19 {collision} {walls}")
    got = sim.LBM(sh, om, collision=collision, walls=WALLSETS3[walls]).run(dev(f0), n)
    ref = run_oracle(f0, n, om, WALLSETS3[walls], collision, None, None)
    assert_close(got, ref, 3e-5, f"{collision}/{walls}")


@pytest.mark.parametrize("forcing", ["guo", "edm"])
@pytest.mark.parametrize("collision", ["bgk", "mrt"])
def test_fused3d_forcing(sim, collision, forcing):
    sh, n, om = (12, 10, 9), 6, 1.7
    f0 = field(sh)
    g = np.zeros((3,) + sh, dtype=np.float32)
    g[:, 3:8, 2:7, 2:6] = 2e-4 * rng.standard_normal((3, 5, 5, 4))
    walls = WALLSETS3["inlet_outlet"]
    got = sim.LBM(sh, om, collision=collision, walls=walls, forcing=forcing, g=dev(g)).run(dev(f0), n)
    ref = run_oracle(f0, n, om, walls, collision, forcing, g.astype(np.float64))
    assert_close(got, ref, 3e-5, f"{collision}+{forcing}")


@pytest.mark.parametrize("storage,tol", [("f16s", 3e-3), ("bf16s", 3e-2), ("e3m12s", 1.5e-3)])
def test_fused3d_half(sim, storage, tol):
    sh, n, om = (16, 12, 10), 12, 1.8
    f0 = field(sh)
    walls = WALLSETS3["inlet_outlet"]
    got = sim.LBM(sh, om, collision="bgk", walls=walls, storage=storage).run(dev(f0), n)
    ref = run_oracle(f0, n, om, walls, "bgk", None, None)
    assert_close(got, ref, tol, storage)


def test_fused3d_streamed_stores(sim):
    """A full step whose work buffer exceeds 256 MiB (the last-level cache)
    writes the field with non-temporal stores: 160 x 160 x 140 in FP32
    (272 MB), MRT, the six-wall box, 2 steps against the oracle."""
    sh, n, om = (160, 160, 140), 2, 1.8
    f0 = field(sh)
    walls = WALLSETS3["box"]
    say("fused d3q19 streamed stores 160x160x140")
    got = sim.LBM(sh, om, collision="mrt", walls=walls).run(dev(f0), n)
    ref = run_oracle(f0, n, om, walls, "mrt", None, None)
    assert_close(got, ref, 3e-5, "streamed stores 3d")


# ------------------------------------------------------------------ sphere IB-LBM (3D)

def small_sphere(cls, **kw):
    base = dict(nx=40, ny=28, nz=26, D=8.0, cx=14.0, cy=14.0, cz=13.0, n_markers=256, U0=0.05, Re=40.0, n_iter=2,
                pad=3)
    base.update(kw)
    return cls(**base)


@pytest.mark.parametrize("variant", [dict(), dict(n_iter=3, collision="mrt", forcing="guo"),
                                     dict(n_iter=1, inlet="nebb", outlet="eq"), dict(n_iter=0)],
                         ids=["bgk_edm_2it", "mrt_guo_3it", "nebb_1it", "no_ib"])
@pytest.mark.parametrize("pipelined", [False, True], ids=["sequential", "pipelined"])
def test_sphere_matches_oracle(sim, variant, pipelined):
    """Both launch forms: IB kernels then the bulk pass on the caller's stream,
    or the IB of step t + 1 on a second stream beside the bulk pass of step t."""
    oc, pc = small_sphere(S.Sphere3D, **variant), small_sphere(sim.SphereCase, **variant)
    steps = 12
    say(f"sphere {variant} pipelined={pipelined}")
    solver = sim.Sphere3D(pc, pipelined=pipelined)
    f0 = oc.initial_f(np.float64)
    got, hh = solver.run(dev(f0), steps)
    W, I = oc.stencil(np.float32)
    assert np.array_equal(to_np(solver.I), I), "stencil indices"
    f = f0.astype(np.float64)
    hs = []
    for _ in range(steps):
        f, h = S.sphere_step(oc, f, W.astype(np.float64), I)
        hs.append(h)
    assert_close(got, f, 5e-5, "sphere f")
    if oc.n_iter > 0:
        assert_close(hh, np.array(hs), 1e-4, "sphere force history")
    else:
        assert torch.all(hh == 0)


@pytest.mark.parametrize("storage,tol", [("f16s", 3e-3), ("e3m12s", 1.5e-3)])
def test_sphere_16bit_matches_oracle(sim, storage, tol):
    """The sphere step in 16-bit shifted storage (f - feq(1, (U0, 0, 0)) kept)."""
    oc, pc = small_sphere(S.Sphere3D), small_sphere(sim.SphereCase)
    steps = 12
    say(f"sphere {storage}")
    solver = sim.Sphere3D(pc, storage=storage)
    f0 = oc.initial_f(np.float64)
    got, hh = solver.run(dev(f0), steps)
    W, I = oc.stencil(np.float32)
    f = f0.astype(np.float64)
    for _ in range(steps):
        f, h = S.sphere_step(oc, f, W.astype(np.float64), I)
    assert_close(got, f, tol, f"sphere f ({storage})")


def test_sphere_transpose_is_exact(sim):
    """The transposed stencil lists every in-range slot once, per cell in ascending order."""
    pc = small_sphere(sim.SphereCase)
    solver = sim.Sphere3D(pc)
    I = to_np(solver.I).ravel()
    cells = pc.ib_size ** 3
    ptr, slot = to_np(solver.cell_ptr), to_np(solver.cell_slot)
    valid = np.nonzero((I >= 0) & (I < cells))[0]
    assert ptr[0] == 0 and ptr[-1] == valid.size
    for k in range(cells):
        lst = slot[ptr[k]:ptr[k + 1]]
        assert np.all(np.diff(lst) > 0)
        assert np.all(I[lst] == k)
    assert np.array_equal(np.sort(slot[:ptr[-1]]), valid)
