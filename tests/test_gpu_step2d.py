"""GPU parity of the fused D2Q9 step (vivsim_amd.sim) against the oracle's
composition of the reference functions (oracle/steps.py).

The fused kernel evaluates the MRT operator in moment space and the walls in
a pull scheme, so results agree with the float64 composition to float32
rounding: tolerances below are relative L-inf errors of the population field."""

import numpy as np
import pytest
import torch

from oracle import lbm2d as O
from oracle import steps as S
from tests.gpu_util import assert_close, dev, rel_err, say, to_np

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(77)


@pytest.fixture(scope="module")
def sim(gpu_lib):
    from vivsim_amd import sim
    return sim


def rand_state(nx, ny, u0=0.04):
    rho = 1 + 0.02 * rng.standard_normal((nx, ny))
    u = u0 * rng.standard_normal((2, nx, ny))
    return (O.get_equilibrium(rho, u) + 1e-3 * rng.standard_normal((9, nx, ny))).astype(np.float32)


WALLSETS = {
    "none": [],
    "cavity_nee": S.cavity_walls(0.08),
    "cavity_nee_lid_first": S.cavity_walls(0.08, order=("top", "left", "right", "bottom")),
    "viv_like": [dict(kind="eq", mode="plain", loc="left", ux_wall=0.05),
                 dict(kind="nee", mode="pressure", loc="right", rho_wall=1.0)],
    "channel": [dict(kind="nebb", mode="velocity", loc="left", ux_wall=0.04),
                dict(kind="nebb", mode="pressure", loc="right", rho_wall=1.0),
                dict(kind="bb", loc="bottom"),
                dict(kind="specular", loc="top", ux_wall=0.01)],
    "mixed": [dict(kind="nebb", mode="force_corrected", loc="top", ux_wall=0.03, gx_wall=1e-4),
              dict(kind="eq", mode="velocity", loc="left", ux_wall=0.02),
              dict(kind="nee", mode="velocity", loc="bottom", ux_wall=0.01),
              dict(kind="eq", mode="pressure", loc="right", rho_wall=1.0)],
}


def run_oracle(f, n, omega, walls, collision, forcing, g):
    f = f.astype(np.float64)
    op = O.get_mrt_collision_operator(omega) if collision == "mrt" else None
    fop = O.get_mrt_forcing_operator(omega) if collision == "mrt" else None
    for _ in range(n):
        f = S.lbm_step2d(f, omega, walls, collision, forcing, g, op, fop)
    return f


@pytest.mark.parametrize("walls", list(WALLSETS))
@pytest.mark.parametrize("collision", ["bgk", "mrt", "kbc", "reg"])
def test_walls_and_collisions(sim, walls, collision):
    nx, ny, n, om = 26, 21, 12, 1.75
    f0 = rand_state(nx, ny)
    say(f"fused d2q9 {collision} walls={walls}")
    got = sim.LBM((nx, ny), om, collision=collision, walls=WALLSETS[walls]).run(dev(f0), n)
    ref = run_oracle(f0, n, om, WALLSETS[walls], collision, None, None)
    assert_close(got, ref, 2e-5, f"{collision}/{walls}")


@pytest.mark.parametrize("forcing", ["guo", "edm"])
@pytest.mark.parametrize("collision", ["bgk", "mrt", "kbc", "reg"])
def test_forcing(sim, collision, forcing):
    nx, ny, n, om = 30, 24, 10, 1.6
    f0 = rand_state(nx, ny)
    g = np.zeros((2, nx, ny), dtype=np.float32)
    g[:, 8:20, 6:18] = 2e-4 * rng.standard_normal((2, 12, 12))     # localised force (IB-like window)
    g += np.float32(1e-5)                                            # plus a uniform part
    walls = WALLSETS["viv_like"]
    say(f"fused d2q9 {collision}+{forcing}")
    got = sim.LBM((nx, ny), om, collision=collision, walls=walls, forcing=forcing, g=dev(g)).run(dev(f0), n)
    ref = run_oracle(f0, n, om, walls, collision, forcing, g.astype(np.float64))
    assert_close(got, ref, 2e-5, f"{collision}+{forcing}")


def test_uniform_force_equals_field(sim):
    nx, ny = 20, 18
    f0 = rand_state(nx, ny)
    walls = WALLSETS["channel"]
    a = sim.LBM((nx, ny), 1.5, collision="bgk", walls=walls, forcing="guo", g_uniform=(2e-5, 0.0)).run(dev(f0), 6)
    g = np.zeros((2, nx, ny), dtype=np.float32)
    g[0] = 2e-5
    b = sim.LBM((nx, ny), 1.5, collision="bgk", walls=walls, forcing="guo", g=dev(g)).run(dev(f0), 6)
    assert torch.equal(a, b)


def test_step_phases_match_run(sim):
    """collide -> full^(n-1) -> stream through the low-level entry equals run()."""
    nx, ny = 24, 20
    f0 = dev(rand_state(nx, ny))
    m = sim.LBM((nx, ny), 1.8, collision="mrt", walls=WALLSETS["cavity_nee"])
    a = m.run(f0, 4)
    w0, w1 = torch.empty_like(f0), torch.empty_like(f0)
    m.step(f0, w0, "collide", src_f32=True)
    m.step(w0, w1, "full")
    m.step(w1, w0, "full")
    m.step(w0, w1, "full")
    out = torch.empty_like(f0)
    m.step(w1, out, "stream", dst_f32=True)
    assert torch.equal(a, out)
    assert torch.equal(m.run(f0, 0), f0)


def test_cavity_1000_steps(sim):
    """Lid-driven cavity 256^2, BGK, NEE x4, U0 = 0.1, nu = 0.01, 1000 steps:
    float32 fused run vs float64 composition, relative L-inf <= 1e-5."""
    n, steps, U0 = 256, 1000, 0.1
    om = O.get_omega(0.01)
    walls = S.cavity_walls(U0)
    f0 = O.get_equilibrium(np.ones((n, n)), np.zeros((2, n, n)))
    say("cavity 256^2 x 1000 (oracle ~20 s)")
    got = sim.LBM((n, n), om, collision="bgk", walls=walls).run(dev(f0), steps)
    ref = run_oracle(f0.astype(np.float32), steps, om, walls, "bgk", None, None)
    e = assert_close(got, ref, 1e-5, "cavity f")
    r1, u1 = O.get_macroscopic(to_np(got).astype(np.float64))
    r2, u2 = O.get_macroscopic(ref)
    assert rel_err(u1, u2) < 1e-4
    say(f"cavity rel err f {e:.2e}")


def test_streamed_stores(sim):
    """A full step whose work buffer exceeds 256 MiB (the last-level cache)
    writes the field with non-temporal stores: 2760 x 2740 in FP32 (272 MB),
    KBC, the mixed wall set, 2 steps against the oracle."""
    nx, ny, n, om = 2760, 2740, 2, 1.7
    f0 = rand_state(nx, ny)
    walls = WALLSETS["mixed"]
    say("fused d2q9 streamed stores 2760x2740")
    got = sim.LBM((nx, ny), om, collision="kbc", walls=walls).run(dev(f0), n)
    ref = run_oracle(f0, n, om, walls, "kbc", None, None)
    assert_close(got, ref, 2e-5, "streamed stores 2d")


@pytest.mark.parametrize("storage,tol", [("f16s", 2e-3), ("bf16s", 2e-2), ("e3m12s", 1e-3)])
def test_half_storage(sim, storage, tol):
    nx, ny, n, om = 40, 32, 30, 1.7
    f0 = rand_state(nx, ny, u0=0.03)
    walls = WALLSETS["viv_like"]
    got = sim.LBM((nx, ny), om, collision="mrt", walls=walls, storage=storage).run(dev(f0), n)
    ref = run_oracle(f0, n, om, walls, "mrt", None, None)
    e = assert_close(got, ref, tol, storage)
    say(f"{storage}: rel err {e:.2e}")


def test_e3m12_codes_match_model(sim):
    """The device E3M12 conversions against tests/storage_model.py, bit for bit.
    Decode: a STREAM pass on a periodic grid writes decode(code) + s_i pulled
    along c_i (s_i = w_i here, read back from an all-zero code field).
    Encode: a BGK COLLIDE pass with omega = 1e-30 leaves f unchanged in float32
    and stores encode(f - s_i); the deviations span every block of the code
    and values beyond its range (saturation)."""
    from tests import storage_model as SM
    nx, ny = 48, 40
    lb = sim.LBM((nx, ny), 1e-30, collision="bgk", storage="e3m12s")
    wa, wb = lb.work()
    out = torch.empty((9, nx, ny), dtype=torch.float32, device="cuda")
    wa.zero_()
    lb.step(wa, out, "stream", dst_f32=True)
    s = to_np(out)[:, 0, 0].astype(np.float32)
    assert np.all(to_np(out) == s[:, None, None])

    codes = rng.integers(0, 0x10000, size=(9, nx, ny)).astype(np.uint16)
    wa.copy_(torch.from_numpy(codes.view(np.int16).reshape(-1)).to("cuda"))
    lb.step(wa, out, "stream", dst_f32=True)
    ref = O.streaming(SM.decode(codes) + s[:, None, None])
    assert ref.dtype == np.float32 and np.array_equal(to_np(out).view(np.uint32), ref.view(np.uint32))

    n = 9 * nx * ny
    x = np.concatenate([rng.uniform(-0.26, 0.26, n // 2),
                        np.sign(rng.standard_normal(n // 4)) * 2.0 ** rng.uniform(-30, -2, n // 4),
                        SM.magnitudes()[rng.integers(0, 0x7fff, n - n // 2 - n // 4)]])
    x = rng.permutation(x).reshape(9, nx, ny)
    x[0] = -x[1:].sum(0)   # density 1: the (unused) equilibrium stays finite
    f = (x + s[:, None, None]).astype(np.float32)
    lb.step(dev(f), wb, "collide", src_f32=True)
    want = SM.encode(f - s[:, None, None])
    got = to_np(wb).astype(np.int16).view(np.uint16).reshape(9, nx, ny)
    bad = np.count_nonzero(got != want)
    say(f"e3m12: {codes.size} decodes, {want.size} encodes, {bad} mismatches")
    assert bad == 0


# ------------------------------------------------------------------ VIV (IB-LBM)

def small_case(cls, **kw):
    base = dict(nx=160, ny=120, D=16.0, cx=48.0, cy=60.0, n_markers=64, U0=0.05, Re=60.0, Ur=5.0, mass_ratio=2.0,
                n_iter=3, pad=6, collision="mrt", forcing="guo", dof=(False, True))
    base.update(kw)
    return cls(**base)


@pytest.mark.parametrize("variant", [dict(), dict(collision="kbc", forcing="edm"), dict(collision="reg", forcing="edm"),
                                     dict(collision="bgk", forcing="guo", dof=(True, True), kernel="cosine_4pt"),
                                     dict(kernel="peskin_3pt", n_iter=1, inlet="fc_nebb", outlet="eq"),
                                     dict(cx=12.0, n_markers=48), dict(cx=150.0, cy=8.0), dict(n_markers=600)],
                         ids=["mrt_guo", "kbc_edm", "reg_edm", "bgk_guo_2dof_cos", "p3_nebb", "window_on_inlet",
                              "window_on_outlet_clamped", "dense_markers"])
@pytest.mark.parametrize("mode", ["pipelined", "one_launch", "separate"])
def test_viv_matches_oracle(sim, variant, mode):
    """Every launch form: the pipelined step (IB kernel of the next step on a
    second stream; windows 2+ cells inside the grid only), the one-launch step
    and the separate IB and bulk launches.  dense_markers: 600 markers, one
    marker per thread of the IB workgroup (more than half its threads)."""
    oc = small_case(S.VIV2D, **variant)
    pc = small_case(sim.VIVCase, **variant)
    if mode == "pipelined" and (pc.ib_x0 < 3 or pc.ib_y0 < 3 or pc.ib_x0 + pc.ib_size + 3 > pc.nx):
        pytest.skip("the pipelined step needs the window 2+ cells inside the grid")
    steps = 40
    f0 = oc.initial_f(np.float64)
    v0 = np.array([0.0, 1e-3])
    solver = sim.VIV2D(pc, fused=mode != "separate", pipelined=mode == "pipelined")
    solver.set_body(v=v0)
    say(f"viv {variant} {mode}")
    got, hd, hh = solver.run(dev(f0), steps)
    solver.check()
    f, d, v, a, dh, hhist = S.viv_run(oc, f0, steps, v=v0)
    assert_close(got, f, 5e-5, "viv f")
    assert_close(hh, hhist, 1e-4, "viv h history")
    assert_close(hd, dh, 1e-4, "viv d history")
    a_, v_, d_ = solver.state()
    assert np.allclose(d_, d, rtol=1e-4, atol=1e-12)


@pytest.mark.parametrize("storage,tol", [("f16s", 3e-3), ("bf16s", 3e-2), ("e3m12s", 1.5e-3)])
@pytest.mark.parametrize("variant", [dict(), dict(collision="bgk", nx=168, ny=100), dict(collision="reg", forcing="edm")],
                         ids=["mrt_guo", "bgk_odd_tiles", "reg_edm"])
def test_viv_16bit_matches_oracle(sim, variant, storage, tol):
    """The fused VIV step in 16-bit storage on small grids (one tile per bulk
    block); the two-tile blocks of large grids are covered below."""
    oc = small_case(S.VIV2D, **variant)
    pc = small_case(sim.VIVCase, **variant)
    steps = 40
    f0 = oc.initial_f(np.float64)
    v0 = np.array([0.0, 1e-3])
    solver = sim.VIV2D(pc, storage=storage)
    solver.set_body(v=v0)
    say(f"viv16 {variant} {storage}")
    got, hd, hh = solver.run(dev(f0), steps)
    solver.check()
    f, d, v, a, dh, hhist = S.viv_run(oc, f0, steps, v=v0)
    assert_close(got, f, tol, f"viv {storage} f")


@pytest.mark.parametrize("storage,tol", [("f16s", 3e-3), ("e3m12s", 1.5e-3)])
def test_viv_16bit_pair_blocks_match_oracle(sim, storage, tol):
    """Grids of at least 4096 tiles run the 16-bit bulk blocks on two tiles
    each: 2052 x 2000 (4104 tiles of 4 x 256; a partial last tile row, the wall
    columns and a lone last tile take the one-cell path), 3 steps."""
    kw = dict(nx=2052, ny=2000, cx=256.0, cy=1000.0)
    oc, pc = S.VIV2D(**kw), sim.VIVCase(**kw)
    steps = 3
    f0 = oc.initial_f(np.float64)
    v0 = np.array([0.0, 1e-3])
    solver = sim.VIV2D(pc, storage=storage)
    solver.set_body(v=v0)
    say(f"viv16 pair blocks {storage}")
    got, hd, hh = solver.run(dev(f0), steps)
    solver.check()
    f, d, v, a, dh, hhist = S.viv_run(oc, f0, steps, v=v0)
    assert_close(got, f, tol, f"viv {storage} f (pair blocks)")


@pytest.mark.parametrize("storage,tol,nx,ny", [("f32", 5e-5, 2760, 2740), ("f16s", 3e-3, 3880, 3872)])
def test_viv_streamed_stores_match_oracle(sim, storage, tol, nx, ny):
    """Pipelined full steps whose work buffer exceeds 256 MiB (the last-level
    cache) write the field with non-temporal stores: FP32 at 2760 x 2740
    (272 MB), FP16S at 3880 x 3872 (270 MB; tile pairs, a partial last tile
    row), 3 steps against the oracle."""
    kw = dict(nx=nx, ny=ny, cx=256.0, cy=ny / 2)
    oc, pc = S.VIV2D(**kw), sim.VIVCase(**kw)
    steps = 3
    f0 = oc.initial_f(np.float64)
    v0 = np.array([0.0, 1e-3])
    solver = sim.VIV2D(pc, storage=storage)
    assert solver.ib.fused == 2
    solver.set_body(v=v0)
    say(f"viv streamed stores {storage} {nx}x{ny}")
    got, hd, hh = solver.run(dev(f0), steps)
    solver.check()
    f, d, v, a, dh, hhist = S.viv_run(oc, f0, steps, v=v0)
    assert_close(got, f, tol, f"viv {storage} f (streamed stores)")
    if storage == "f32":
        assert_close(hh, hhist, 1e-4, "viv f32 h (streamed stores)")


def test_viv_split_runs(sim):
    """Runs of 1 + 14 + 25 steps continue the body state and the step counter
    exactly as one run of 40 steps (the oracle's 40-step run)."""
    oc, pc = small_case(S.VIV2D, dof=(True, True)), small_case(sim.VIVCase, dof=(True, True))
    f0 = oc.initial_f(np.float64)
    v0 = np.array([2e-4, 1e-3])
    solver = sim.VIV2D(pc)
    solver.set_body(v=v0)
    f = dev(f0)
    hs = []
    for n in (1, 14, 25):
        f, _, hh = solver.run(f, n)
        hs.append(to_np(hh))
    solver.check()
    fr, d, v, a, dh, hhist = S.viv_run(oc, f0, 40, v=v0)
    assert_close(f, fr, 5e-5, "split f")
    assert_close(np.concatenate(hs), hhist, 1e-4, "split h")
    a_, v_, d_ = solver.state()
    assert np.allclose(d_, d, rtol=1e-4, atol=1e-12)
    assert np.allclose(v_, v, rtol=1e-4, atol=1e-12)


def test_viv_zero_iterations_is_plain_step(sim):
    """With n_iter = 0 the IB adds nothing: the VIV run equals the plain fused run
    to float32 rounding (its window cells take the forced collision path,
    every other cell the unforced one)."""
    pc = small_case(sim.VIVCase, n_iter=0)
    f0 = dev(small_case(S.VIV2D).initial_f(np.float32))
    a, _, hh = sim.VIV2D(pc).run(f0, 10)
    b = sim.LBM((pc.nx, pc.ny), pc.omega, collision=pc.collision, walls=pc.walls(), forcing=pc.forcing).run(f0, 10)
    e = assert_close(a, to_np(b), 1e-6, "viv n_iter=0 vs plain")
    say(f"max rel diff {e:.2e}")
    assert torch.all(hh == 0)


def test_pipelined_window_jump_is_reported(sim):
    """A body moving two or more cells in a step moves the IB window out of the
    hand-off buffer of the pipelined step: the kernels report it (every access
    stays in bounds) through the pinned status word, ``check`` raises, the next
    run raises without a sync, and ``set_body`` clears the report."""
    pc = small_case(sim.VIVCase)
    f0 = dev(small_case(S.VIV2D).initial_f(np.float32))
    solver = sim.VIV2D(pc, pipelined=True)
    solver.set_body(v=(0.0, 2.5))
    say("pipelined window jump")
    f, _, _ = solver.run(f0, 4, history=False)
    with pytest.raises(RuntimeError, match="pipelined VIV step"):
        solver.check()
    with pytest.raises(RuntimeError, match="pipelined VIV step"):
        solver.run(f0, 1)
    solver.set_body(v=(0.0, 1e-3))
    solver.run(f0, 4)
    solver.check()
