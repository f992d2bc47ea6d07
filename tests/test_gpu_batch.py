"""GPU parity of the batched VIV sweep (vsx_viv2d_batch_*): every case of a
batch must reproduce the oracle's VIV run of that case, also when the run is
captured in a HIP graph and replayed, and with autotuned tile widths."""

import numpy as np
import pytest
import torch

from oracle import steps as S
from tests.gpu_util import assert_close, dev, say, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim(gpu_lib):
    from vivsim_amd import sim
    return sim


# grids of different sizes; D = ny / 8, markers round(pi D / 0.75), as in the sweep
SHAPES = [(96, 64, 2.0, 1.5, 120.0), (128, 64, 6.0, 4.0, 300.0), (160, 96, 4.0, 8.0, 500.0),
          (96, 128, 3.0, 2.5, 900.0)]


def sweep_case(cls, nx, ny, mr, ur, re, **kw):
    D = ny / 8
    base = dict(nx=nx, ny=ny, D=D, cx=nx / 4, cy=ny / 2, n_markers=int(round(np.pi * D / 0.75)), Re=re, Ur=ur,
                mass_ratio=mr, n_iter=3, pad=6, collision="kbc", forcing="edm", dof=(False, True))
    base.update(kw)
    return cls(**base)


def oracle_runs(cases, steps, v0):
    out = []
    for c in cases:
        f0 = c.initial_f(np.float64)
        out.append(S.viv_run(c, f0, steps, v=v0))
    return out


@pytest.mark.parametrize("tile_y", [None, 64, 128])
def test_batch_matches_oracle(sim, tile_y):
    steps = 16
    v0 = np.array([0.0, 1e-3])
    pcs = [sweep_case(sim.VIVCase, *s) for s in SHAPES]
    ocs = [sweep_case(S.VIV2D, *s) for s in SHAPES]
    say(f"batch of {len(pcs)} tile_y={tile_y}")
    b = sim.VIVBatch(pcs, tile_y=tile_y, hist_cap=steps)
    b.set_bodies(v=v0)
    f0 = torch.zeros(b.total, dtype=torch.float32, device="cuda")
    for k, c in enumerate(ocs):
        b.view(f0, k).copy_(dev(c.initial_f(np.float64)))
    got = b.run(f0, steps)
    b.check()
    assert b.clock() == steps
    for k, (f, d, v, a, dh, hh) in enumerate(oracle_runs(ocs, steps, v0)):
        assert_close(b.view(got, k), f, 5e-5, f"case {k} f")
        assert_close(b.hist_h[k][:steps], hh, 1e-4, f"case {k} h")
        assert_close(b.hist_d[k][:steps], dh, 1e-4, f"case {k} d")
        _, _, d_ = b.state(k)
        assert np.allclose(d_, d, rtol=1e-4, atol=1e-12)


@pytest.mark.parametrize("storage,tol", [("f16s", 3e-3), ("bf16s", 3e-2), ("e3m12s", 1.5e-3)])
def test_batch_16bit_matches_oracle(sim, storage, tol):
    """The batch in 16-bit shifted storage (each case keeps f - feq(1, (U0, 0)))."""
    steps = 16
    v0 = np.array([0.0, 1e-3])
    pcs = [sweep_case(sim.VIVCase, *s) for s in SHAPES]
    ocs = [sweep_case(S.VIV2D, *s) for s in SHAPES]
    say(f"batch of {len(pcs)} {storage}")
    b = sim.VIVBatch(pcs, storage=storage, tile_y=64, hist_cap=steps)
    b.set_bodies(v=v0)
    f0 = torch.zeros(b.total, dtype=torch.float32, device="cuda")
    for k, c in enumerate(ocs):
        b.view(f0, k).copy_(dev(c.initial_f(np.float64)))
    got = b.run(f0, steps)
    b.check()
    for k, (f, d, v, a, dh, hh) in enumerate(oracle_runs(ocs, steps, v0)):
        assert_close(b.view(got, k), f, tol, f"case {k} f ({storage})")


def test_batch_equals_single_case_runs(sim):
    """A batch is the set of its cases: each case equals its own VIV2D run."""
    steps = 12
    pcs = [sweep_case(sim.VIVCase, *s) for s in SHAPES[:3]]
    b = sim.VIVBatch(pcs)
    b.set_bodies(v=(0.0, 1e-3))
    f0 = b.initial_f()
    got = b.run(f0, steps)
    for k, c in enumerate(pcs):
        one = sim.VIV2D(c)
        one.set_body(v=(0.0, 1e-3))
        ref, _, _ = one.run(b.view(f0, k).clone(), steps)
        assert_close(b.view(got, k), to_np(ref).astype(np.float64), 1e-5, f"case {k}")


def test_batch_streamed_stores_equal_single_case_runs(sim):
    """A batch whose work buffer exceeds 256 MiB (the last-level cache) writes
    its full steps with non-temporal stores: two 2752 x 1536 cases (304 MB in
    FP32), each equal to its own VIV2D run (whose 152 MB buffer takes plain
    stores)."""
    steps = 4
    pcs = [sweep_case(sim.VIVCase, 2752, 1536, 2.0, 4.0, 150.0, D=40.0, n_markers=168, cx=688.0, cy=768.0),
           sweep_case(sim.VIVCase, 2752, 1536, 3.0, 6.0, 300.0, D=40.0, n_markers=168, cx=700.0, cy=760.0)]
    b = sim.VIVBatch(pcs)
    assert b.total * 4 > 256 * 2 ** 20
    b.set_bodies(v=(0.0, 1e-3))
    f0 = b.initial_f()
    say("batch streamed stores 2 x 2752x1536")
    got = b.run(f0, steps)
    b.check()
    for k, c in enumerate(pcs):
        one = sim.VIV2D(c)
        one.set_body(v=(0.0, 1e-3))
        ref, _, _ = one.run(b.view(f0, k).clone(), steps)
        assert_close(b.view(got, k), to_np(ref).astype(np.float64), 1e-5, f"streamed case {k}")


def test_batch_graph_replays(sim):
    """Replays of a captured 8-step graph continue the clock, body slots and
    histories: 3 replays == one 24-step run."""
    pcs = [sweep_case(sim.VIVCase, *s) for s in SHAPES[:2]]
    ref_b = sim.VIVBatch(pcs, hist_cap=24)
    ref_b.set_bodies(v=(0.0, 1e-3))
    f0 = ref_b.initial_f()
    ref = ref_b.run(f0, 24)
    b = sim.VIVBatch(pcs, hist_cap=24)
    b.set_bodies(v=(0.0, 1e-3))
    fa = f0.clone()
    fb = torch.empty_like(fa)
    g_ab = b.graph(fa, fb, 8)
    g_ba = b.graph(fb, fa, 8)
    say("graph replays")
    g_ab.launch()
    g_ba.launch()
    g_ab.launch()
    torch.cuda.synchronize()
    b.check()
    assert b.clock() == 24
    for k in range(len(pcs)):
        assert_close(b.view(fb, k), to_np(ref_b.view(ref, k)).astype(np.float64), 1e-5, f"graph case {k}")
        assert_close(b.hist_h[k], to_np(ref_b.hist_h[k]), 1e-5, f"graph h {k}")


def test_autotuned_tiles(sim):
    from vivsim_amd import tune
    c = sweep_case(sim.VIVCase, *SHAPES[2])
    t = tune.best_tile_y(c, steps=4, reps=1)
    assert t in tune.candidates(c.ny)
    b = sim.VIVBatch([c], tile_y="auto")
    assert b.tile_y == [t]
    b.set_bodies(v=(0.0, 1e-3))
    f = b.run(b.initial_f(), 6)
    b.check()
    assert torch.isfinite(f).all()


def test_batch_rejects_bad_input(sim):
    c = sweep_case(sim.VIVCase, *SHAPES[0])
    with pytest.raises(ValueError):
        sim.VIVBatch([c], tile_y=96)
    with pytest.raises(ValueError):
        sim.VIVBatch([c, sweep_case(sim.VIVCase, *SHAPES[1], collision="mrt", forcing="guo")])


def test_batch_split_mode_matches_oracle(sim):
    """Windows of 84 x 84 cells (D = 64, 268 markers) need more LDS than two
    bulk workgroups per CU leave: the batch then runs each step as an IB launch
    for all cases followed by a bulk launch.  Both cases must match the oracle,
    and a replayed graph must continue the run."""
    steps = 10
    v0 = np.array([0.0, 1e-3])
    shapes = [(192, 512, 2.0, 5.0, 300.0), (160, 512, 4.0, 3.0, 600.0)]
    pcs = [sweep_case(sim.VIVCase, *s, pad=10) for s in shapes]
    ocs = [sweep_case(S.VIV2D, *s, pad=10) for s in shapes]
    assert pcs[0].ib_size == 84
    say("batch split mode")
    b = sim.VIVBatch(pcs, hist_cap=2 * steps)
    b.set_bodies(v=v0)
    f0 = b.initial_f()
    got = b.run(f0, steps)
    b.check()
    for k, (f, d, v, a, dh, hh) in enumerate(oracle_runs(ocs, steps, v0)):
        assert_close(b.view(got, k), f, 5e-5, f"split case {k} f")
        assert_close(b.hist_h[k][:steps], hh, 1e-4, f"split case {k} h")
    # graph replays continue the clock: a 5-step graph replayed twice after the run
    ref = sim.VIVBatch(pcs, hist_cap=2 * steps)
    ref.set_bodies(v=v0)
    ref_out = ref.run(f0, 2 * steps)
    fa = got.clone()
    fb = torch.empty_like(fa)
    g_ab = b.graph(fa, fb, 5)
    g_ba = b.graph(fb, fa, 5)
    g_ab.launch()
    g_ba.launch()
    torch.cuda.synchronize()
    b.check()
    assert b.clock() == 2 * steps
    for k in range(len(pcs)):
        assert_close(b.view(fa, k), to_np(ref.view(ref_out, k)).astype(np.float64), 1e-5, f"split graph case {k}")
