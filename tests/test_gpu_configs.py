"""Parity at the workloads' own sizes: the 2D VIV case of the headline
benchmark (1024 x 1024, D = 40 at (256, 512), 512 markers, MRT + Guo, Peskin
4-point, 3 MDF iterations, FP32) against the float64 oracle, and the fused
one-launch step against the separate-launch path over a long run."""

import numpy as np
import pytest
import torch

from oracle import lbm2d as L2
from oracle import steps as S
from tests.gpu_util import rel_err, say, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim(gpu_lib):
    from vivsim_amd import sim
    return sim


def test_viv_1024_headline_case_matches_oracle(sim):
    """50 steps of the headline case (seeded +-1 % inflow perturbation, body
    started with a 1e-3 U0 transverse velocity) on the GPU and in the float64
    oracle: population field, velocity field, force and displacement histories."""
    steps = 50
    oc, pc = S.VIV2D(), sim.VIVCase()
    assert (oc.nx, oc.ny, oc.D, oc.cx, oc.cy, oc.n_markers) == (1024, 1024, 40.0, 256.0, 512.0, 512)
    f0 = oc.initial_f(np.float64)
    v0 = np.array([0.0, 1e-3 * oc.U0])
    solver = sim.VIV2D(pc)
    solver.set_body(v=tuple(v0))
    say("headline VIV 1024^2, 50 steps: GPU")
    got, hd, hh = solver.run(torch.as_tensor(f0, dtype=torch.float32, device="cuda"), steps)
    solver.check()
    say("headline VIV 1024^2, 50 steps: oracle (float64)")
    f, d, v, a, dh, hhist = S.viv_run(oc, f0, steps, v=v0)
    g = to_np(got).astype(np.float64)
    e_f = rel_err(g, f)
    e_u = rel_err(L2.get_macroscopic(g)[1], L2.get_macroscopic(f)[1])
    e_h = rel_err(hh, hhist)
    e_d = rel_err(hd, dh)
    say(f"headline VIV 1024^2, 50 steps: rel err f {e_f:.2e}, u {e_u:.2e}, h {e_h:.2e}, d {e_d:.2e}")
    assert e_f < 1e-5 and e_u < 1e-5 and e_h < 1e-4 and e_d < 1e-4


def test_launch_forms_agree_long_run(sim):
    """300 steps at 1024 x 1024 in every launch form: the pipelined step
    (hand-off between kernels on two streams) and the one-launch step (hand-off
    between workgroups of one launch) against the separate IB and bulk launches
    (no hand-off), which sum their spreads in a different order; fields and
    histories agree to FP32 rounding and no wait timed out."""
    steps = 300
    pc = sim.VIVCase()
    f0 = torch.as_tensor(S.VIV2D().initial_f(np.float32), dtype=torch.float32, device="cuda")
    out = {}
    for mode in ("pipelined", "one_launch", "separate"):
        solver = sim.VIV2D(pc, fused=mode != "separate", pipelined=mode == "pipelined")
        solver.set_body(v=(0.0, 1e-3 * pc.U0))
        say(f"{mode}: {steps} steps at 1024^2")
        f, hd, hh = solver.run(f0, steps)
        solver.check()
        out[mode] = (to_np(f), to_np(hd), to_np(hh), solver.state())
    ref = out["separate"]
    for mode in ("pipelined", "one_launch"):
        e_f, e_d, e_h = (rel_err(out[mode][k], ref[k]) for k in range(3))
        say(f"{mode} vs separate, {steps} steps: rel err f {e_f:.2e}, d {e_d:.2e}, h {e_h:.2e}")
        # the transverse displacement integrates the lift, which over these first steps is
        # a small difference of forces: its relative spread is the largest
        assert e_f < 1e-5 and e_h < 1e-5 and e_d < 1e-3


def test_sphere_drag_coefficient_matches_oracle(sim):
    """The sphere case (D3Q19 BGK + EDM, Peskin-4 with 2 MDF iterations, EQ inlet,
    pressure-NEE outlet, periodic lateral faces) scaled to 48 x 24 x 24 (D = 6,
    300 markers, Re = 40) and run for 1000 steps, past the impulsive start's
    acoustic transient: the drag coefficient history of the FP32 GPU run against
    the float64 oracle, and its mean over the last 200 steps."""
    from oracle import steps as S
    kw = dict(nx=48, ny=24, nz=24, D=6.0, cx=16.0, cy=12.0, cz=12.0, n_markers=300, Re=40.0, pad=3)
    steps = 1000
    oc, pc = S.Sphere3D(**kw), sim.SphereCase(**kw)
    solver = sim.Sphere3D(pc)
    say("sphere 48x24x24, 1000 steps: GPU")
    _, hh = solver.run(solver.initial_f(), steps)
    cd_gpu = np.array([pc.drag_coefficient(x) for x in to_np(hh)[:, 0]])
    say("sphere 48x24x24, 1000 steps: oracle (float64)")
    f = oc.initial_f(np.float64)
    W, I = oc.stencil(np.float64)
    hs = []
    for _ in range(steps):
        f, h = S.sphere_step(oc, f, W, I)
        hs.append(h[0])
    cd_ref = np.array([pc.drag_coefficient(x) for x in hs])
    e_hist = float(np.abs(cd_gpu - cd_ref).max() / np.abs(cd_ref).max())
    m_gpu, m_ref = cd_gpu[-200:].mean(), cd_ref[-200:].mean()
    say(f"sphere Cd (mean of the last 200 steps): GPU {m_gpu:.6f}, oracle {m_ref:.6f}, "
        f"rel diff {abs(m_gpu - m_ref) / abs(m_ref):.2e}; history rel err {e_hist:.2e}")
    assert abs(m_gpu - m_ref) <= 1e-4 * abs(m_ref) and e_hist <= 1e-3
    assert m_ref > 0    # drag points downstream
