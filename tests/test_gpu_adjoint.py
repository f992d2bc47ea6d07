"""GPU parity of the hand-written VIV adjoint (vsx_viv2d_adjoint) against
torch.autograd through the float64 restatement (oracle/viv_torch.py, the
stand-in for jax.grad of the reference) and against central finite
differences of the HIP forward run."""

import numpy as np
import pytest
import torch

from oracle import steps as S
from oracle import viv_torch as VT
from tests.gpu_util import say, to_np

pytestmark = pytest.mark.gpu

GEOM = dict(nx=80, ny=64, D=10.0, cx=18.0, cy=32.0, n_markers=40, U0=0.05, Re=60.0, n_iter=3, pad=6,
            dof=(False, True))
T = 40
V0 = (0.0, 1e-3)


@pytest.fixture(scope="module")
def sim(gpu_lib):
    from vivsim_amd import sim
    return sim


@pytest.fixture(scope="module")
def reference():
    """float64 autograd gradients of J at delta = 0.02 sin(3 theta)."""
    c = S.VIV2D(**GEOM)
    m = VT.VIVModel(c)
    th = 2 * np.pi * np.arange(c.n_markers) / c.n_markers
    delta = torch.tensor(0.02 * np.sin(3 * th), requires_grad=True)
    K = torch.tensor(c.stiffness, dtype=torch.float64, requires_grad=True)
    U = torch.tensor(c.U0, dtype=torch.float64, requires_grad=True)
    f0 = torch.tensor(c.initial_f(np.float64))
    J, _ = m.objective(f0, T, delta, K, U, v0=V0)
    J.backward()
    return dict(J=J.item(), gd=delta.grad.numpy(), gk=K.grad.item(), gu=U.grad.item(), delta=delta.detach().numpy(),
                K=c.stiffness, U=c.U0)


def test_adjoint_matches_autograd(sim, reference):
    c = sim.VIVCase(**GEOM)
    for seg in (7, 40):
        say(f"adjoint seg={seg}")
        adj = sim.VIVAdjoint(c, T, seg=seg, v0=V0)
        J, gd, gk, gu = adj.evaluate(reference["delta"], reference["K"], reference["U"])
        gd = to_np(gd)
        r = reference
        assert abs(J - r["J"]) <= 2e-4 * abs(r["J"]), (J, r["J"])
        assert np.linalg.norm(gd - r["gd"]) <= 2e-2 * np.linalg.norm(r["gd"]), (gd[:4], r["gd"][:4])
        assert abs(gk - r["gk"]) <= 2e-2 * abs(r["gk"]), (gk, r["gk"])
        assert abs(gu - r["gu"]) <= 5e-2 * abs(r["gu"]) + 1e-9, (gu, r["gu"])


def test_adjoint_matches_finite_differences(sim, reference):
    """Central differences of the HIP forward (FP32 field, FP64 displacement history): steps of
    1 % of K and 0.02 cells for delta keep the difference well above the FP32 noise of J."""
    c = sim.VIVCase(**GEOM)
    adj = sim.VIVAdjoint(c, T, seg=8, v0=V0)
    d0, k0, u0 = reference["delta"], reference["K"], reference["U"]
    J, gd, gk, gu = adj.evaluate(d0, k0, u0)
    gd = to_np(gd)
    h = 1e-2
    fd_k = (adj.objective(d0, k0 * (1 + h), u0) - adj.objective(d0, k0 * (1 - h), u0)) / (2 * h * k0)
    assert abs(fd_k - gk) <= 5e-2 * abs(gk), (fd_k, gk)
    for i in (3, 17):
        e = np.zeros_like(d0)
        e[i] = 2e-2
        fd = (adj.objective(d0 + e, k0, u0) - adj.objective(d0 - e, k0, u0)) / (2 * 2e-2)
        assert abs(fd - gd[i]) <= 5e-2 * np.abs(gd).max(), (i, fd, gd[i])


def test_autograd_function(sim, reference):
    c = sim.VIVCase(**GEOM)
    adj = sim.VIVAdjoint(c, T, v0=V0)
    delta = torch.tensor(reference["delta"], requires_grad=True)
    K = torch.tensor(reference["K"], dtype=torch.float64, requires_grad=True)
    U = torch.tensor(reference["U"], dtype=torch.float64, requires_grad=True)
    J = sim.VIVObjective.apply(delta, K, U, adj)
    (3.0 * J).backward()
    assert np.allclose(delta.grad.numpy(), 3.0 * reference["gd"], rtol=0, atol=6e-2 * np.abs(reference["gd"]).max())
    assert abs(K.grad.item() - 3.0 * reference["gk"]) <= 6e-2 * abs(3.0 * reference["gk"])
