"""The differentiable float64 restatement of the VIV step (oracle/viv_torch.py):
its forward equals the NumPy oracle's, and its autograd gradients equal
central finite differences of the objective."""

import numpy as np
import torch

from oracle import steps as S
from oracle import viv_torch as VT


def small():
    return S.VIV2D(nx=48, ny=40, D=8.0, cx=14.0, cy=20.0, n_markers=32, U0=0.05, Re=60.0, n_iter=3, pad=6,
                   dof=(False, True))


def test_forward_matches_numpy_oracle():
    c = small()
    f0 = c.initial_f(np.float64)
    v0 = np.array([0.0, 1e-3])
    fr, d, v, a, dh, hh = S.viv_run(c, f0, 6, v=v0)
    m = VT.VIVModel(c)
    z = torch.zeros(c.n_markers, dtype=torch.float64)
    J, hist = m.objective(torch.tensor(f0), 6, z, torch.tensor(c.stiffness, dtype=torch.float64),
                          torch.tensor(c.U0, dtype=torch.float64), v0=tuple(v0))
    assert np.allclose(hist.numpy(), dh, rtol=1e-12, atol=1e-18)
    assert np.isclose(J.item(), np.mean(dh[:, 1] ** 2), rtol=1e-12)


def test_autograd_matches_central_differences():
    c = small()
    f0 = torch.tensor(c.initial_f(np.float64))
    m = VT.VIVModel(c)
    T = 14
    rng = np.random.default_rng(5)
    delta = torch.tensor(0.05 * rng.standard_normal(c.n_markers), requires_grad=True)
    K = torch.tensor(c.stiffness, dtype=torch.float64, requires_grad=True)
    U = torch.tensor(c.U0, dtype=torch.float64, requires_grad=True)
    J, _ = m.objective(f0, T, delta, K, U, v0=(0.0, 1e-3))
    J.backward()

    def Jof(dl, k, u):
        with torch.no_grad():
            return m.objective(f0, T, dl, k, u, v0=(0.0, 1e-3))[0].item()

    dl, k, u = delta.detach(), K.detach(), U.detach()
    h = 1e-6
    fd_k = (Jof(dl, k + h * k, u) - Jof(dl, k - h * k, u)) / (2 * h * k)
    fd_u = (Jof(dl, k, u + h) - Jof(dl, k, u - h)) / (2 * h)
    assert np.isclose(K.grad.item(), fd_k.item(), rtol=1e-5)
    assert np.isclose(U.grad.item(), fd_u, rtol=1e-4, atol=1e-14)
    for i in (0, 7, 19):
        e = torch.zeros(c.n_markers, dtype=torch.float64)
        e[i] = h
        fd = (Jof(dl + e, k, u) - Jof(dl - e, k, u)) / (2 * h)
        assert np.isclose(delta.grad[i].item(), fd, rtol=1e-4, atol=1e-12), i
