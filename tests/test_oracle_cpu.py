"""The C restatement of the cavity and VIV steps (oracle/c/viv_cpu.c, OpenMP;
the benchmark's CPU baseline) against the NumPy oracle on the same inputs."""

import numpy as np
import pytest

from oracle import cpu
from oracle import lbm2d as L2
from oracle import steps as S


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max())


def test_cavity_matches_numpy_oracle():
    """Lid-driven cavity, BGK, NEE on all four walls (lid on top), 60 steps on
    48 x 40: the float32 C step against the float64 NumPy step."""
    nx, ny, u0, nu = 48, 40, 0.1, 0.01
    omega = L2.get_omega(nu)
    f0 = L2.get_equilibrium(np.ones((nx, ny)), np.zeros((2, nx, ny)))
    fc = cpu.cavity_run(f0, omega, u0, 60)
    f = f0.copy()
    walls = S.cavity_walls(u0)
    for _ in range(60):
        f = S.lbm_step2d(f, omega, walls, "bgk")
    assert rel(fc, f) < 2e-6


@pytest.mark.parametrize("dof", [(False, True), (True, True)])
def test_viv_matches_numpy_oracle(dof):
    """The VIV step (MRT + Guo, Peskin-4 MDF with 3 iterations, FP64 Newmark),
    25 steps of a 160 x 120 case: field, displacement and force histories
    against the float64 NumPy oracle."""
    case = S.VIV2D(nx=160, ny=120, D=16.0, cx=48.0, cy=60.0, n_markers=64, Re=60.0, pad=6, dof=dof)
    f0 = case.initial_f(np.float64)
    v0 = np.array([2e-4, 1e-3])
    run = cpu.VivRunner(case)
    f = f0.astype(np.float32)
    body = np.array([0.0, 0.0, v0[0], v0[1], 0.0, 0.0])
    hd, hh = run.run(f, 25, body)
    fr, d, v, a, dh, hhist = S.viv_run(case, f0, 25, v=v0)
    assert rel(f, fr) < 1e-5
    assert rel(hh, hhist) < 1e-4 and rel(hd, dh) < 1e-4
    assert np.allclose(body[4:6], d, rtol=1e-4, atol=1e-12)
