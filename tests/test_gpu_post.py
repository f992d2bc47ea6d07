"""GPU parity of the field diagnostics (vivsim_amd.post) and the grid-refinement
transfers (vivsim_amd.multigrid) against the oracle."""

import numpy as np
import pytest
import torch

from oracle import multigrid as OMG
from oracle import post as OP
from tests.gpu_util import assert_close, dev, to_np

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(77)

FIELD_OPS = ["velocity_magnitude", "velocity_gradient", "vorticity", "vorticity_magnitude", "divergence",
             "strain_rate", "strain_rate_magnitude", "kinetic_energy", "enstrophy", "q_criterion"]


@pytest.fixture(scope="module")
def V(gpu_lib):
    import vivsim_amd
    return vivsim_amd


@pytest.mark.parametrize("shape", [(2, 37, 29), (3, 11, 9, 13)])
@pytest.mark.parametrize("op", FIELD_OPS)
def test_field_diagnostics(V, shape, op):
    u = rng.standard_normal(shape).astype(np.float32)
    got = getattr(V.post, op)(dev(u))
    ref = getattr(OP, op)(u.astype(np.float64))
    assert tuple(got.shape) == ref.shape
    assert_close(got, ref, 2e-6, op)


@pytest.mark.parametrize("shape", [(2, 300, 257), (3, 40, 33, 20)])
def test_means(V, shape):
    u = rng.standard_normal(shape).astype(np.float32)
    for op in ("mean_kinetic_energy", "mean_enstrophy"):
        got = float(getattr(V.post, op)(dev(u)).item())
        ref = float(getattr(OP, op)(u.astype(np.float64)))
        assert abs(got - ref) <= 1e-6 * abs(ref), (op, got, ref)


def test_pressure_coefficients_aliases(V):
    rho = (1 + 0.01 * rng.standard_normal((13, 7))).astype(np.float32)
    assert_close(V.post.pressure(dev(rho)), OP.pressure(rho.astype(np.float64)), 1e-7, "pressure")
    h = rng.standard_normal((50, 2))
    cd, cl = V.post.force_coefficients(torch.as_tensor(h, device="cuda"), 40.0, 0.05)
    rcd, rcl = OP.force_coefficients(h, 40.0, 0.05)
    assert np.allclose(to_np(cd), rcd, rtol=1e-15) and np.allclose(to_np(cl), rcl, rtol=1e-15)
    u = rng.standard_normal((2, 9, 8)).astype(np.float32)
    assert torch.equal(V.post.calculate_curl(dev(u)), V.post.vorticity(dev(u)))
    assert_close(V.post.calculate_vorticity_dimensionless(dev(u), 40.0, 0.05),
                 OP.vorticity(u.astype(np.float64)) * 40.0 / 0.05, 2e-6, "dimensionless vorticity")


def test_gradient_needs_two_points(V):
    with pytest.raises(ValueError):
        V.post.vorticity(dev(np.zeros((2, 1, 5), np.float32)))


@pytest.mark.parametrize("dir", ["left", "right", "up", "down", "north"])
def test_refinement_transfers_bit_exact(V, dir):
    nxc, nyc = 9, 12
    nxf, nyf = (14, 2 * nyc) if dir in ("left", "right", "north") else (2 * nxc, 10)
    ff = rng.standard_normal((9, nxf, nyf)).astype(np.float32)
    fc = rng.standard_normal((9, nxc, nyc)).astype(np.float32)
    got = V.multigrid.fine_to_coarse(dev(ff), dev(fc), dir)
    assert np.array_equal(to_np(got), OMG.fine_to_coarse(ff, fc, dir))
    got = V.multigrid.coarse_to_fine(dev(fc), dev(ff), dir)
    assert np.array_equal(to_np(got), OMG.coarse_to_fine(fc, ff, dir))


def test_refinement_host_helpers(V):
    assert V.multigrid.get_omega(0.02, 1) == OMG.get_omega(0.02, 1)
    assert V.multigrid.coord_to_indices(3.3, 7.9, 1, 2, 2) == OMG.coord_to_indices(3.3, 7.9, 1, 2, 2)
    f, r, u = V.multigrid.init_grid(10, 6, level=-1, buffer_x=1)
    assert tuple(f.shape) == (9, 6, 3) and tuple(u.shape) == (2, 6, 3) and bool((r == 1).all())
