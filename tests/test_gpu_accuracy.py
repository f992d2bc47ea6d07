"""Accuracy of the storage formats against the float64 oracle over a long run:
the VIV case scaled to 256 x 256 (D = 20, 256 markers, MRT + Guo, Re = 100,
U_r = 5, mass ratio 2), 1000 steps.  Errors are stated for the lift
coefficient history, the transverse displacement history and the final
population, density and velocity fields, for FP32 and the 16-bit shifted storages (FP32 arithmetic: FP16, BF16 and
the E3M12 code), which keep f_i - feq_i(1, (U0, 0)); "f16s_rest" keeps
f_i - w_i instead.
The measured values are also written as JSON (accuracy.json) into the run's
output folder when it exists."""

import json
import os

import numpy as np
import pytest
import torch

from oracle import lbm2d as L2
from oracle import steps as S
from tests.gpu_util import say, to_np

pytestmark = pytest.mark.gpu

GEOM = dict(nx=256, ny=256, D=20.0, cx=64.0, cy=128.0, n_markers=256, Re=100.0, Ur=5.0, mass_ratio=2.0, pad=8)
STEPS = 1000
V0 = np.array([0.0, 5e-4])


@pytest.fixture(scope="module")
def fp64():
    c = S.VIV2D(**GEOM)
    f, d, v, a, dh, hh = S.viv_run(c, c.initial_f(np.float64), STEPS, v=V0)
    rho, u = L2.get_macroscopic(f)
    return c, dh, hh, u, rho, f


def rel(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30))


def test_storage_accuracy(gpu_lib, fp64):
    from vivsim_amd import lbm, sim
    c, dh, hh, u64, rho64, f64 = fp64
    cl = lambda h: h[:, 1] / (0.5 * c.U0 ** 2 * c.D)
    out = {}
    for name in ("f32", "f16s", "bf16s", "e3m12s", "f16s_rest"):
        storage = name.split("_")[0]
        pc = sim.VIVCase(**GEOM)
        solver = sim.VIV2D(pc, storage=storage)
        if name.endswith("_rest"):
            for a in range(3):
                solver.lbm.p.store_u[a] = 0.0
        solver.set_body(v=tuple(V0))
        f0 = torch.as_tensor(c.initial_f(np.float64), dtype=torch.float32, device="cuda")
        f, hd, hh_g = solver.run(f0, STEPS)
        solver.check()
        rg, ug = lbm.get_macroscopic(f)
        out[name] = dict(lift_coefficient=rel(cl(to_np(hh_g)), cl(hh)), displacement=rel(to_np(hd)[:, 1], dh[:, 1]),
                         velocity=rel(to_np(ug), u64), density=rel(to_np(rg), rho64),
                         populations=rel(to_np(f), f64))
        say(f"{name}: {out[name]}")
    out["case"] = dict(GEOM, steps=STEPS, collision="mrt", forcing="guo", reference="float64 NumPy oracle")
    if os.path.isdir("gpurun_out"):
        with open("gpurun_out/accuracy.json", "w") as fh:
            json.dump(out, fh, indent=1)
    # FP32 fields over 1000 steps within 1e-5 of the float64 oracle (max error relative to the
    # largest reference value); the lift and displacement histories, integrals over the body whose
    # early values are small, within 1e-4
    for k in ("populations", "density", "velocity"):
        assert out["f32"][k] < 1e-5, (k, out["f32"][k])
    assert out["f32"]["lift_coefficient"] < 1e-4 and out["f32"]["displacement"] < 1e-4
    assert out["f16s"]["velocity"] < 2e-2 and out["bf16s"]["velocity"] < 2e-1
    assert out["e3m12s"]["lift_coefficient"] < 1e-3 and out["e3m12s"]["velocity"] < 1e-3
