#!/usr/bin/env python3
"""Throughput of the fused 2D IB-LBM vortex-induced-vibration step on one MI355X.

Workload (the headline configuration): D2Q9, 1024 x 1024 cells, MRT collision
with Guo forcing, a D = 40 cylinder at (256, 512) carried by 512 markers,
Peskin 4-point kernel, 3 multi-direct-forcing iterations, FP64 body-force
reduction and Newmark update, equilibrium inlet, pressure-NEE outlet,
periodic top / bottom, FP32 populations.  A "step" is one full IB-LBM time
step; ``value`` is GLUPS = cells * steps / seconds / 1e9 with the state
resident in HBM.

Also reported on the same line:
  roofline      : the fused step kernel's achieved algorithmic bandwidth
                  (72 B per cell in FP32 / 36 B in 16-bit storage) from HIP events
                  recorded around every launch in the timed region, vs 8 TB/s
  variants      : FP16 / BF16 / E3M12 shifted-storage GLUPS of the same step,
                  the step as two launches, 4096^2 in FP32 and 16-bit storage, the
                  3D sphere, a 64-case batch and the adjoint / forward ratio
  cpu_baseline  : the oracle (NumPy FP32 restatement of the reference step),
                  single-threaded on a bounded sample of the same workload
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_viv_fused launch on the 1024^2 FP32 workload, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes scaled by a 256 MiB copy (profiles/k_viv_fused_viv2d_1024x1024_f32_pmc.md); a profiler run
# cannot be part of the timed bench itself.
TRAFFIC_1024_F32 = 79692224


def kernel_times(lib, n):
    ms = (ctypes.c_float * n)()
    got = ctypes.c_int32(0)
    rc = lib.vsx_timer_stop(ms, n, ctypes.byref(got))
    assert rc == 0, lib.vsx_last_error()
    return np.array(ms[: got.value], dtype=np.float64)


def run_gpu(args, storage, steps, warmup, fused=True):
    import torch
    from vivsim_amd import _lib, lbm, sim

    case = sim.VIVCase(nx=args.nx, ny=args.ny)
    solver = sim.VIV2D(case, storage=storage, fused=fused)
    solver.set_body(v=(0.0, 1e-3 * case.U0))
    u = torch.as_tensor(case.initial_u(), dtype=torch.float32, device="cuda")
    f = lbm.get_equilibrium(torch.ones(case.nx, case.ny, device="cuda"), u)
    f, _, _ = solver.run(f, warmup, history=False)
    torch.cuda.synchronize()
    # throughput: the plain run, no instrumentation in the stream
    t0 = time.perf_counter()
    out, hd, hh = solver.run(f, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    assert torch.isfinite(out).all().item(), "non-finite state"
    solver.check()
    # kernel duration: a second run with a HIP event pair around every fused launch
    lib = _lib.lib()
    n_timed = min(steps, 400)
    assert lib.vsx_timer_start(n_timed) == 0
    solver.run(out, n_timed, history=False)
    torch.cuda.synchronize()
    kt = kernel_times(lib, n_timed)
    solver.check()
    cells = case.nx * case.ny
    sec = t1 - t0
    bulk_cells = case.nx * case.ny                  # every cell, wall columns included, is in the fused launch
    bpc = 72 if storage == "f32" else 36
    mean_ms = float(kt.mean())
    return dict(glups=cells * steps / sec / 1e9, ms_per_step=sec / steps * 1e3, kernel_ms=mean_ms,
                kernel_launches=int(kt.size), bytes_per_launch=bpc * bulk_cells,
                achieved_gbs=bpc * bulk_cells / (mean_ms * 1e-3) / 1e9,
                lift_last=float(hh[-1, 1].item()), d_last=float(hd[-1, 1].item()))


def timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def run_sphere(steps):
    """3D flow past a sphere: D3Q19 BGK + EDM, 256x128x128, D = 24, 4096 Fibonacci markers, 2 MDF iterations."""
    from vivsim_amd import sim
    case = sim.SphereCase()
    solver = sim.Sphere3D(case)
    f = solver.initial_f()
    f, _ = solver.run(f, 10, history=False)
    (out, hh), sec = timed(lambda: solver.run(f, steps))
    cells = case.nx * case.ny * case.nz
    fx = float(hh[-1, 0].item())
    return dict(glups=round(cells * steps / sec / 1e9, 3), ms_per_step=round(sec / steps * 1e3, 4), steps=steps,
                cd_last=round(case.drag_coefficient(fx), 5), config="sphere3d_d3q19_bgk_edm_256x128x128_D24_m4096_mdf2")


def sweep_cases(n=64, seed=0):
    """The batched sweep: grids from {256x256, 512x256, 512x512, 1024x512} drawn with a seed, D = NY / 8,
    markers round(pi D / 0.75), U_r ~ U[3, 8], Re ~ U[100, 1000], mass ratio ~ U[1, 10]; KBC + EDM, 3 MDF
    iterations, cylinder at (NX / 4, NY / 2)."""
    from vivsim_amd import sim
    rng = np.random.default_rng(seed)
    grids = [(256, 256), (512, 256), (512, 512), (1024, 512)]
    out = []
    for _ in range(n):
        nx, ny = grids[rng.integers(len(grids))]
        D = ny / 8
        out.append(sim.VIVCase(nx=nx, ny=ny, D=D, cx=nx / 4, cy=ny / 2, n_markers=int(round(np.pi * D / 0.75)),
                               Ur=float(rng.uniform(3, 8)), Re=float(rng.uniform(100, 1000)),
                               mass_ratio=float(rng.uniform(1, 10)), collision="kbc", forcing="edm", n_iter=3))
    return out


def run_batch(steps, chunk=100):
    import torch
    from vivsim_amd import sim, tune
    cases = sweep_cases()
    b = sim.VIVBatch(cases, tile_y="auto", hist_cap=steps + 2 * chunk)
    b.set_bodies(v=(0.0, 5e-5))
    fa = b.initial_f()
    fb = torch.empty_like(fa)
    g_ab, g_ba = b.graph(fa, fb, chunk), b.graph(fb, fa, chunk)
    g_ab.launch()
    g_ba.launch()
    reps = max(1, steps // (2 * chunk))

    def go():
        for _ in range(reps):
            g_ab.launch()
            g_ba.launch()

    _, sec = timed(go)
    b.check()
    n = 2 * chunk * reps
    return dict(glups=round(b.cells * n / sec / 1e9, 3), ms_per_step=round(sec / n * 1e3, 4), steps=n, cases=len(cases),
                cells_per_step=b.cells, graph_steps=chunk, tile_y=sorted(set(b.tile_y)),
                config="viv2d_batch64_kbc_edm_mdf3_hipgraph")


def run_adjoint(steps):
    """Adjoint / forward wall-time ratio: 512x256, D = 32 at (128, 128), 256 markers, MRT + Guo, Re = 150."""
    from vivsim_amd import sim
    case = sim.VIVCase(nx=512, ny=256, D=32.0, cx=128.0, cy=128.0, n_markers=256, Re=150.0, Ur=5.0, mass_ratio=2.0)
    adj = sim.VIVAdjoint(case, steps, v0=(0.0, 1e-3 * case.U0))
    delta = np.zeros(case.n_markers)
    adj.objective(delta, case.stiffness, case.U0)           # warm-up
    J_f, t_f = timed(lambda: adj.objective(delta, case.stiffness, case.U0))
    (J, gd, gk, gu), t_a = timed(lambda: adj.evaluate(delta, case.stiffness, case.U0))
    return dict(ratio=round(t_a / t_f, 3), forward_ms=round(t_f * 1e3, 2), adjoint_ms=round(t_a * 1e3, 2),
                steps=steps, segment=adj.seg, J=J, dJ_dK=gk, dJ_du_in=gu,
                dJ_ddelta_norm=float(gd.norm().item()), config="viv2d_adjoint_512x256_D32_m256_mrt_guo_re150")


def run_cpu(args, seconds_budget):
    """Oracle VIV step (float32, one thread) on the same 1024^2 workload."""
    from threadpoolctl import threadpool_limits

    from oracle import steps as S

    case = S.VIV2D(nx=args.nx, ny=args.ny)
    f = case.initial_f(np.float32)
    d, v, a = np.zeros(2), np.array([0.0, 1e-3 * case.U0]), np.zeros(2)
    n = 0
    with threadpool_limits(1):
        t0 = time.perf_counter()
        while True:
            f, d, v, a, h = S.viv_step(case, f, d, v, a)
            n += 1
            if time.perf_counter() - t0 > seconds_budget or n >= args.cpu_max_steps:
                break
        sec = time.perf_counter() - t0
    return dict(value=case.nx * case.ny * n / sec / 1e9, unit="GLUPS", cores=1, kind="port",
                sample=f"{n} oracle VIV steps (NumPy float32, 1 thread) on the {args.nx}x{args.ny} workload, "
                       f"{sec:.1f} s")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16s", "bf16s", "e3m12s"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-max-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 5 device-to-device copies of 256 MiB (calibration of HBM byte counters)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        a = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "copies": 5}))
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs the IB-LBM step on a single MI355X "
              "(multi-GPU slab decomposition is design-only); rerun with --gpus 1", file=sys.stderr)
        return 0

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)

    main_run = run_gpu(args, args.storage, args.steps, args.warmup)
    variants = {}
    if not args.no_variants:
        for st in ("f32", "f16s", "bf16s", "e3m12s"):
            if st != args.storage:
                r = run_gpu(args, st, args.steps, args.warmup)
                variants[st] = dict(glups=round(r["glups"], 3), ms_per_step=round(r["ms_per_step"], 5),
                                    kernel_us=round(r["kernel_ms"] * 1e3, 3), achieved_gbs=round(r["achieved_gbs"], 1))
        # the same FP32 step as two launches (IB kernel, then the bulk kernel)
        r = run_gpu(args, args.storage, args.steps, args.warmup, fused=False)
        variants["separate_launches"] = dict(glups=round(r["glups"], 3), ms_per_step=round(r["ms_per_step"], 5),
                                             bulk_kernel_us=round(r["kernel_ms"] * 1e3, 3))
        # beyond the 256 MB last-level cache: 4096^2 (2 x 1.2 GB populations), same body
        big = argparse.Namespace(**{**vars(args), "nx": 4096, "ny": 4096})
        r = run_gpu(big, "f32", 200, 20)
        variants["grid_4096"] = dict(glups=round(r["glups"], 3), ms_per_step=round(r["ms_per_step"], 5),
                                     kernel_us=round(r["kernel_ms"] * 1e3, 3), achieved_gbs=round(r["achieved_gbs"], 1),
                                     frac=round(r["achieved_gbs"] / HBM_PEAK_GBS, 4))
        # 16-bit storage where the streaming, not the IB workgroup, bounds the step
        for st in ("f16s", "e3m12s"):
            r16 = run_gpu(big, st, 200, 20)
            variants[f"grid_4096_{st}"] = dict(glups=round(r16["glups"], 3), ms_per_step=round(r16["ms_per_step"], 5),
                                               kernel_us=round(r16["kernel_ms"] * 1e3, 3),
                                               achieved_gbs=round(r16["achieved_gbs"], 1),
                                               frac=round(r16["achieved_gbs"] / HBM_PEAK_GBS, 4),
                                               speedup_vs_f32=round(r16["glups"] / r["glups"], 3))
        variants["sphere3d"] = run_sphere(100)
        variants["batch64"] = run_batch(400)
        variants["adjoint"] = run_adjoint(args.adjoint_steps)
    cpu = None if args.no_cpu else run_cpu(args, args.cpu_seconds)

    line = {
        "metric": "GLUPS & %HBM peak, 2D IB-LBM 1024² FP32 vs FP16-store; adjoint/fwd time ratio",
        "value": round(main_run["glups"], 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage} storage, f32 arithmetic",
        "data": "synthetic (seeded +-1% velocity perturbation on a uniform inflow)",
        "config": {"workload": f"viv2d_d2q9_mrt_guo_{args.nx}x{args.ny}_D40_m512_mdf3",
                   "nx": args.nx, "ny": args.ny, "markers": 512, "mdf_iterations": 3, "collision": "mrt",
                   "forcing": "guo", "kernel": "peskin_4pt", "storage": args.storage, "structure": "fp64 newmark"},
        "roofline": {"bound": "hbm", "achieved": round(main_run["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(main_run["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_1024_F32 if (args.nx, args.ny, args.storage) == (1024, 1024, "f32") else None,
                     "kernel": "k_viv_fused (IB workgroup + pull-stream/MRT/Guo tiles, one launch per step)",
                     "kernel_us": round(main_run["kernel_ms"] * 1e3, 3),
                     "bytes_per_launch": main_run["bytes_per_launch"], "launches": main_run["kernel_launches"]},
        "variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
