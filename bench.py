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
  roofline      : the bulk pass kernel's achieved algorithmic bandwidth
                  (72 B per cell in FP32 / 36 B in 16-bit storage) from HIP events
                  recorded around every launch of a second run, vs 8 TB/s
  variants      : FP16 / BF16 / E3M12 shifted-storage GLUPS of the same step,
                  the step in its other launch forms (one launch per step, separate
                  IB and bulk launches), 4096^2 in FP32 and 16-bit storage, the
                  3D sphere, a 64-case batch and the adjoint / forward ratio
  cpu_baseline  : the reference's step restated in C (oracle/c, float32, OpenMP
                  over the host cores) on a bounded sample of the same workload
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def kernel_times(lib, n):
    ms = (ctypes.c_float * n)()
    got = ctypes.c_int32(0)
    rc = lib.vsx_timer_stop(ms, n, ctypes.byref(got))
    assert rc == 0, lib.vsx_last_error()
    return np.array(ms[: got.value], dtype=np.float64)


def run_gpu(args, storage, steps, warmup, mode="default"):
    """mode: "default" (the VIV2D default: pipelined, IB kernel of the next step on a second
    stream), "one_launch" or "separate"."""
    import torch
    from vivsim_amd import _lib, lbm, sim

    case = sim.VIVCase(nx=args.nx, ny=args.ny)
    kw = dict(default={}, one_launch=dict(pipelined=False), separate=dict(fused=False))[mode]
    solver = sim.VIV2D(case, storage=storage, **kw)
    solver.set_body(v=(0.0, 1e-3 * case.U0))
    u = torch.as_tensor(case.initial_u(), dtype=torch.float32, device="cuda")
    f0 = lbm.get_equilibrium(torch.ones(case.nx, case.ny, device="cuda"), u)
    # caller-owned fields and histories, as a loop of runs keeps them: the warm-up
    # run writes the buffer the timed run reads and fills the same histories
    f = torch.empty_like(f0)
    hist = (torch.empty((max(steps, warmup), 2), dtype=torch.float64, device="cuda"),
            torch.empty((max(steps, warmup), 2), dtype=torch.float64, device="cuda"))
    # The W warm-up steps run as two calls (W // 2, then W - W // 2 steps) when W >= 4. The first
    # call of a solver loads every kernel of a run; the second still carries a one-time cost of its
    # own (40-60 us at 1024^2, measured), which would otherwise land in the timed call; with three
    # or more steps the second call runs every kernel of a run (first, full and last steps).
    if warmup >= 4:
        solver.run(f0, warmup // 2, out=f, history=hist)
        solver.run(f, warmup - warmup // 2, out=f0, history=hist)
        f, out = f0, f
    else:
        f, _, _ = solver.run(f0, warmup, out=f, history=hist)
        out = f0
    torch.cuda.synchronize()
    # throughput: the plain run, no instrumentation in the stream
    t0 = time.perf_counter()
    out, hd, hh = solver.run(f, steps, out=out, history=hist)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    assert torch.isfinite(out).all().item(), "non-finite state"
    solver.check()
    # kernel duration: a second run with a HIP event pair around every bulk launch
    lib = _lib.lib()
    n_timed = min(steps, 400)
    assert lib.vsx_timer_start(n_timed) == 0
    solver.run(out, n_timed, history=False)
    torch.cuda.synchronize()
    kt = kernel_times(lib, n_timed)
    solver.check()
    cells = case.nx * case.ny
    sec = t1 - t0
    bpc = 72 if storage == "f32" else 36
    # the full-step launches: the run's first launch (it reads the caller's field and its
    # window tiles wait for the run's first IB result) and its last (it pushes into the
    # caller's output) are kernels of their own
    full = kt[1:-1] if kt.size >= 3 else kt
    mean_ms = float(full.mean())
    return dict(glups=cells * steps / sec / 1e9, ms_per_step=sec / steps * 1e3, kernel_ms=mean_ms,
                kernel_launches=int(full.size), bytes_per_launch=bpc * cells,
                achieved_gbs=bpc * cells / (mean_ms * 1e-3) / 1e9, mode=solver.ib.fused,
                lift_last=float(hh[steps - 1, 1].item()), d_last=float(hd[steps - 1, 1].item()))


def timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def run_sphere(steps, storage="f32", cd_steps=2000):
    """3D flow past a sphere: D3Q19 BGK + EDM, 256x128x128, D = 24, 4096 Fibonacci markers, 2 MDF iterations.
    GLUPS from `steps` timed steps; the drag coefficient C_d = F_x / (1/2 U0^2 pi D^2 / 4) of the marker force
    history, averaged over the last 200 of a `cd_steps` run from the initial state (the impulsive start's
    acoustic transient has passed; its agreement with the float64 oracle is tested on a reduced grid,
    tests/test_gpu_configs.py)."""
    from vivsim_amd import sim
    case = sim.SphereCase()
    solver = sim.Sphere3D(case, storage=storage)
    f0 = solver.initial_f()
    _, hh = solver.run(f0, cd_steps)
    cd = [case.drag_coefficient(x) for x in hh[-200:, 0].cpu().numpy()]
    f, _ = solver.run(f0, 10, history=False)
    (out, hh), sec = timed(lambda: solver.run(f, steps))
    cells = case.nx * case.ny * case.nz
    return dict(glups=round(cells * steps / sec / 1e9, 3), ms_per_step=round(sec / steps * 1e3, 4), steps=steps,
                storage=storage, cd_mean=round(float(np.mean(cd)), 5), cd_steps=cd_steps,
                config="sphere3d_d3q19_bgk_edm_256x128x128_D24_m4096_mdf2")


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
    """The reference's VIV step restated in C (oracle/c/viv_cpu.c, OpenMP over the
    host cores OMP_NUM_THREADS allows) on the same 1024^2 workload, for a bounded
    number of steps."""
    from oracle import cpu
    from oracle import steps as S

    case = S.VIV2D(nx=args.nx, ny=args.ny)
    run = cpu.VivRunner(case)
    f = case.initial_f(np.float32)
    body = np.array([0.0, 0.0, 0.0, 1e-3 * case.U0, 0.0, 0.0])
    run.run(f, 2, body)                                    # warm-up (first touch of the buffers)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds_budget and n < args.cpu_max_steps:
        run.run(f, 5, body)
        n += 5
    sec = time.perf_counter() - t0
    return dict(value=case.nx * case.ny * n / sec / 1e9, unit="GLUPS", cores=cpu.threads(), kind="port",
                sample=f"{n} VIV steps of the {args.nx}x{args.ny} workload by the C restatement of the reference "
                       f"step (float32, OpenMP, {cpu.threads()} threads), {sec:.1f} s")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16s", "bf16s", "e3m12s"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-max-steps", type=int, default=5000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--form", default="default", choices=["default", "one_launch", "separate"],
                    help="launch form of the headline run (default: the VIV2D default, the pipelined step); "
                         "byte-counter passes use one_launch, whose kernels never wait for one another")
    ap.add_argument("--only", choices=["grid_4096", "grid_4096_f16s", "grid_4096_e3m12s", "sphere3d", "sphere3d_f16s",
                                       "batch64", "adjoint"],
                    help="profiling aid: run only this variant (no headline line) and print its result")
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

    if args.only:
        big = argparse.Namespace(**{**vars(args), "nx": 4096, "ny": 4096})
        fns = {"grid_4096": lambda: run_gpu(big, "f32", 200, 20),
               "grid_4096_f16s": lambda: run_gpu(big, "f16s", 200, 20),
               "grid_4096_e3m12s": lambda: run_gpu(big, "e3m12s", 200, 20),
               "sphere3d": lambda: run_sphere(100), "sphere3d_f16s": lambda: run_sphere(100, "f16s"),
               "batch64": lambda: run_batch(400), "adjoint": lambda: run_adjoint(args.adjoint_steps)}
        print(json.dumps({args.only: fns[args.only]()}), flush=True)
        return 0

    main_run = run_gpu(args, args.storage, args.steps, args.warmup, args.form)
    print(f"headline: {main_run['glups']:.3f} GLUPS, {main_run['ms_per_step'] * 1e3:.2f} us / step, "
          f"bulk kernel {main_run['kernel_ms'] * 1e3:.2f} us", file=sys.stderr, flush=True)
    variants = {}

    def variant(name, fn):
        """A variant that fails is recorded as such; it never loses the headline line."""
        try:
            variants[name] = fn()
        except Exception as e:  # noqa: BLE001
            variants[name] = {"error": f"{type(e).__name__}: {e}"}
        print(f"variant {name}: {variants[name]}", file=sys.stderr, flush=True)

    def short(r, **extra):
        d = dict(glups=round(r["glups"], 3), ms_per_step=round(r["ms_per_step"], 5),
                 kernel_us=round(r["kernel_ms"] * 1e3, 3), achieved_gbs=round(r["achieved_gbs"], 1))
        d.update(extra)
        return d

    if not args.no_variants:
        for st in ("f32", "f16s", "bf16s", "e3m12s"):
            if st != args.storage:
                variant(st, lambda st=st: short(run_gpu(args, st, args.steps, args.warmup)))
        # the same step in the other launch forms
        variant("one_launch", lambda: short(run_gpu(args, args.storage, args.steps, args.warmup, "one_launch")))
        variant("separate_launches", lambda: short(run_gpu(args, args.storage, args.steps, args.warmup, "separate")))
        # beyond the 256 MB last-level cache: 4096^2 (2 x 1.2 GB populations), same body
        big = argparse.Namespace(**{**vars(args), "nx": 4096, "ny": 4096})
        r32 = {}

        def big_f32():
            r32.update(run_gpu(big, "f32", 200, 20))
            return short(r32, frac=round(r32["achieved_gbs"] / HBM_PEAK_GBS, 4))
        variant("grid_4096", big_f32)
        # 16-bit storage where the streaming, not the IB, bounds the step
        for st in ("f16s", "e3m12s"):
            def big16(st=st):
                r16 = run_gpu(big, st, 200, 20)
                return short(r16, frac=round(r16["achieved_gbs"] / HBM_PEAK_GBS, 4),
                             speedup_vs_f32=round(r16["glups"] / r32["glups"], 3) if r32 else None)
            variant(f"grid_4096_{st}", big16)
        variant("sphere3d", lambda: run_sphere(100))
        variant("sphere3d_f16s", lambda: run_sphere(100, "f16s"))
        variant("batch64", lambda: run_batch(400))
        variant("adjoint", lambda: run_adjoint(args.adjoint_steps))
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
                     # HBM bytes per launch come from rocprofv3 FETCH_SIZE / WRITE_SIZE passes, which cannot
                     # run inside this timed process; the measured figures are kept under profiles/
                     "traffic": None,
                     "kernel": {2: "k_viv_fused, full-step launches (pull-stream + walls + MRT/Guo bulk pass; "
                                   "pipelined form: the IB kernel of the next step runs on a second stream)",
                                1: "k_viv_fused (one launch per step: IB workgroup + pull-stream/MRT/Guo tiles)",
                                0: "k_bulk (bulk pass of the separate IB and bulk launches)"}[main_run["mode"]],
                     "kernel_us": round(main_run["kernel_ms"] * 1e3, 3),
                     "bytes_per_launch": main_run["bytes_per_launch"], "launches": main_run["kernel_launches"]},
        "variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
