"""Benchmark of the batched parameter-sweep path on one MI355X.

Workload (default): the four-level interferometer on a 2048 x 2048 (eps, A) grid, RK4 over
t in [0, 100] with dt = 0.05 (2000 steps), mean output, in the FP32 form (FP32 RHS/state, FP64
time table and mean accumulator).  One "step" of this benchmark is one full sweep
(4 194 304 points x 2000 RK steps).  The FP64 sweep of the same grid is run once as the
error reference (its parity with the CPU oracle is covered by the test-suite), and the plain
FP64 form (time terms evaluated per point, as the reference kernel does) is timed beside it.

Prints one JSON line: metric "param-point RK steps/s", value = points x steps x K / time of the
K timed sweeps (inputs resident, results left on the device).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-quantum-interferometry-solver_amd"))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}   # MI355X published vector peaks


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_flops_per_point_step(compiled):
    """4 RHS evaluations (2 flops per state coefficient, 1 per constant term) + 13 flops per
    state element for the RK4 combinations + the observable's flops."""
    import sympy as sp
    plan = compiled.plan
    d = plan.state_size
    rhs = 0
    for e in plan.rhs_out:
        terms = sp.Add.make_args(sp.expand(e))
        for term in terms:
            rhs += 2 if term.free_symbols & set(plan.states) else 1
    obs = sum(1 for t in sp.Add.make_args(sp.expand(plan.obs_out[0]))
              if t.free_symbols & set(plan.states))
    return 4 * rhs + 13 * d + obs, rhs


def run_sweeps(w, fp64, k, time_table=True):
    import torch
    from gqis import mesolve_2D
    res, info = None, []
    lat = []
    for _ in range(k):
        t0 = time.perf_counter()
        res, ti = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                             var_arrays=w["var_arrays"], fp64=fp64, return_device=True,
                             return_timing_info=True, time_table=time_table,
                             ignore_non_finite_output=True)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        info.append(ti)
    return res, info, lat


def cpu_baseline(seconds_budget=15.0):
    """FP64 C oracle (RK4, OpenMP over points) on a bounded sample of the same workload:
    whole eps columns (2048 A values each) at the full 2000 steps, as many columns as fit the
    time budget (at least one)."""
    import numpy as np
    from oracle import cbridge, models
    threads = len(os.sched_getaffinity(0))
    threads = min(threads, 16)
    eps = np.linspace(-15.0, 15.0, 2048)
    amp = np.linspace(0.0, 30.0, 2048)
    probe = models.four_level_chain(eps[:1], amp[:64])
    t0 = time.perf_counter()
    cbridge.rk4_affine(probe, 0.05, 2000, threads=1)
    per_point = (time.perf_counter() - t0) / 64
    cols = max(1, int(seconds_budget * threads / (per_point * 2048)))
    cols = min(cols, 64)
    m = models.four_level_chain(eps[:cols], amp)
    t0 = time.perf_counter()
    cbridge.rk4_affine(m, 0.05, 2000, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": cols * 2048 * 2000 / dt, "unit": "param-point RK steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{cols} eps column(s) x 2048 A values x 2000 RK4 steps, FP64 C oracle "
                      f"(OpenMP, {threads} threads), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the sweep shards by parameter rows with no "
              "communication, and multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from gqis import workloads
    from gqis.solver import _compile_rhs

    w = workloads.four_level(n_eps=args.grid, n_amp=args.grid)
    npts = args.grid * args.grid
    nsteps = len(w["tlist"]) - 1
    compiled, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                               sweep_symbols=tuple(w["var_arrays"]))
    flops_ps, rhs_flops = algorithmic_flops_per_point_step(compiled)

    log("FP64 reference sweep ...")
    ref64, info64, lat64 = run_sweeps(w, True, 2)
    ref64 = ref64.cpu().numpy()
    # RK4 at dt = 0.05 is unstable in a corner of the grid (|2 eps + 2 + A cos t| dt beyond the
    # stability limit): there the state leaves the set of density matrices and rounding
    # differences are amplified without bound.  Those points are reported, not gated.
    from gqis import mesolve_2D
    fin = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], fp64=True, output_mode="final_rho",
                     return_device=True, ignore_non_finite_output=True)
    stable = (fin.abs().amax(dim=-1) <= 1.0).cpu().numpy()
    del fin
    fp64_kernel_s = info64[-1]["gpu_kernel_s"]
    plain = None
    if not args.no_plain:
        log("plain FP64 sweep (time terms per point) ...")
        _, infop, _ = run_sweeps(w, True, 2, time_table=False)
        plain = infop[-1]["gpu_kernel_s"]

    log(f"FP32 warmup x{args.warmup} ...")
    run_sweeps(w, False, args.warmup)
    torch.cuda.synchronize()
    log(f"FP32 timed x{args.steps} ...")
    t0 = time.perf_counter()
    res32, info32, lat32 = run_sweeps(w, False, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    res32 = res32.cpu().numpy()

    finite = np.isfinite(ref64) & np.isfinite(res32)
    gate = finite & stable
    err = np.abs(res32.astype(np.complex128) - ref64)
    max_err = float(np.max(err[gate])) if gate.any() else float("nan")
    frac_over = float(np.mean(err[gate] > 1e-6)) if gate.any() else float("nan")
    max_err_all = float(np.max(err[finite])) if finite.any() else float("nan")
    kernel_s = float(np.mean([i["gpu_kernel_s"] for i in info32]))
    point_steps = npts * nsteps
    achieved = flops_ps * point_steps / kernel_s / 1e12
    out = {
        "metric": "param-point RK steps/s at <=1e-6 err vs FP64; p50/p99 sweep latency",
        "value": point_steps * args.steps / elapsed,
        "unit": "param-point RK steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"four-level interferometer {args.grid}x{args.grid}, RK4 "
                               f"2000 steps dt=0.05, mean of |0><0|",
                   "points": npts, "rk_steps": nsteps, "solver": "rk4",
                   "precision": "fp32 RHS/state, fp64 time table + mean accumulator"},
        "error_vs_fp64": {"max_abs_stable": max_err, "frac_stable_points_over_1e-6": frac_over,
                          "rk4_unstable_points": int(np.sum(~stable)),
                          "max_abs_all_points": max_err_all,
                          "nonfinite_points": int(np.sum(~finite))},
        "latency_s": {"p50": float(np.percentile(lat32, 50)), "p99": float(np.percentile(lat32, 99)),
                      "samples": len(lat32)},
        "fp64_kernel_s": fp64_kernel_s, "plain_fp64_kernel_s": plain, "fp32_kernel_s": kernel_s,
        "speedup_vs_plain_fp64": (plain / kernel_s) if plain else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["fp32"], "traffic": None,
                     "flops_per_point_step": flops_ps, "kernel": "gq_rk4"},
    }
    if not args.no_cpu:
        log("CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
