"""Benchmark of the batched parameter-sweep path on one MI355X.

Workload (default): the four-level interferometer on a 2048 x 2048 (eps, A) grid, RK4 over
t in [0, 100] with dt = 0.05 (2000 steps), mean output.  One "step" of this benchmark is one
full sweep (4 194 304 points x 2000 RK steps).

Three forms of the same sweep are run:
  * fp64  -- FP64 kernel with the time-only drive terms tabulated once per launch (the form the
             headline value is measured on: it is the fastest one within 1e-6 of the FP64 oracle);
  * plain -- FP64 kernel evaluating the drive per point and stage, as the reference kernel does;
  * fp32  -- FP32 RHS and state, FP64 time table and mean accumulator (faster, but its error vs
             FP64 is reported, and it exceeds 1e-6 at about 1% of the points).
The FP64 result is the error reference (its parity with the CPU oracle is covered by the tests).

Prints one JSON line: value = points x steps x K / wall time of the K timed sweeps (inputs
resident in HBM, results left on the device).
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


def _expr_flops(e):
    """Arithmetic operations of an expression (integer powers as repeated products; calls of
    transcendental functions are counted apart, not here)."""
    import sympy as sp
    ops = sp.count_ops(e, visual=True)
    total = 0
    for term in sp.Add.make_args(ops):
        coeff, sym = term.as_coeff_Mul()
        name = str(sym)
        if name == "POW":
            continue
        if name in ("ADD", "SUB", "MUL", "NEG", "DIV"):
            total += int(coeff)
    for p in e.atoms(sp.Pow):
        if p.exp.is_Integer and p.exp > 1:
            total += int(p.exp) - 1
    return total


def adaptive_flops_per_attempt(plan, tableau, d):
    """Algorithmic FP64 FLOP of one attempted embedded-RK step of a d-state ODE: the stage
    right-hand sides and stage values (every stage after the first, plus the FSAL derivative),
    the stage inputs y + h sum a_sj k_j, the solution, the error estimate(s) and the RMS norm
    (2 per multiply-add; the step controller counted as 10).  Returns (FLOP, transcendental
    calls): one call per time atom and stage time."""
    ns = len(tableau["c"])
    f_rhs = sum(_expr_flops(e) for e in plan.rhs_out) + sum(_expr_flops(e)
                                                           for _, e in plan.rhs_temps)
    f_stage = sum(_expr_flops(e) for _, e in plan.stage)
    flops = ns * (f_rhs + f_stage)
    for row in tableau["a"][1:]:
        flops += d * (2 * sum(1 for v in row if v != 0.0) + 2)
    flops += d * (2 * sum(1 for v in tableau["b"] if v != 0.0) + 2)
    for key in ("e", "e5", "e3"):
        if key in tableau:
            flops += d * 2 * sum(1 for v in tableau[key] if v != 0.0)
    flops += d * 6 + 10
    return flops, (ns - 1) * len(plan.time_atoms)


def run_sweeps(w, fp64, k, time_table=True, **kw):
    import torch
    from gqis import mesolve_2D
    res, info = None, []
    lat = []
    for _ in range(k):
        t0 = time.perf_counter()
        res, ti = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                             var_arrays=w["var_arrays"], fp64=fp64, return_device=True,
                             return_timing_info=True, time_table=time_table,
                             ignore_non_finite_output=True, **kw)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        info.append(ti)
    return res, info, lat


def measured_traffic(kernel):
    """HBM bytes per launch of a four-level 2048 x 2048 kernel from its committed rocprofv3
    counter runs (profiles/pmc_<kernel>_four_level_2048.json: FETCH_SIZE and WRITE_SIZE scaled
    by a 1 GiB copy), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_four_level_2048.json")
    try:
        with open(path) as fh:
            return json.load(fh)["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def _four_level_liouvillian(u, tunnel=0.5):
    """Lindblad superoperator (row-major vec) of the four-level model at the instantaneous
    detuning u = eps + A cos(t) / 2: H = diag(u, -u, u + 2, -u - 2) + tunnelling on (i, i+1),
    relaxation sqrt(0.02) |i><i+1|, dephasing sqrt(0.01) diag(1, -1, 1, -1)."""
    import numpy as np
    h = np.diag([u, -u, u + 2.0, -u - 2.0]).astype(complex)
    for i in range(3):
        h[i, i + 1] = h[i + 1, i] = tunnel
    ops = []
    for i in range(3):
        op = np.zeros((4, 4), complex)
        op[i, i + 1] = np.sqrt(0.02)
        ops.append(op)
    ops.append(np.sqrt(0.01) * np.diag([1.0, -1.0, 1.0, -1.0]).astype(complex))
    eye = np.eye(4)
    sup = -1j * (np.kron(h, eye) - np.kron(eye, h.T))
    for c in ops:
        cc = c.conj().T @ c
        sup += np.kron(c, c.conj()) - 0.5 * np.kron(cc, eye) - 0.5 * np.kron(eye, cc.T)
    return sup


def rk4_stable_mask(eps, amp, dt):
    """(eps, A) points whose RK4 step is linearly stable at every drive phase: the generator
    depends on t only through u = eps + A cos(t) / 2, and the largest RK4 amplification
    |1 + z + z^2/2 + z^3/6 + z^4/24| over its eigenvalues z = lambda dt grows with |u| on each
    side, so a point is stable when [eps - A/2, eps + A/2] lies inside the interval
    [u_lo, u_hi] where the amplification stays <= 1 (end points by bisection)."""
    import numpy as np

    def growth(u):
        z = np.linalg.eigvals(_four_level_liouvillian(u)) * dt
        return float(np.max(np.abs(1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24)))

    def edge(sign):
        lo, hi = 0.0, 200.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if growth(sign * mid) <= 1.0 + 1e-12 else (lo, mid)
        return sign * lo
    u_lo, u_hi = edge(-1.0), edge(1.0)
    e, a = np.asarray(eps)[:, None], np.asarray(amp)[None, :]
    return (e - a / 2 >= u_lo) & (e + a / 2 <= u_hi)


# The GPU box gives one GPU's process a share of 16 host cores, whatever the machine has; CPU
# baselines use at most that many threads and say so.
CPU_SHARE = 16


def host_cpu():
    """CPU model, logical and physical core counts, and the cores this process may use."""
    model, cores = None, set()
    phys = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model is None:
                    model = val
                elif key == "physical id":
                    phys = val
                elif key == "core id":
                    cores.add((phys, val))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0))
    return {"model": model, "logical_cores": os.cpu_count(), "physical_cores": len(cores) or None,
            "usable_cores": usable, "threads": min(usable, CPU_SHARE),
            "thread_cap": f"min(usable cores, {CPU_SHARE}): the per-GPU host share of the box"}


def gpu_state():
    """Clocks, power and temperature of GPU 0 from rocm-smi (None where it is unavailable)."""
    import subprocess
    try:
        out = subprocess.run(["rocm-smi", "--showclocks", "--showpower", "--showtemp", "--json"],
                             capture_output=True, text=True, timeout=30).stdout
        cards = json.loads(out)
    except (OSError, ValueError, subprocess.SubprocessError):
        return None
    card = cards.get("card0") if isinstance(cards, dict) else None
    if not isinstance(card, dict):
        return None
    keep = ("clock", "power", "temperature")
    return {k: v for k, v in card.items() if any(w in k.lower() for w in keep)}


def cpu_baseline(grid=2048, seconds_budget=15.0):
    """FP64 C oracle (RK4, OpenMP over points) on a bounded sample of the same workload:
    eps columns spread evenly over the grid (2048 A values each) at the full 2000 steps, as
    many as fit the time budget (1 to 64).  Returns the baseline record and the sample
    (column indices, oracle map) for the GPU maps' error."""
    import numpy as np
    from oracle import cbridge, models
    host = host_cpu()
    threads = host["threads"]
    eps = np.linspace(-15.0, 15.0, grid)
    amp = np.linspace(0.0, 30.0, grid)
    probe = models.four_level_chain(eps[:1], amp[:64])
    t0 = time.perf_counter()
    cbridge.rk4_affine(probe, 0.05, 2000, threads=1)
    per_point = (time.perf_counter() - t0) / 64
    cols = max(1, int(seconds_budget * threads / (per_point * grid)))
    cols = min(cols, 64, grid)
    idx = np.unique(np.linspace(0, grid - 1, cols).round().astype(int))
    m = models.four_level_chain(eps[idx], amp)
    t0 = time.perf_counter()
    ref = np.asarray(cbridge.rk4_affine(m, 0.05, 2000, threads=threads)).reshape(len(idx), grid)
    dt = time.perf_counter() - t0
    record = {"value": len(idx) * grid * 2000 / dt, "unit": "param-point RK steps/s",
              "cores": threads, "kind": "port",
              "sample": f"{len(idx)} eps columns spread over the {grid} x {grid} grid x {grid} "
                        f"A values "
                        f"x 2000 RK4 steps, FP64 C oracle (OpenMP, {threads} threads), {dt:.1f} s",
              "host": host}
    return record, idx, ref


def map_error(gpu_map, idx, ref, stable):
    """Max absolute and relative (to the map's largest value) deviation of a GPU map's columns
    ``idx`` from the oracle's, over the RK4-stable points and over all finite points."""
    import numpy as np
    got = np.asarray(gpu_map)[idx].astype(np.complex128)
    err = np.abs(got - ref)
    fin = np.isfinite(err)
    gate = fin & np.asarray(stable)[idx]
    scale = float(np.max(np.abs(ref[gate]))) if gate.any() else float("nan")
    worst = float(np.max(err[gate])) if gate.any() else float("nan")
    return {"max_abs": worst, "max_rel_to_map_max": worst / scale, "map_max": scale,
            "frac_over_1e-6": float(np.mean(err[gate] > 1e-6)) if gate.any() else None,
            "max_abs_all_finite": float(np.max(err[fin])) if fin.any() else None,
            "points": int(gate.sum()), "columns": len(idx)}


def reference_cpu_path():
    """The reference's own CPU path restated (per-point Python RK4 on a lambdified RHS, spawn
    pool over eps columns, workers = cores - 1 as the reference defaults to) on a bounded
    sample: 16 eps columns x 64 A values spread over the 2048 x 2048 grid, full 2000 steps.
    The sample's results are compared with the C oracle's."""
    import numpy as np
    from oracle import cbridge, models, python_cpu
    cores = host_cpu()["threads"]
    eps = np.linspace(-15.0, 15.0, 2048)[::128]
    amp = np.linspace(0.0, 30.0, 2048)[::32]
    res, compute_s, workers = python_cpu.sweep("four_level", eps, amp, 0.05, 2000,
                                               workers=max(1, cores - 1))
    ref = np.asarray(cbridge.rk4_affine(models.four_level_chain(eps, amp), 0.05, 2000)).real
    ok = np.isfinite(ref.reshape(res.shape))
    dev = float(np.max(np.abs(res - ref.reshape(res.shape))[ok]))
    return {"value": res.size * 2000 / compute_s, "unit": "param-point RK steps/s",
            "cores": workers, "kind": "port",
            "sample": f"{len(eps)} eps x {len(amp)} A values of the 2048 x 2048 grid x 2000 RK4 "
                      f"steps, per-point Python loop on a lambdified RHS, {workers} spawn "
                      f"workers, {compute_s:.1f} s after worker preparation",
            "max_abs_dev_vs_c_oracle": dev}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30,
                    help="timed sweeps (at least 30 for the latency percentiles)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--profile", choices=["fp64", "fp32", "plain", "mixed", "mixed64",
                                          "tsit5a", "dop853a"],
                    default=None,
                    help="run only `--steps` sweeps of one form (for rocprofv3), no JSON line")
    ap.add_argument("--requests", type=int, default=1000, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: Poisson rate (req/s)")
    ap.add_argument("--workload", choices=["four_level", "lzs", "duffing", "serving", "scaling",
                                           "resonator", "multiqubit"],
                    default="four_level",
                    help="four_level: the headline sweep; lzs: the 64x64 two-level map with the "
                         "CPU oracle beside it; duffing: adaptive Tsit5/DOP853 sweep; serving: "
                         "Poisson request stream; scaling: Hilbert-space N-level models; "
                         "resonator: the reference's published qubit-resonator benchmarks")
    ap.add_argument("--photons", type=int, nargs="*", default=[2, 3, 4, 5, 6, 7, 8, 9, 10],
                    help="resonator: Fock states of the cavity (2 to 10)")
    ap.add_argument("--no-adaptive", action="store_true", help="resonator: RK4 rows only")
    ap.add_argument("--qubits", type=int, nargs="*", default=[3, 4, 5],
                    help="multiqubit: chain lengths (3 to 5)")
    ap.add_argument("--solvers", nargs="*",
                    default=["rk4", "tsit5", "dop853", "lserk4", "lserk4f", "tsit5a"],
                    help="multiqubit: solvers of the published rows")
    ap.add_argument("--levels", type=int, nargs="*", default=None,
                    help="scaling: only these N (4, 8, 16, 32)")
    ap.add_argument("--forms", default="",
                    help="scaling: only these forms, comma-separated (e.g. mfma,mfma_f16)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="with --profile: also run one device copy of this many bytes "
                         "(the counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the sweep shards by parameter rows with no "
              "communication, and multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from gqis import workloads
    from gqis.solver import _compile_rhs
    if args.workload == "duffing":
        return bench_duffing(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "scaling":
        return bench_scaling(args)
    if args.workload == "lzs":
        return bench_lzs(args)
    if args.workload == "resonator":
        return bench_resonator(args)
    if args.workload == "multiqubit":
        return bench_multiqubit(args)

    w = workloads.four_level(n_eps=args.grid, n_amp=args.grid)
    npts = args.grid * args.grid
    nsteps = len(w["tlist"]) - 1
    compiled, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                               sweep_symbols=tuple(w["var_arrays"]))
    if args.profile:
        fp64 = {"fp32": False, "mixed": "mixed", "mixed64": "mixed"}.get(args.profile, True)
        table = args.profile != "plain"
        kw = {"hiprtc_options": ("-DGQ_MIXED_ACC64=1",)} if args.profile == "mixed64" else {}
        run_sweeps(w, fp64, 1, time_table=table, **kw)          # compile + warm
        _, info, _ = run_sweeps(w, fp64, args.steps, time_table=table, **kw)
        log(f"profile {args.profile}: kernel {np.mean([i['gpu_kernel_s'] for i in info]):.6f} s")
        if args.copy_bytes:
            src = torch.empty(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            log(f"copy kernel: {args.copy_bytes} bytes read and written")
        return
    flops_ps, rhs_flops = algorithmic_flops_per_point_step(compiled)
    eps_v, amp_v = (np.asarray(v) for v in w["var_arrays"].values())
    dt = float(w["tlist"][1] - w["tlist"][0])

    log("FP64 reference sweep (time terms tabulated) ...")
    ref64, info64, _ = run_sweeps(w, True, 2)
    ref64 = ref64.cpu().numpy()
    fp64_kernel_s = float(np.mean([i["gpu_kernel_s"] for i in info64]))
    # Points whose RK4 step stays inside the method's linear stability region at every drive
    # phase (frozen-coefficient analysis).  Outside it RK4 amplifies every perturbation
    # exponentially, so no form of other precision can follow the FP64 map there; the gate is
    # evaluated on the stable points, and the count outside is reported.
    stable = rk4_stable_mask(eps_v, amp_v, dt)
    from gqis import mesolve_2D
    fin = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], fp64=True, output_mode="final_rho",
                     return_device=True, ignore_non_finite_output=True)
    bounded = (fin.abs().amax(dim=-1) <= 1.0).cpu().numpy()
    del fin
    plain = None
    if not args.no_plain:
        log("plain FP64 sweep (time terms per point) ...")
        _, infop, _ = run_sweeps(w, True, 2, time_table=False)
        plain = infop[-1]["gpu_kernel_s"]

    def form_error(res):
        fin_ = np.isfinite(ref64) & np.isfinite(res)
        err = np.abs(res.astype(np.complex128) - ref64)
        gate = fin_ & stable
        loose = fin_ & bounded
        return {"max_abs_err_stable": float(np.max(err[gate])),
                "frac_stable_points_over_1e-6": float(np.mean(err[gate] > 1e-6)),
                "meets_1e-6": bool(np.max(err[gate]) <= 1e-6),
                "max_abs_err_bounded_final_state": float(np.max(err[loose])),
                "bounded_final_state_points_over_1e-6": int(np.sum(err[loose] > 1e-6)),
                "nonfinite_points": int(np.sum(~fin_))}

    log("FP32 form ...")
    run_sweeps(w, False, 1)
    res32, info32, _ = run_sweeps(w, False, 3)
    res32 = res32.cpu().numpy()
    k32 = float(np.mean([i["gpu_kernel_s"] for i in info32]))
    forms = {"fp32": {"kernel_s": k32, "value": npts * nsteps / k32, **form_error(res32)}}
    del res32
    log("mixed form with an FP64 stage sum ...")
    acc64 = {"hiprtc_options": ("-DGQ_MIXED_ACC64=1",)}
    run_sweeps(w, "mixed", 1, **acc64)
    resm, infom, _ = run_sweeps(w, "mixed", 3, **acc64)
    km = float(np.mean([i["gpu_kernel_s"] for i in infom]))
    forms["mixed_acc64"] = {"kernel_s": km, "value": npts * nsteps / km,
                            **form_error(resm.cpu().numpy())}
    del resm

    log(f"headline (mixed form): warmup x{args.warmup} ...")
    run_sweeps(w, "mixed", args.warmup)
    torch.cuda.synchronize()
    gpu_before = gpu_state()
    log(f"headline timed x{args.steps} ...")
    t0 = time.perf_counter()
    res_t, info_t, lat_t = run_sweeps(w, "mixed", args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gpu_after = gpu_state()
    res_t = res_t.cpu().numpy()
    kernel_s = float(np.mean([i["gpu_kernel_s"] for i in info_t]))
    head_err = form_error(res_t)
    forms["mixed_acc32"] = {"kernel_s": kernel_s, "value": npts * nsteps / kernel_s, **head_err}
    point_steps = npts * nsteps
    achieved = flops_ps * point_steps / kernel_s / 1e12
    achieved64 = flops_ps * point_steps / fp64_kernel_s / 1e12
    out = {
        "metric": "param-point RK steps/s at <=1e-6 err vs FP64; p50/p99 sweep latency",
        "value": point_steps * args.steps / elapsed,
        "unit": "param-point RK steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"four-level interferometer {args.grid}x{args.grid}, RK4 "
                               f"2000 steps dt=0.05, mean of |0><0|",
                   "points": npts, "rk_steps": nsteps, "solver": "rk4",
                   "precision": "mixed: FP32 right-hand side and stage sum; FP64 state, sweep "
                                "values, statics, time table and stage values (fp64='mixed')"},
        "error_vs_fp64": {"reference": "FP64 GPU map (tabulated time terms), full grid",
                          "stable_points": int(stable.sum()),
                          "rk4_linearly_unstable_points": int((~stable).sum()),
                          "bounded_final_state_points": int(bounded.sum()), **head_err},
        "error_vs_fp64_oracle": None,
        "latency_s": {"p50": float(np.percentile(lat_t, 50)),
                      "p99": float(np.percentile(lat_t, 99)), "samples": len(lat_t)},
        "kernel_s": {"mixed_acc32": kernel_s, "fp64_tabulated": fp64_kernel_s,
                     "plain_fp64": plain, "fp32": k32, "mixed_acc64": km},
        "speedup_vs_plain_fp64": ({k: plain / v["kernel_s"] for k, v in forms.items()} |
                                  {"fp64_tabulated": plain / fp64_kernel_s}) if plain else None,
        "forms": forms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["fp32"],
                     "traffic": (measured_traffic("gq_rk4_mixed_f32")
                                 if args.grid == 2048 else None),
                     "flops_per_point_step": flops_ps, "kernel": "gq_rk4_mixed_f32",
                     "peak_note": "FP32 vector peak: the bound of a kernel whose FLOPs are "
                                  "mostly FP32 (its FP64 share runs at half that rate)"},
        "roofline_fp64_tabulated": {"bound": "valu", "achieved": achieved64,
                                    "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                                    "frac": achieved64 / PEAK_TFLOPS["fp64"],
                                    "traffic": (measured_traffic("gq_rk4_f64")
                                                if args.grid == 2048 else None),
                                    "kernel": "gq_rk4_f64"},
        "gpu_state": {"before_timed": gpu_before, "after_timed": gpu_after},
    }
    if not args.no_cpu:
        log("CPU baseline (FP64 C oracle sample; also the error reference) ...")
        out["cpu_baseline"], idx, ref = cpu_baseline(args.grid)
        # the timed headline map and the FP64 map against the oracle on the same columns
        mask = stable
        out["error_vs_fp64_oracle"] = map_error(res_t, idx, ref, mask)
        out["error_vs_fp64_oracle"]["form"] = "mixed_acc32 (the timed map)"
        out["error_vs_fp64_oracle"]["fp64_tabulated"] = map_error(ref64, idx, ref, mask)
        log("reference CPU path ...")
        out["cpu_reference_path"] = reference_cpu_path()
    print(json.dumps(out), flush=True)


def bench_duffing(args):
    """Forced Duffing sweep over (gamma, x0), t in [0, 500], output step 0.1, adaptive Tsit5 and
    DOP853 at rtol 1e-8 / atol 1e-10: final state of every point plus x(t) at the 5000 output
    times for an 8 x 8 subset (a second call on the subset: per-point arithmetic does not depend
    on the batch).  FP64 throughout, and FP32 RHS with FP64 state/controller beside it."""
    import numpy as np
    import torch
    from gqis import odesolve_2D, workloads
    side = args.grid if args.grid != 2048 else 1024
    w = workloads.duffing(n_gamma=side, n_x0=side)
    out = {"metric": "accepted adaptive point-steps/s", "unit": "accepted point-steps/s",
           "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "vs_baseline": None,
           "config": {"workload": f"forced Duffing {side}x{side}, t 0..500, dt_out 0.1, "
                                  "rtol 1e-8 atol 1e-10", "points": side * side}}
    finals = {}
    from gqis import tableaux
    from gqis.gqis_ode_api import ode_model
    from gqis.solver import _compile_rhs
    model = ode_model(w["rhs"], w["state_symbols"], w["y0"], w["var_arrays"])
    compiled, _ = _compile_rhs(model.equations, {}, [], model.observable,
                               sweep_symbols=tuple(model.sweeps), cse_simplify=True,
                               pre_expand=False, **model.symbol_kwargs())
    tabs = {"tsit5a": tableaux.TSIT5, "dop853a": tableaux.DOP853}
    if args.profile in tabs:
        ao = {"step_rtol": 1e-8, "step_atol": 1e-10, "max_substeps": 4096}
        kw = dict(var_arrays=w["var_arrays"], solver=args.profile, fp64=True, adaptive_options=ao,
                  return_device=True, return_timing_info=True, ignore_non_finite_output=True)
        odesolve_2D(w["rhs"], w["state_symbols"], w["y0"], w["tlist"][:11], **kw)
        for _ in range(max(1, args.steps)):
            _, info = odesolve_2D(w["rhs"], w["state_symbols"], w["y0"], w["tlist"], **kw)
            log(f"profile duffing {args.profile}: kernel {info['gpu_kernel_s']:.6f} s")
        return
    for solver in ("tsit5a", "dop853a"):
        for label, fp64, opts in (("fp64", True, {}), ("fp32_rhs", False, {"state_fp64": True})):
            ao = {"step_rtol": 1e-8, "step_atol": 1e-10, "max_substeps": 4096,
                  "count_steps": True, **opts}
            kw = dict(var_arrays=w["var_arrays"], solver=solver, fp64=fp64, adaptive_options=ao,
                      return_device=True, return_timing_info=True, ignore_non_finite_output=True)
            log(f"duffing {solver} {label} ...")
            odesolve_2D(w["rhs"], w["state_symbols"], w["y0"], w["tlist"][:11],
                        **kw)  # compile + warm on a short interval
            t0 = time.perf_counter()
            res, info = odesolve_2D(w["rhs"], w["state_symbols"], w["y0"], w["tlist"], **kw)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            acc = info["adaptive_steps"]["accepted"].astype(np.int64)
            rej = info["adaptive_steps"]["rejected"].astype(np.int64)
            k = info["gpu_kernel_s"]
            finals[(solver, label)] = res.cpu().numpy()
            out[f"{solver}_{label}"] = {
                "kernel_s": k, "wall_s": wall, "accepted_steps": int(acc.sum()),
                "rejected_steps": int(rej.sum()),
                "accepted_per_point": {"mean": float(acc.mean()), "max": int(acc.max()),
                                       "min": int(acc.min())},
                "value": float(acc.sum() / k),
                "nonfinite_points": int(np.sum(~np.isfinite(finals[(solver, label)]).any(-1)))}
            if fp64:
                f_att, n_trans = adaptive_flops_per_attempt(compiled.plan, tabs[solver], 2)
                attempts = int(acc.sum() + rej.sum())
                ach = f_att * attempts / k / 1e12
                out[f"{solver}_{label}"]["roofline"] = {
                    "bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS["fp64"],
                    "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS["fp64"],
                    "flops_per_attempted_step": f_att, "attempted_steps": attempts,
                    "transcendentals_per_attempted_step": n_trans,
                    "kernel": "gq_adaptive_f64"}
        a, b = finals[(solver, "fp64")], finals[(solver, "fp32_rhs")]
        d = np.abs(a - b).max(-1).ravel()
        out[f"{solver}_fp32_rhs_vs_fp64"] = {"median": float(np.nanmedian(d)),
                                             "p90": float(np.nanpercentile(d, 90)),
                                             "frac_below_1e-6": float(np.mean(d < 1e-6))}
    x = finals[("tsit5a", "fp64")]
    y = finals[("dop853a", "fp64")]
    dd = np.abs(x - y).max(-1).ravel()
    out["tsit5a_vs_dop853a_fp64"] = {"median": float(np.nanmedian(dd)),
                                     "frac_below_1e-6": float(np.mean(dd < 1e-6))}
    # x(t) at the 5000 output times for an 8 x 8 subset
    gi = np.arange(0, side, side // 8)
    sub = {k: v[gi] for k, v in w["var_arrays"].items()}
    t0 = time.perf_counter()
    fin_s, tr, times = odesolve_2D(w["rhs"], w["state_symbols"], w["y0"], w["tlist"],
                                   var_arrays=sub, solver="tsit5a", fp64=True,
                                   adaptive_options={"step_rtol": 1e-8, "step_atol": 1e-10,
                                                     "max_substeps": 4096},
                                   return_time_trace=True, ignore_non_finite_output=True)
    out["trace_subset"] = {"points": 64, "samples": int(tr.shape[2]),
                           "wall_s": time.perf_counter() - t0,
                           "final_matches_full_sweep": bool(np.array_equal(
                               fin_s, x.reshape(side, side, 2)[np.ix_(gi, gi)]))}
    t_best = out["tsit5a_fp64"]
    out["value"] = t_best["value"]
    out["roofline"] = t_best["roofline"]
    out["dtype"] = "f64"
    print(json.dumps(out), flush=True)


def bench_scaling(args):
    """Hilbert-space scaling: seeded random N-level models (N = 4, 8, 16, 32), 256 x 256 grid,
    RK4 2000 steps (dt = 0.01).  N = 4 runs the per-point register kernel; N = 8, 16, 32 the
    matrix-core path (FP32 MFMA, and FP16 / BF16 hi/lo split products), N = 8 and 16 also
    the per-point kernels, with one lane or 4 / 8 lanes (wavefronts) per point.  Errors are against the FP64
    GPU paths (per-point FP64 for N = 4 and 8, FP64 Hilbert-space path for N = 16, 32) on a
    16 x 16 subset of the grid."""
    import numpy as np
    import torch
    from gqis import mesolve_2D, workloads
    side = 256 if args.grid == 2048 else args.grid
    out = {"metric": "param-point RK steps/s (Hilbert-space scaling)",
           "unit": "param-point RK steps/s", "n_gpus": 1, "data": "synthetic",
           "config": {"workload": f"random N-level, {side}x{side}, RK4 2000 steps dt 0.01",
                      "points": side * side}, "results": {}}
    plan = ((4, ["lanes"]),
            (8, ["lanes", "lanes_g4", "lanes_g8", "mfma", "mfma_f16", "lanes64", "lanes_g8_64",
                 "mfma64"]),
            (16, ["lanes_g8", "mfma", "mfma_f16", "mfma_bf16", "mfma64"]),
            (32, ["mfma", "mfma_f16", "mfma_bf16", "mfma64"]))
    if args.levels:
        plan = tuple((n, [f for f in forms if not args.forms or f in args.forms.split(",")])
                     for n, forms in plan if n in args.levels)
    for n, forms in plan:
        w = workloads.random_n_level(n, n_eps=side, n_amp=side)
        sub = {k: v[:: side // 16] for k, v in w["var_arrays"].items()}
        args_m = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
        ref = mesolve_2D(*args_m, var_arrays=sub, fp64=True,
                         liouville="lanes" if n <= 8 else "mfma")
        for form in forms:
            fp64 = form.endswith("64")
            lv = form[:-2] if fp64 else form
            lv = lv.rstrip("_")
            groups = 1
            if lv.startswith("lanes_g"):
                lv, groups = "lanes", int(lv[len("lanes_g"):])
            kw = dict(var_arrays=w["var_arrays"], liouville=lv, return_device=True,
                      return_timing_info=True, ignore_non_finite_output=True, fp64=fp64,
                      lanes_per_point=groups)
            log(f"scaling N={n} {form} ...")
            mesolve_2D(*args_m, **kw)
            _, info = mesolve_2D(*args_m, **kw)
            torch.cuda.synchronize()
            sub_res = mesolve_2D(*args_m, var_arrays=sub, liouville=lv, fp64=fp64,
                                 lanes_per_point=groups)
            k = info["gpu_kernel_s"]
            ps = side * side * 2000
            d2 = n * n
            flops = (4 * 8 * n ** 3 + 4 * 12 * d2 + 13 * (d2 - 1)
                     if form.startswith("mfma") else None)
            entry = {"kernel_s": k, "value": ps / k,
                     "max_abs_err_vs_fp64": float(np.max(np.abs(sub_res - ref)))}
            if flops:
                ach = flops * ps / k / 1e12
                peak = 78.6 if fp64 else 157.3
                entry.update({"flops_per_point_step": flops, "achieved_tflops": ach,
                              f"mfma_{'fp64' if fp64 else 'fp32'}_peak_frac": ach / peak})
                if form not in ("mfma", "mfma64"):
                    # three half-precision products per FP32-equivalent product; from 16
                    # levels the two-sided kernel runs both sides of the commutator on the
                    # matrix cores (twice the products of H_eff rho alone) for the 16 x 16
                    # blocks on and above the diagonal (3 of 4 at 32 levels)
                    sides = 2 if n >= 16 else 1
                    blocks = {32: 0.75}.get(n, 1.0)
                    issued = sides * blocks * 3 * 4 * 8 * n ** 3 * ps / k / 1e12
                    entry.update({"half_mfma_issued_tflops": issued,
                                  "half_mfma_issued_frac": issued / 2500.0})
            out["results"][f"N={n} {form}"] = entry
    print(json.dumps(out), flush=True)


def bench_lzs(args):
    """The two-level LZS map (64 x 64, t in [0, 200] with 2001 samples, RK4, FP64, mean of sz)
    on the GPU and, on the host's cores, the FP64 C oracle over the same full grid; reports both
    rates and the GPU map's deviation from the oracle's."""
    import numpy as np
    import torch
    from gqis import mesolve_2D, workloads
    from oracle import cbridge, models
    w = workloads.lzs_two_level()
    args_m = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    got = mesolve_2D(*args_m, var_arrays=w["var_arrays"], fp64=True)
    times = []
    for _ in range(max(3, args.steps)):
        _, info = mesolve_2D(*args_m, var_arrays=w["var_arrays"], fp64=True, return_device=True,
                             return_timing_info=True)
        times.append(info["gpu_kernel_s"])
    torch.cuda.synchronize()
    eps, amp = w["var_arrays"].values()
    m = models.lzs_two_level(eps, amp)
    threads = host_cpu()["threads"]
    t0 = time.perf_counter()
    ref = cbridge.rk4_affine(m, 0.1, 2000, threads=threads)
    cpu_s = time.perf_counter() - t0
    point_steps = 64 * 64 * 2000
    k = float(np.median(times))
    out = {"metric": "param-point RK steps/s (two-level LZS map)", "unit": "param-point RK steps/s",
           "value": point_steps / k, "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
           "dtype": "f64", "vs_baseline": None,
           "config": {"workload": "two-level LZS 64x64, RK4 2000 steps dt 0.1, mean of sz",
                      "points": 64 * 64},
           "kernel_s": k, "max_abs_err_vs_cpu_oracle": float(np.max(np.abs(got.ravel() - ref))),
           "cpu_baseline": {"value": point_steps / cpu_s, "unit": "param-point RK steps/s",
                            "cores": threads, "kind": "port",
                            "sample": f"full 64x64 grid, 2000 RK4 steps, FP64 C oracle "
                                      f"(OpenMP, {threads} threads), {cpu_s:.3f} s"}}
    if not args.no_cpu:
        # the reference's own CPU path for this map (per-point Python RK4 on a lambdified RHS,
        # spawn pool over eps columns, cores - 1 workers), restated, over the full grid
        from oracle import python_cpu
        log("reference CPU path over the full 64x64 map ...")
        py_res, py_s, workers = python_cpu.sweep("lzs_two_level", eps, amp, 0.1, 2000,
                                                 workers=max(1, threads - 1))
        out["cpu_reference_path"] = {
            "value": point_steps / py_s, "unit": "param-point RK steps/s", "cores": workers,
            "kind": "port",
            "sample": f"full 64x64 grid, 2000 RK4 steps, per-point Python loop on a lambdified "
                      f"RHS, {workers} spawn workers, {py_s:.1f} s after worker preparation",
            "max_abs_dev_vs_gpu": float(np.max(np.abs(py_res.ravel() - got.real.ravel())))}
    print(json.dumps(out), flush=True)


# The reference's published timings on its own qubit-resonator benchmark (RTX 3080, FP32):
# Benchmarks/results/Benchmark_02_full_benchmark.csv:62 (2048 x 2048, RK4 5840 steps, solve-call
# wall time) and Benchmarks/results/Benchmark_04_hilbert_space_scaling_calibrated.csv:57-65, 75-83
# (1024 x 1024, kernel time: RK4 8611 steps; Tsit5A at rtol = atol = 1.77828e-4, 160 output
# intervals, mean accepted steps per point in the second entry).
RESONATOR_2048_RK4_S = 0.971001387
RESONATOR_1024_RK4_S = {2: 0.345004028, 3: 0.849800171, 4: 1.70617957, 5: 17.3547266,
                        6: 71.9960313, 7: 179.748641, 8: 302.470219, 9: 614.485,
                        10: 662.597125}
RESONATOR_1024_TSIT5A = {2: (0.272495605, 1294.33534), 3: (0.551332886, 1260.34076),
                         4: (5.4231543, 1241.68607), 5: (27.9381484, 1229.23429),
                         6: (85.9629922, 1219.59529), 7: (117.030633, 1211.54963),
                         8: (151.767547, 1204.46933), 9: (204.656266, 1198.04485),
                         10: (250.943922, 1192.12925)}


def bench_resonator(args):
    """The reference's own qubit-resonator benchmarks re-run here, same model, grid, step
    counts and precision (FP32), beside its published RTX 3080 times: the 2048 x 2048 four-level
    map (RK4, 5840 steps; call wall time and kernel time) and the Hilbert-space scaling rows at
    1024 x 1024 (RK4 8611 steps, Tsit5A at the published tolerance) for ``--photons``."""
    import numpy as np
    import torch
    from gqis import mesolve_2D, workloads
    rows = {}

    def call(w, tl, **kw):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], tl,
                               var_arrays={k: v.astype(np.float32)
                                           for k, v in w["var_arrays"].items()},
                               return_timing_info=True, **kw)
        return out, info, time.perf_counter() - t0

    if 2 in args.photons and not args.no_plain:
        w = workloads.qubit_resonator(2, n_eps=2048, n_amp=2048)
        tl = w["tlist"].astype(np.float32)
        log("qubit-resonator 2048x2048, RK4 5840 steps, FP32 ...")
        call(w, tl)                                         # symbolic model, compile
        walls, kern = [], []
        for _ in range(max(3, min(args.steps, 10))):
            out, info, wall = call(w, tl)
            walls.append(wall)
            kern.append(info["gpu_kernel_s"])
        _, info64, _ = call(w, tl, fp64=True)
        rows["2048x2048 rk4 5840 steps"] = {
            "call_wall_s": float(np.median(walls)), "kernel_s": float(np.median(kern)),
            "kernel_s_fp64": info64["gpu_kernel_s"], "published_rtx3080_s": RESONATOR_2048_RK4_S,
            "speedup_call_wall": RESONATOR_2048_RK4_S / float(np.median(walls)),
            "point_steps_per_s": 2048 * 2048 * 5840 / float(np.median(kern))}
        log(json.dumps(rows["2048x2048 rk4 5840 steps"]))
    for nph in args.photons:
        # 8611 RK4 steps over 40 periods; Tsit5A with four output points per period
        w = workloads.qubit_resonator(nph, n_eps=1024, n_amp=1024, steps_per_period=8611 / 40)
        tl = w["tlist"].astype(np.float32)
        log(f"qubit-resonator {2 * nph} levels ({(2 * nph) ** 2 - 1} states), 1024x1024 ...")
        t0 = time.perf_counter()
        call(w, tl[:2], output_mode="final")                # symbolic model only
        prep = time.perf_counter() - t0
        out, info, _ = call(w, tl)                          # compile
        state_before = gpu_state()
        out, info, _ = call(w, tl)
        state_after = gpu_state()
        row = {"levels": 2 * nph, "states": (2 * nph) ** 2 - 1, "symbolic_s": prep,
               "gpu_state_before_rk4": state_before, "gpu_state_after_rk4": state_after,
               "rk4_kernel_s": info["gpu_kernel_s"], "rk4_kernel": info.get("kernel"),
               "rk4_lanes_per_point": info.get("lanes_per_point"),
               "rk4_published_rtx3080_s": RESONATOR_1024_RK4_S[nph],
               "rk4_speedup": RESONATOR_1024_RK4_S[nph] / info["gpu_kernel_s"]}
        log(json.dumps(row))
        if not args.no_adaptive:
            tla = np.linspace(0.0, float(w["tlist"][-1]), 161)
            opts = {"step_rtol": 1.77828e-4, "step_atol": 1.77828e-4, "max_substeps": 128,
                    "count_steps": True}
            call(w, tla, solver="tsit5a", adaptive_options=opts, ignore_non_finite_output=True)
            out, info, _ = call(w, tla, solver="tsit5a", adaptive_options=opts,
                                ignore_non_finite_output=True)
            acc = info["adaptive_steps"]["accepted"]
            rej = info["adaptive_steps"]["rejected"]
            pub_s, pub_steps = RESONATOR_1024_TSIT5A[nph]
            row.update({"tsit5a_kernel_s": info["gpu_kernel_s"],
                        "tsit5a_kernel": info.get("kernel"),
                        "tsit5a_lanes_per_point": info.get("lanes_per_point"),
                        "tsit5a_mean_accepted": float(acc.mean()),
                        "tsit5a_rejected_frac": float(rej.sum() / max(acc.sum(), 1)),
                        "tsit5a_finite": bool(np.all(np.isfinite(out))),
                        "tsit5a_published_rtx3080_s": pub_s,
                        "tsit5a_published_mean_accepted": pub_steps,
                        "tsit5a_speedup": pub_s / info["gpu_kernel_s"]})
            log(json.dumps(row))
        rows[f"1024x1024 {2 * nph} levels"] = row
    print(json.dumps({"metric": "kernel time on the reference's qubit-resonator benchmarks",
                      "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f32",
                      "data": "synthetic", "vs_baseline": None,
                      "config": {"workload": "qubit-resonator (reference Benchmark 02 / 04), "
                                             "wd500 regime, 40 drive periods, mean of a"},
                      "results": rows}), flush=True)


# The reference's published multi-qubit chain rows (RTX 3080, FP32, 1024 x 1024 grid, 40 drive
# periods, calc_s = GPU time of the full-grid solve): Benchmarks/results/
# Benchmark_04_hilbert_space_scaling_calibrated.csv:94-96 (dop853), 99-101 (lserk4), 104-106
# (lserk4f), 109-111 (rk4; the 5-qubit run stopped at the 900 s limit), 114-116 (tsit5),
# 119-121 (tsit5a at accuracy 3.16228e-5: seconds and mean accepted steps per point).
MULTI_QUBIT_STEPS = {"rk4": 10240, "tsit5": 4305, "dop853": 1810, "lserk4": 7241, "lserk4f": 7241}
MULTI_QUBIT_PUBLISHED_S = {
    "dop853": {3: 17.4067402, 4: 296.216663, 5: 1657.27025},
    "lserk4": {3: 2.20887549, 4: 255.716538, 5: 2544.99575},
    "lserk4f": {3: 2.00129834, 4: 209.329688, 5: 2290.759},
    "rk4": {3: 2.17273145, 4: 390.043406, 5: None},
    "tsit5": {3: 3.36533301, 4: 357.227966, 5: 1907.40088},
    "tsit5a": {3: (8.92895996, 1269.62088), 4: (146.657688, 1251.5546),
               5: (880.714938, 1339.77004)}}


def bench_multiqubit(args):
    """The reference's multi-qubit chain rows of Benchmark 04 (n qubits, Hilbert dimension 2^n,
    sx gaps, ZZ neighbours, drive on qubit 0, relaxation and dephasing on every qubit) on the
    published grid (1024 x 1024), 40 drive periods, published step counts and FP32, beside the
    published RTX 3080 kernel times.  Up to 4 qubits every solver runs on the per-point
    kernels; at 5 qubits (1023 reduced states) every solver runs on the sparse Hilbert-space
    form, in 8 row blocks of the grid (kernel times summed; a block is 128 x 1024 points)."""
    import numpy as np
    import torch
    from gqis import mesolve_2D, workloads
    side = 1024 if args.grid == 2048 else args.grid
    w_drive = 1.14
    duration = 40 * 2.0 * np.pi / w_drive
    rows = {}

    def call(w, tl, var_arrays, **kw):
        torch.cuda.synchronize()
        out, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], tl,
                               var_arrays=var_arrays, return_timing_info=True,
                               ignore_non_finite_output=True, return_device=True, **kw)
        torch.cuda.synchronize()
        return out, info

    for nq in args.qubits:
        w = workloads.multi_qubit(nq, n_eps=side, n_amp=side)
        va = {k: v.astype(np.float32) for k, v in w["var_arrays"].items()}
        eps_s, amp_s = list(va)
        for solver in args.solvers:
            label = f"{nq} qubits {solver}"
            pub = MULTI_QUBIT_PUBLISHED_S[solver][nq]
            kw = {}
            if solver == "tsit5a":
                tl = np.linspace(0.0, duration, 161)
                kw = {"solver": "tsit5a",
                      "adaptive_options": {"step_rtol": 3.16228e-5, "step_atol": 3.16228e-5,
                                           "max_substeps": 128, "count_steps": True}}
            else:
                tl = np.linspace(0.0, duration, MULTI_QUBIT_STEPS[solver] + 1)
                kw = {"solver": solver}
            log(f"multi-qubit {nq} qubits ({4 ** nq - 1} states), {solver}, {side}x{side} ...")
            small = {eps_s: va[eps_s][:2], amp_s: va[amp_s][:2]}
            t0 = time.perf_counter()
            call(w, tl[:3], small, **kw)                     # symbolic model and kernel build
            prep = time.perf_counter() - t0
            blocks = 8 if nq >= 5 else 1
            kernel_s, kernels, finite, acc_sum, rej_sum = 0.0, set(), True, 0, 0
            step = side // blocks
            for b in range(blocks):
                part = {eps_s: va[eps_s][b * step:(b + 1) * step], amp_s: va[amp_s]}
                out, info = call(w, tl, part, **kw)
                kernel_s += info["gpu_kernel_s"]
                kernels.add(info.get("kernel"))
                finite &= bool(torch.isfinite(out).all().item())
                if solver == "tsit5a":
                    acc_sum += int(info["adaptive_steps"]["accepted"].sum())
                    rej_sum += int(info["adaptive_steps"]["rejected"].sum())
                if blocks > 1:
                    log(f"  block {b + 1}/{blocks}: {info['gpu_kernel_s']:.2f} s")
            row = {"states": 4 ** nq - 1, "kernel_s": kernel_s, "kernel": sorted(kernels),
                   "lanes_per_point": info.get("lanes_per_point"), "prepare_and_build_s": prep,
                   "finite": finite, "row_blocks": blocks}
            if solver == "tsit5a":
                row.update({"mean_accepted": acc_sum / side / side,
                            "rejected_frac": rej_sum / max(acc_sum, 1),
                            "published_rtx3080_s": pub[0],
                            "published_mean_accepted": pub[1], "speedup": pub[0] / kernel_s})
            else:
                row.update({"steps": MULTI_QUBIT_STEPS[solver],
                            "point_steps_per_s": side * side * MULTI_QUBIT_STEPS[solver] / kernel_s,
                            "published_rtx3080_s": pub,
                            "speedup": (pub / kernel_s) if pub else None})
                if pub is None:
                    row["published_note"] = "the reference's run stopped at its 900 s limit"
            rows[label] = row
            log(json.dumps({label: row}))
    print(json.dumps({"metric": "kernel time on the reference's multi-qubit chain benchmarks",
                      "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f32",
                      "data": "synthetic", "vs_baseline": None,
                      "config": {"workload": f"multi-qubit chain (reference Benchmark 04), "
                                             f"{side}x{side}, 40 drive periods, mean of qubit 0 "
                                             f"excited"},
                      "results": rows}), flush=True)


def bench_serving(args):
    """Poisson stream (--rate requests/s, seeded) of sweep requests: two-level LZS or N-level chain
    models (N in {2, 4, 8}), square grids of side 32/128/512/1024, random amplitude ranges,
    RK4 or Tsit5A, FP32 or FP64, served with continuous batching on 4 streams under a 16 GiB
    HBM budget.  Reports throughput and arrival-to-host-result latency percentiles."""
    import numpy as np
    from gqis import compile_lindblad_rhs, workloads
    from gqis.serving import SweepRequest, SweepServer
    stream = workloads.serving_stream(count=args.requests, rate=args.rate)
    models = {}
    for d in stream:
        key = (d["style"], d["n"])
        if key not in models:
            w = (workloads.lzs_two_level() if d["style"] == "lzs"
                 else workloads.chain(d["n"], eps_range=(-10.0, 10.0)))
            models[key] = (compile_lindblad_rhs(w["H"], w["Drive"], w["Col_Ops"],
                                                w["mean_operator"], list(w["var_arrays"])), w)
    reqs = []
    for d in stream:
        c, w = models[(d["style"], d["n"])]
        eps = np.linspace(-10.0, 10.0, d["side"])
        amp = np.linspace(0.0, d["amp_max"], d["side"])
        reqs.append(SweepRequest(c, w["tlist"], (eps, amp), solver=d["solver"], fp64=d["fp64"],
                                 arrival=d["arrival"]))
    srv = SweepServer(streams=4, hbm_budget_bytes=16 << 30)
    log(f"serving: compiling {len(models)} models x solver/precision variants ...")
    srv.warm(reqs)
    log(f"serving {len(reqs)} requests ...")
    t0 = time.perf_counter()
    done = srv.serve(reqs, realtime=True,
                     progress=lambda f, s, t: log(f"  {f} finished, {s} submitted, {t:.1f} s"))
    wall = time.perf_counter() - t0
    lat = np.array([r.done_at - r.arrival for r in done])
    pts = sum(r.shape[0] * r.shape[1] * (len(r.tlist) - 1) for r in done)
    out = {"metric": "sweep-request latency (arrival -> host result) under a Poisson stream",
           "unit": "s", "n_gpus": 1, "data": "synthetic", "vs_baseline": None,
           "config": {"workload": f"Poisson {args.rate:g} req/s, LZS / N-level chain (N 2,4,8), "
                                  "sides 32..1024, rk4/tsit5a, fp32/fp64", "requests": len(done),
                      "rate": args.rate},
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "mean_s": float(lat.mean()), "max_s": float(lat.max()),
           "requests_per_s": len(done) / wall, "param_point_steps_per_s": pts / wall,
           "wall_s": wall, "launches": srv.launches,
           "gpu_kernel_s_sum": srv.gpu_kernel_s, "host_launch_s": srv.host_launch_s,
           "gpu_by_kind": {f"D={k[0]} {k[1]} {k[2]}": {"launches": v[0], "points": v[1],
                                                        "gpu_s": v[2]}
                           for k, v in sorted(srv.gpu_by_kind.items())},
           "mean_batch_size": float(np.mean([r.batch_size for r in done])),
           "offered_point_steps_per_s": pts / max(r.arrival for r in done),
           "peak_hbm_bytes": srv.peak_in_use,
           "nonfinite_requests": int(sum(1 for r in done if r.status))}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
