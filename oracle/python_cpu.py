"""The reference's CPU path for the sweep, restated: a per-point Python RK4 loop on a lambdified
reduced Lindblad RHS, parallel over eps columns in a spawn process pool.

TEST INFRASTRUCTURE / CPU BASELINE ONLY (see ``oracle/__init__.py``).  It restates
``Benchmarks/Benchmark_02_four_level_Interferometry.py:199-263`` (RHS preparation and the
per-point loop) and ``:348-380`` (the pool over columns, preparation time taken out of the
wall time) of the reference.  As there, the right-hand side comes from the package's own
SymPy front end (``build_reduced_lindblad_rhs``, the drive substituted into H) and is turned
into a NumPy function with ``sympy.lambdify``; each worker prepares it once.  The loop is the
reference's: k2 and k3 at t + dt/2, k4 at t + dt, the observable summed after every step with
index >= warmup and divided by ``max(num_steps - warmup, 1)``; a point whose state stops being
finite returns NaN.  Its results equal the C oracle's to rounding (``tests/test_oracle.py``).
"""
import multiprocessing as mp
import os
import queue
import time

import numpy as np

_W = {}


def prepare(model):
    """(rhs(t, px, py, *y) -> list, obs(t, px, py, *y) -> scalar, state size, rho0 index) for a
    benchmark model name of ``gqis.workloads``."""
    import sympy as sp
    from gqis import workloads
    from gqis.symbolic import build_reduced_lindblad_rhs
    w = getattr(workloads, model)(n_eps=1, n_amp=1)
    h = sp.Matrix(w["H"]).subs(w["Drive"])
    n = h.shape[0]
    eqs, mean_re, _, meta = build_reduced_lindblad_rhs(n, h, w["Col_Ops"], w["mean_operator"])
    u = [sp.Symbol(f"u{i}", real=True) for i in range(len(meta["rho_syms"]))]
    sub = dict(zip(meta["rho_syms"], u))
    px, py = list(w["var_arrays"])
    args = (workloads.T, px, py, *u)
    rhs = sp.lambdify(args, [e.subs(sub) for e in eqs], "numpy")
    obs = sp.lambdify(args, mean_re.subs(sub), "numpy")
    return rhs, obs, len(u)


def solve_point(rhs, obs, size, px, py, dt, num_steps, warmup=0, t0=0.0):
    """Time-averaged observable of one parameter point (reference loop, FP64)."""
    dt2, dt6 = 0.5 * dt, dt / 6.0
    y = np.zeros(size)
    y[0] = 1.0                       # |0><0|
    acc = 0.0
    with np.errstate(over="ignore", invalid="ignore"):
        for step in range(num_steps):
            t = t0 + step * dt
            k1 = np.asarray(rhs(t, px, py, *y), np.float64)
            k2 = np.asarray(rhs(t + dt2, px, py, *(y + dt2 * k1)), np.float64)
            k3 = np.asarray(rhs(t + dt2, px, py, *(y + dt2 * k2)), np.float64)
            k4 = np.asarray(rhs(t + dt, px, py, *(y + dt * k3)), np.float64)
            y = y + dt6 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
            if not np.all(np.isfinite(y)):
                return np.nan
            if step >= warmup:
                acc += float(obs(t + dt, px, py, *y))
    return acc / max(num_steps - warmup, 1)


def _init(model, times):
    start = time.perf_counter()
    _W["fns"] = prepare(model)
    times.put(time.perf_counter() - start)


def _column(task):
    j, px, pys, dt, num_steps, warmup = task
    rhs, obs, size = _W["fns"]
    return j, [solve_point(rhs, obs, size, px, py, dt, num_steps, warmup) for py in pys]


def sweep(model, px_values, py_values, dt, num_steps, *, warmup=0, workers=None):
    """The map over (px, py) in a spawn pool, one task per px column.  Returns (result (nx, ny),
    compute seconds = wall time minus the slowest worker's preparation, workers)."""
    workers = workers or max(1, (os.cpu_count() or 2) - 1)
    ctx = mp.get_context("spawn")
    times = ctx.Queue()
    out = np.empty((len(px_values), len(py_values)))
    start = time.perf_counter()
    with ctx.Pool(workers, initializer=_init, initargs=(model, times)) as pool:
        tasks = [(j, float(px), [float(v) for v in py_values], dt, num_steps, warmup)
                 for j, px in enumerate(px_values)]
        for j, col in pool.imap_unordered(_column, tasks):
            out[j] = col
    wall = time.perf_counter() - start
    prep = []
    while True:
        try:
            prep.append(times.get_nowait())
        except queue.Empty:
            break
    return out, max(0.0, wall - (max(prep) if prep else 0.0)), workers
