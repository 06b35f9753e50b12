"""FP64 integrators with the reference kernel's semantics (oracle; test infrastructure only).

All functions integrate many parameter points at once: the state ``y`` has shape ``(D, P)``
and ``f(t, y)`` returns ``dy/dt`` for every point at the absolute time ``t`` (a scalar for the
fixed-step methods, shape ``(P,)`` for the adaptive ones).  ``obs(t, y)`` returns the complex
observable of every point.

Semantics restated from the reference (no code shared):

* time: ``dt`` is the uniform ``tlist`` step; stage times come from the integer step index,
  ``t_mid = (step + 0.5) * dt`` and ``t_end = t_mid + dt/2``; every drive evaluation uses
  ``t_in + t`` (``gqis/cuda_solvers.py:470-517``, ``gqis/N_Level_Kernel.cu:55``);
* RK4 update with k1 of a step reused from the end of the previous one
  (``gqis/cuda_solvers.py:479-515``);
* outputs (``gqis/solver.py:2126-2193``): ``mean`` is the right-endpoint average of the
  observable over the steps with index >= warmup, times ``1/(num_steps - warmup)``;
  ``mean_state`` the same for the state (divided); ``final`` the observable of the final state
  with the final drives; ``final_rho`` the state; the trace stores after steps
  ``offset + k*stride``;
* ``tsit5a``: adaptive Tsit5 between output points (``gqis/cuda_solvers.py:705-908``) with
  its FSAL reuse, reject/restore, last-step snapping, RMS error norm, step factor limits and the
  trapezoid mean (``gqis/solver.py:2156-2176``).
"""
import numpy as np

MODES = ("mean", "final", "final_rho", "mean_state")


def _trace_slot(step, stride, offset, n_trace):
    if step < offset or (step - offset) % stride:
        return None
    k = (step - offset) // stride
    return k if k < n_trace else None


def rk4(f, y0, dt, num_steps, *, t_in=0.0, warmup_steps=0, obs=None, mode="mean",
        trace_stride=1, trace_offset=0, n_trace=0, trace_kind="obs"):
    """Fixed-step classical RK4 (``gqis/cuda_solvers.py:470-517``).

    Returns ``(result, trace)``; ``trace`` is None when ``n_trace == 0``.
    """
    y = np.array(y0, dtype=np.float64, copy=True)
    npts = y.shape[1]
    dt = float(dt)
    dt2 = dt * 0.5
    dt6 = dt / 6.0
    if mode == "mean" and warmup_steps >= num_steps:
        mode = "final"
    if mode == "mean_state" and warmup_steps >= num_steps:
        mode = "final_rho"
    avg = np.zeros(npts, complex)
    state_avg = np.zeros_like(y)
    trace = None
    if n_trace:
        shape = (npts, n_trace) if trace_kind == "obs" else (npts, n_trace, y.shape[0])
        trace = np.zeros(shape, complex if trace_kind == "obs" else np.float64)
    t_drive = 0.0
    k = f(t_in + 0.0, y)
    for step in range(num_steps):
        t_mid = (step + 0.5) * dt
        tmp = y + dt2 * k
        acc = k.copy()
        k = f(t_in + t_mid, tmp)
        tmp = y + dt2 * k
        acc = acc + 2.0 * k
        k = f(t_in + t_mid, tmp)
        tmp = y + dt * k
        acc = acc + 2.0 * k
        t_mid = t_mid + dt2
        t_drive = t_mid
        k = f(t_in + t_drive, tmp)
        acc = acc + k
        y = y + dt6 * acc
        if step + 1 < num_steps:
            k = f(t_in + t_drive, y)
        if step >= warmup_steps:
            if mode == "mean":
                avg = avg + obs(t_in + t_drive, y)
            elif mode == "mean_state":
                state_avg = state_avg + y
        if trace is not None:
            slot = _trace_slot(step, trace_stride, trace_offset, n_trace)
            if slot is not None:
                if trace_kind == "obs":
                    trace[:, slot] = obs(t_in + t_drive, y)
                else:
                    trace[:, slot, :] = y.T
    if mode == "mean":
        result = avg * (1.0 / (num_steps - warmup_steps))
    elif mode == "final":
        result = obs(t_in + t_drive, y)
    elif mode == "final_rho":
        result = y.T.copy()
    else:
        result = (state_avg / (num_steps - warmup_steps)).T.copy()
    return result, trace


# Tsitouras 5(4) coefficients (Ch. Tsitouras, Comput. Math. Appl. 62 (2011) 770-775), the
# values the reference lists at gqis/cuda_solvers.py:705-721.
TSIT5_C = (0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0)
TSIT5_A = (
    (),
    (0.161,),
    (-0.008480655492356989, 0.335480655492357),
    (2.8971530571054935, -6.359448489975075, 4.3622954328695815),
    (5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525),
    (5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
     -0.028269050394068383),
)
TSIT5_B = (0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742,
           -3.290069515436081, 2.324710524099774)
TSIT5_BTILDE = (-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
                0.015151515151515152)


def tsit5a(f, y0, dt_out, num_steps, *, rtol=1e-5, atol=1e-5, max_substeps=64, t_in=0.0,
           warmup_steps=0, obs=None, mode="final_rho", trace_stride=1, trace_offset=0,
           n_trace=0, trace_kind="obs"):
    """Adaptive Tsit5 between the output points (``gqis/cuda_solvers.py:731-908``), in FP64.

    ``f(t, y, idx)`` and ``obs(t, y, idx)`` receive per-point times and the indices of the points
    in ``y`` (the active subset). Returns ``(result, trace, accepted,
    rejected)``; a point that exceeds ``max_substeps`` in one output interval becomes NaN.
    """
    y = np.array(y0, dtype=np.float64, copy=True)
    d, npts = y.shape
    c, a, b, e = TSIT5_C, TSIT5_A, TSIT5_B, TSIT5_BTILDE
    if mode == "mean" and warmup_steps >= num_steps:
        mode = "final"
    if mode == "mean_state" and warmup_steps >= num_steps:
        mode = "final_rho"
    inv_avg_time = 1.0 / ((num_steps - warmup_steps) * dt_out)
    h = np.full(npts, float(dt_out))
    failed = np.zeros(npts, bool)
    accepted = np.zeros(npts, np.int64)
    rejected = np.zeros(npts, np.int64)
    t_last = np.zeros(npts)
    allp = np.arange(npts)
    k1 = f(np.full(npts, t_in), y, allp)
    avg = np.zeros(npts, complex)
    state_avg = np.zeros_like(y)
    obs_prev = obs(np.full(npts, t_in), y, allp) if mode == "mean" else None
    trace = None
    if n_trace:
        shape = (npts, n_trace) if trace_kind == "obs" else (npts, n_trace, d)
        trace = np.zeros(shape, complex if trace_kind == "obs" else np.float64)
    for step in range(num_steps):
        failed_before = failed.copy()
        t_start = step * dt_out
        tau = np.zeros(npts)
        substeps = np.zeros(npts, np.int64)
        active = ~failed
        while np.any(active):
            idx = np.nonzero(active)[0]
            remaining = dt_out - tau[idx]
            last = h[idx] * 1.01 >= remaining
            hs = np.where(last, remaining, h[idx])
            t = t_start + tau[idx]
            yp = y[:, idx]
            ks = [k1[:, idx]]
            for s in range(1, 6):
                ys = yp + hs * sum(a[s][j] * ks[j] for j in range(s))
                ks.append(f(t_in + t + c[s] * hs, ys, idx))
            ynew = yp + hs * sum(b[j] * ks[j] for j in range(6))
            t_end = t_in + t_start + tau[idx] + hs
            k7 = f(t_end, ynew, idx)
            err = hs * (sum(e[j] * ks[j] for j in range(6)) + e[6] * k7)
            scale = rtol * np.maximum(np.abs(yp), np.abs(ynew)) + atol
            with np.errstate(all="ignore"):
                err_norm = np.sqrt(np.sum((err / scale) ** 2, axis=0) / d)
                factor = 0.9 * err_norm ** -0.2
            factor = np.where(np.isnan(factor), 0.2, np.clip(factor, 0.2, 5.0))
            ok = err_norm <= 1.0
            acc_idx, rej_idx = idx[ok], idx[~ok]
            # accepted points
            y[:, acc_idx] = ynew[:, ok]
            k1[:, acc_idx] = k7[:, ok]
            accepted[acc_idx] += 1
            t_last[acc_idx] = t_end[ok]
            if mode == "mean" and ok.any():
                o = obs(t_end[ok], ynew[:, ok], acc_idx)
                if step >= warmup_steps:
                    avg[acc_idx] += 0.5 * hs[ok] * inv_avg_time * (o + obs_prev[acc_idx])
                obs_prev[acc_idx] = o
            elif mode == "mean_state" and ok.any() and step >= warmup_steps:
                state_avg[:, acc_idx] += (0.5 * hs[ok] * inv_avg_time) * (ynew[:, ok] + yp[:, ok])
            lo = last[ok]
            tau[acc_idx] = np.where(lo, dt_out, tau[acc_idx] + hs[ok])
            h[acc_idx] = np.where(lo, np.maximum(h[acc_idx], hs[ok] * factor[ok]), hs[ok] * factor[ok])
            # rejected points: restore y, recompute k1 at the step start
            rejected[rej_idx] += 1
            h[rej_idx] = hs[~ok] * factor[~ok]
            if rej_idx.size:
                k1[:, rej_idx] = f(t_in + t[~ok], y[:, rej_idx], rej_idx)
            substeps[idx] += 1
            over = idx[substeps[idx] > max_substeps]
            failed[over] = True
            y[:, over] = np.nan
            active = (~failed) & (tau < dt_out)
        if trace is not None:
            # A point that failed in an earlier interval stops there and stores nothing more.
            slot = _trace_slot(step, trace_stride, trace_offset, n_trace)
            live = ~failed_before
            if slot is not None and live.any():
                if trace_kind == "obs":
                    trace[live, slot] = obs(t_in + t_last[live], y[:, live], allp[live])
                else:
                    trace[live, slot, :] = y[:, live].T
        if failed.all():
            break
    if mode == "mean":
        result = np.where(failed, np.nan, avg)
    elif mode == "mean_state":
        result = np.where(failed[None, :], np.nan, state_avg).T.copy()
    elif mode == "final":
        result = obs(t_in + t_last, y, allp)
    else:
        result = y.T.copy()
    return result, trace, accepted, rejected


# DOP853 coefficients of Hairer, Norsett & Wanner (Solving ODEs I, 2nd ed., and the published
# DOP853 code): main formula, and the 5th- and 3rd-order error estimators.  The values are the
# ones SciPy ships in ``scipy/integrate/_ivp/dop853_coefficients.py`` (BSD); the oracle reads
# them from there so the table is not retyped.
def dop853_tableau():
    from scipy.integrate._ivp import dop853_coefficients as co
    n = co.N_STAGES
    return co.C[:n], co.A[:n, :n], co.B, co.E5[:n], co.E3[:n]


def dop853a(f, y0, dt_out, num_steps, *, rtol=1e-8, atol=1e-10, max_substeps=4096, t_in=0.0,
            obs=None, mode="final_rho", trace_stride=1, trace_offset=0, n_trace=0,
            trace_kind="obs", fac_min=1.0 / 3.0, fac_max=6.0, safety=0.9):
    """Adaptive DOP853 between output points (no reference counterpart; new method).

    Same interval structure as ``tsit5a`` (steps end on the ``tlist`` points, the last step
    of an interval snaps to it, rejections restore the state), with Hairer's combined
    5th/3rd-order error norm and a step factor ``clamp(safety * err**(-1/8), fac_min, fac_max)``.
    Only ``final``/``final_rho`` and traces are supported.
    """
    cc, aa, bb, e5, e3 = dop853_tableau()
    y = np.array(y0, dtype=np.float64, copy=True)
    d, npts = y.shape
    h = np.full(npts, float(dt_out))
    failed = np.zeros(npts, bool)
    accepted = np.zeros(npts, np.int64)
    rejected = np.zeros(npts, np.int64)
    t_last = np.zeros(npts)
    allp = np.arange(npts)
    k1 = f(np.full(npts, t_in), y, allp)
    trace = None
    if n_trace:
        shape = (npts, n_trace) if trace_kind == "obs" else (npts, n_trace, d)
        trace = np.zeros(shape, complex if trace_kind == "obs" else np.float64)
    for step in range(num_steps):
        failed_before = failed.copy()
        t_start = step * dt_out
        tau = np.zeros(npts)
        substeps = np.zeros(npts, np.int64)
        active = ~failed
        while np.any(active):
            idx = np.nonzero(active)[0]
            remaining = dt_out - tau[idx]
            last = h[idx] * 1.01 >= remaining
            hs = np.where(last, remaining, h[idx])
            t = t_start + tau[idx]
            yp = y[:, idx]
            ks = [k1[:, idx]]
            for s in range(1, 12):
                ys = yp + hs * sum(aa[s, j] * ks[j] for j in range(s) if aa[s, j] != 0.0)
                ks.append(f(t_in + t + cc[s] * hs, ys, idx))
            ynew = yp + hs * sum(bb[j] * ks[j] for j in range(12) if bb[j] != 0.0)
            t_end = t_in + t_start + tau[idx] + hs
            err5 = sum(e5[j] * ks[j] for j in range(12) if e5[j] != 0.0)
            err3 = sum(e3[j] * ks[j] for j in range(12) if e3[j] != 0.0)
            sk = atol + rtol * np.maximum(np.abs(yp), np.abs(ynew))
            with np.errstate(all="ignore"):
                s5 = np.sum((err5 / sk) ** 2, axis=0)
                s3 = np.sum((err3 / sk) ** 2, axis=0)
                deno = s5 + 0.01 * s3
                deno = np.where(deno > 0.0, deno, 1.0)
                err_norm = np.abs(hs) * s5 / np.sqrt(deno * d)
                factor = safety * err_norm ** (-1.0 / 8.0)
            factor = np.where(np.isnan(factor), fac_min, np.clip(factor, fac_min, fac_max))
            ok = err_norm <= 1.0
            acc_idx, rej_idx = idx[ok], idx[~ok]
            y[:, acc_idx] = ynew[:, ok]
            if acc_idx.size:
                k1[:, acc_idx] = f(t_end[ok], ynew[:, ok], acc_idx)
            accepted[acc_idx] += 1
            t_last[acc_idx] = t_end[ok]
            lo = last[ok]
            tau[acc_idx] = np.where(lo, dt_out, tau[acc_idx] + hs[ok])
            h[acc_idx] = np.where(lo, np.maximum(h[acc_idx], hs[ok] * factor[ok]), hs[ok] * factor[ok])
            rejected[rej_idx] += 1
            h[rej_idx] = hs[~ok] * factor[~ok]
            substeps[idx] += 1
            over = idx[substeps[idx] > max_substeps]
            failed[over] = True
            y[:, over] = np.nan
            active = (~failed) & (tau < dt_out)
        if trace is not None:
            slot = _trace_slot(step, trace_stride, trace_offset, n_trace)
            live = ~failed_before
            if slot is not None and live.any():
                if trace_kind == "obs":
                    trace[live, slot] = obs(t_in + t_last[live], y[:, live], allp[live])
                else:
                    trace[live, slot, :] = y[:, live].T
        if failed.all():
            break
    if mode == "final":
        result = obs(t_in + t_last, y, allp)
    else:
        result = y.T.copy()
    return result, trace, accepted, rejected


# ---------------------------------------------------------------------------------------------
# Other fixed-step methods (reference gqis/cuda_solvers.py:48-702): explicit tableau RK with the
# k1 of a step taken from the end of the previous one, the 2N-storage LSRK4 and Adams-Bashforth 5
# with RK4 start-up.  Stage times t = step*dt + c*dt; outputs after each step as in rk4().
def fixed_tableau(name, step_size=None, frequency=None):
    """(c, a, b) of a fixed-step method; sources: SciPy (DP5 = RK45, DOP853), Tsitouras 2011,
    Alshina et al. 2008 and Anastassi & Simos (fitted coefficient from w*dt)."""
    if name == "dp5":
        from scipy.integrate._ivp.rk import RK45
        return RK45.C, RK45.A, RK45.B
    if name == "dop853":
        c, a, b, _, _ = dop853_tableau()
        return c, a, b
    if name == "tsit5":
        a = np.zeros((6, 6))
        for s, row in enumerate(TSIT5_A):
            a[s, :len(row)] = row
        return np.array(TSIT5_C), a, np.array(TSIT5_B)
    if name == "alshina6":
        c = np.array([0.0, 4 / 7, 5 / 7, 6 / 7, (5 - np.sqrt(5)) / 10, (5 + np.sqrt(5)) / 10, 1.0])
        rows = [(), (4 / 7,), (115 / 112, -5 / 16), (589 / 630, 5 / 18, -16 / 45),
                (229 / 1200 - 29 * np.sqrt(5) / 6000, 119 / 240 - 187 * np.sqrt(5) / 1200,
                 -14 / 75 + 34 * np.sqrt(5) / 375, -3 * np.sqrt(5) / 100),
                (71 / 2400 - 587 * np.sqrt(5) / 12000, 187 / 480 - 391 * np.sqrt(5) / 2400,
                 -38 / 75 + 26 * np.sqrt(5) / 375, 27 / 80 - 3 * np.sqrt(5) / 400,
                 (1 + np.sqrt(5)) / 4),
                (-49 / 480 + 43 * np.sqrt(5) / 160, -425 / 96 + 51 * np.sqrt(5) / 32,
                 52 / 15 - 4 * np.sqrt(5) / 5, -27 / 16 + 3 * np.sqrt(5) / 16,
                 5 / 4 - 3 * np.sqrt(5) / 4, 5 / 2 - np.sqrt(5) / 2)]
        a = np.zeros((7, 7))
        for s, row in enumerate(rows):
            a[s, :len(row)] = row
        return c, a, np.array([1 / 12, 0, 0, 0, 5 / 12, 5 / 12, 1 / 12])
    if name == "anas5":
        p = anas5_parameter(frequency, step_size)
        c = np.array([0.0, 0.1, 1 / 3, 2 / 3, 0.9, 1.0])
        a = np.zeros((6, 6))
        a[1, 0] = 0.1
        a[2, :2] = (-2 / 9, 5 / 9)
        a[3, :3] = (28 / 9, -40 / 9, 2.0)
        a[4, :4] = (-11277 / 8000, 171 / 80, -459 / 2000, 3213 / 8000)
        a[5, :5] = (-4.0 - 0.595 * p, 5.0, 1.89 * p, -2.295 * p, p)
        return c, a, np.array([23 / 216, 0.0, 63 / 136, 9 / 56, 1000 / 3213, -1 / 24])
    raise ValueError(name)


def anas5_parameter(frequency, step_size):
    """Fitted Anas5 coefficient from v = w*dt (reference gqis/cuda_solvers.py:971-990)."""
    import mpmath as mp
    w = 1.0 if frequency is None else float(frequency)
    with mp.workdps(60):
        v = mp.mpf(str(w)) * mp.mpf(str(step_size))
        if v == 0:
            return float(mp.mpf(-800) / 1071)
        tn = mp.tan(v)
        num = -2 * v ** 5 + 6 * tn * v ** 4 + 24 * v ** 3 - 72 * tn * v ** 2 - 144 * v + 144 * tn
        den = v ** 5 * (2 * tn * v + 12 - 20)
        return float((-mp.mpf(8000) / 1071) * num / den)


LSRK4_A = (0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
           -3550918686646 / 2091501179385, -1275806237668 / 842570457699)
LSRK4_B = (1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
           1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
           2277821191437 / 14882151754819)
LSRK4_C = (0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896,
           2006345519317 / 3224310063776, 2802321613138 / 2924317926251)
AB5_BETA = (1901 / 720, -1387 / 360, 109 / 30, -637 / 360, 251 / 720)


def fixed_step(name, f, y0, dt, num_steps, *, t_in=0.0, warmup_steps=0, obs=None, mode="mean",
               trace_stride=1, trace_offset=0, n_trace=0, trace_kind="obs", frequency=None):
    """Any fixed-step solver name of the reference (rk4 goes through rk4())."""
    if name == "rk4":
        return rk4(f, y0, dt, num_steps, t_in=t_in, warmup_steps=warmup_steps, obs=obs, mode=mode,
                   trace_stride=trace_stride, trace_offset=trace_offset, n_trace=n_trace,
                   trace_kind=trace_kind)
    y = np.array(y0, dtype=np.float64, copy=True)
    npts = y.shape[1]
    if mode == "mean" and warmup_steps >= num_steps:
        mode = "final"
    if mode == "mean_state" and warmup_steps >= num_steps:
        mode = "final_rho"
    avg = np.zeros(npts, complex)
    state_avg = np.zeros_like(y)
    trace = None
    if n_trace:
        shape = (npts, n_trace) if trace_kind == "obs" else (npts, n_trace, y.shape[0])
        trace = np.zeros(shape, complex if trace_kind == "obs" else np.float64)
    t_end = 0.0
    hist = []
    if name not in ("lserk4", "lserk4f"):
        k1 = f(t_in, y)
        hist = [k1]
    if name in FIXED_RK:
        c, a, b = fixed_tableau(name, step_size=dt, frequency=frequency)
    for step in range(num_steps):
        t = step * dt
        t_end = t + dt
        if name in FIXED_RK:
            ks = [k1]
            for s in range(1, len(c)):
                ys = y + dt * sum(a[s, j] * ks[j] for j in range(s) if a[s, j] != 0.0)
                ks.append(f(t_in + t + c[s] * dt, ys))
            y = y + dt * sum(b[j] * ks[j] for j in range(len(c)) if b[j] != 0.0)
            if step + 1 < num_steps:
                k1 = f(t_in + t_end, y)
        elif name in ("lserk4", "lserk4f"):
            r = None
            for s in range(5):
                k = f(t_in + t + LSRK4_C[s] * dt, y) if s else f(t_in + (t if step else 0.0), y)
                r = dt * k if s == 0 else LSRK4_A[s] * r + dt * k
                y = y + LSRK4_B[s] * r
        elif name == "ab5":
            if step < 4:
                k = hist[0]
                tmp = y + 0.5 * dt * k
                acc = k.copy()
                k = f(t_in + t + 0.5 * dt, tmp)
                tmp = y + 0.5 * dt * k
                acc = acc + 2 * k
                k = f(t_in + t + 0.5 * dt, tmp)
                tmp = y + dt * k
                acc = acc + 2 * k
                k = f(t_in + t_end, tmp)
                acc = acc + k
                y = y + dt / 6 * acc
            else:
                y = y + dt * sum(beta * hist[j] for j, beta in enumerate(AB5_BETA))
            if step + 1 < num_steps:
                hist = [f(t_in + t_end, y)] + hist[:4]
        else:
            raise ValueError(name)
        if step >= warmup_steps:
            if mode == "mean":
                avg = avg + obs(t_in + t_end, y)
            elif mode == "mean_state":
                state_avg = state_avg + y
        if trace is not None:
            slot = _trace_slot(step, trace_stride, trace_offset, n_trace)
            if slot is not None:
                if trace_kind == "obs":
                    trace[:, slot] = obs(t_in + t_end, y)
                else:
                    trace[:, slot, :] = y.T
    if mode == "mean":
        result = avg * (1.0 / (num_steps - warmup_steps))
    elif mode == "final":
        result = obs(t_in + t_end, y)
    elif mode == "final_rho":
        result = y.T.copy()
    else:
        result = (state_avg / (num_steps - warmup_steps)).T.copy()
    return result, trace


FIXED_RK = ("dp5", "tsit5", "alshina6", "dop853", "anas5")
