"""calibrate_timesteps on the GPU: the calibrated sweep meets the requested accuracy against a
converged FP64 oracle, and the final run equals an uncalibrated sweep on the chosen grid."""
import numpy as np
import pytest
import sympy as sp

import gqis.solver as core
from gqis import mesolve_2D, odesolve_2D, workloads
from oracle import integrate, models

pytestmark = pytest.mark.gpu

# The search checks its Richardson estimate on every 4th point of each axis only; the full grid
# may hold a slightly worse point, so the accuracy check allows a factor of 3 on the limits.
MARGIN = 3.0


def test_rk4_calibration_meets_requested_accuracy(gpu):
    core._TIMESTEP_CALIBRATION_CACHE.clear()
    w = workloads.four_level(n_eps=16, n_amp=12, n_t=101, t_end=20.0)
    # narrower ranges than the benchmark grid: the 100-step start (dt 0.2) is unstable only in
    # the large-|eps| corner, and refinement then reaches the asymptotic RK4 regime
    eps_s, amp_s = list(w["var_arrays"])
    w["var_arrays"] = {eps_s: np.linspace(-3.0, 3.0, 16), amp_s: np.linspace(0.0, 6.0, 12)}
    lim = dict(rms_atol=1e-5, max_atol=1e-4, sample_stride=4)
    got, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                           var_arrays=w["var_arrays"], fp64=True, output_mode="final",
                           calibrate_timesteps=lim, return_timing_info=True)
    rep = info["timestep_calibration"]
    print("calibrated", rep["steps"], rep["details"]["status"], flush=True)
    assert rep["details"]["status"] == "converged" and rep["calibrated_this_call"]
    n = rep["steps"]
    assert 100 < n < 20000
    eps_v, amp_v = w["var_arrays"].values()
    m = models.four_level_chain(eps_v, amp_v)
    # a 4x finer RK4 run is converged far below the limits (error / 256)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 20.0 / (4 * n), 4 * n, obs=m.obs,
                           mode="final")
    err = np.abs(got.reshape(-1) - ref)
    assert float(np.sqrt(np.mean(err ** 2))) <= MARGIN * lim["rms_atol"]
    assert float(err.max()) <= MARGIN * lim["max_atol"]
    # the final run is an ordinary sweep on the chosen grid
    plain = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                       np.linspace(0.0, 20.0, n + 1), var_arrays=w["var_arrays"], fp64=True,
                       output_mode="final")
    np.testing.assert_array_equal(got, plain)
    # a second call reuses the stored step without a new search
    _, again = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                          var_arrays=w["var_arrays"], fp64=True, output_mode="final",
                          calibrate_timesteps=lim, return_timing_info=True)
    assert again["timestep_calibration"]["steps"] == n
    assert not again["timestep_calibration"]["calibrated_this_call"]


def test_tsit5a_calibration_meets_requested_accuracy(gpu):
    core._TIMESTEP_CALIBRATION_CACHE.clear()
    x, y, g = sp.symbols("x y gamma", real=True)
    t = sp.Symbol("t", real=True)
    eqs = [y, -0.2 * y - x - x ** 3 + g * sp.cos(1.2 * t)]
    var = {g: np.linspace(0.1, 0.5, 9)}
    tl = np.linspace(0.0, 10.0, 11)
    lim = dict(rms_atol=1e-5, max_atol=1e-4, sample_stride=2)
    opts = dict(step_rtol=1e-3, step_atol=1e-3, max_substeps=100000)
    got, info = odesolve_2D(eqs, [x, y], [1.0, 0.0], tl, var_arrays=var, solver="tsit5a",
                            fp64=True, output_mode="final", adaptive_options=opts,
                            calibrate_timesteps=lim, return_timing_info=True)
    rep = info["timestep_calibration"]
    print("calibrated step_tol", rep["step_tol"], rep["details"]["status"], flush=True)
    assert rep["details"]["status"] == "converged"
    assert rep["step_rtol"] == pytest.approx(rep["step_tol"])
    assert rep["step_atol"] == pytest.approx(rep["step_tol"])
    fine = odesolve_2D(eqs, [x, y], [1.0, 0.0], tl, var_arrays=var, solver="tsit5a", fp64=True,
                       output_mode="final",
                       adaptive_options=dict(opts, step_rtol=1e-12, step_atol=1e-12))
    err = np.abs(np.asarray(got) - np.asarray(fine))
    assert float(np.sqrt(np.mean(err ** 2))) <= MARGIN * lim["rms_atol"]
    assert float(err.max()) <= MARGIN * lim["max_atol"]


# ---- every output mode, traces and ODE outputs (the behaviours the reference checks on its
# GPU, tests/test_gpu_solver.py:131-266): a coarse start refined until the map meets the limits

EPS_Q, AMP_Q, DRV_Q, DELTA_Q = sp.symbols("eps A Drive Delta", real=True)
OMEGA_Q = 1.14
PERIOD_Q = 2.0 * np.pi / OMEGA_Q
LIMITS = dict(rms_atol=1e-3, max_atol=1e-2, sample_stride=2)


def _qubit(steps_per_period, periods=10, **kw):
    """Driven dissipative qubit on a 16 x 16 (eps, A) grid."""
    t = sp.Symbol("t", real=True)
    sz, sx, sm = sp.Matrix([[1, 0], [0, -1]]), sp.Matrix([[0, 1], [1, 0]]), sp.Matrix([[0, 1], [0, 0]])
    g1 = 0.05 / PERIOD_Q
    g2 = g1 / 2.0 + 0.04 / PERIOD_Q
    return mesolve_2D(0.5 * DELTA_Q * sx + 0.5 * DRV_Q * sz,
                      {DRV_Q: EPS_Q + AMP_Q * sp.sin(OMEGA_Q * t)},
                      [sp.sqrt(g1) * sm, sp.sqrt(g2) * sz], sp.diag(0, 1),
                      np.linspace(0.0, periods * PERIOD_Q, periods * steps_per_period + 1),
                      var_arrays={EPS_Q: np.linspace(-6 * OMEGA_Q, 6 * OMEGA_Q, 16),
                                  AMP_Q: np.linspace(0.0, 6 * OMEGA_Q, 16)},
                      const_values={DELTA_Q: 1.0}, **kw)


def _within(got, ref, rms, mx):
    err = np.abs(np.asarray(got) - np.asarray(ref))
    assert float(np.sqrt(np.mean(err ** 2))) <= rms and float(err.max()) <= mx


@pytest.mark.parametrize("solver", ["rk4", "tsit5a"])
@pytest.mark.parametrize("mode", ["mean", "final", "final_rho"])
def test_calibration_meets_limits_for_every_output(gpu, solver, mode):
    core._TIMESTEP_CALIBRATION_CACHE.clear()
    ref = _qubit(2048, output_mode=mode)
    got = _qubit(8, output_mode=mode, solver=solver, calibrate_timesteps=LIMITS)
    _within(got, ref, LIMITS["rms_atol"], LIMITS["max_atol"])


@pytest.mark.parametrize("solver", ["rk4", "tsit5a"])
def test_calibrated_trace_keeps_the_supplied_sample_times(gpu, solver):
    core._TIMESTEP_CALIBRATION_CACHE.clear()
    _, _, plain_times = _qubit(8, output_mode="final", solver=solver, return_time_trace=True,
                               time_trace_every=8, ignore_non_finite_output=True)
    res, tr, times, info = _qubit(8, output_mode="final", solver=solver, return_time_trace=True,
                                  time_trace_every=8, calibrate_timesteps=LIMITS,
                                  return_timing_info=True)
    np.testing.assert_allclose(times, plain_times, rtol=1e-6)
    np.testing.assert_allclose(times, (np.arange(len(times)) * 8 + 1) * PERIOD_Q / 8, rtol=1e-6)
    if solver == "rk4":
        assert info["timestep_calibration"]["steps"] > 80         # the grid was refined
    ref, ref_tr, ref_times = _qubit(2048, output_mode="final", return_time_trace=True)
    picked = np.searchsorted(ref_times, times - 1e-6 * PERIOD_Q)
    np.testing.assert_allclose(ref_times[picked], times, rtol=1e-5)
    _within(res, ref, LIMITS["rms_atol"], LIMITS["max_atol"])
    _within(tr, ref_tr[:, :, picked], LIMITS["rms_atol"], LIMITS["max_atol"])


@pytest.mark.parametrize("solver", ["rk4", "tsit5a"])
@pytest.mark.parametrize("output", ["final", "mean", "observable", "trace"])
def test_calibration_meets_limits_for_ode_outputs(gpu, solver, output):
    x, v, t, damp, amp = sp.symbols("x v t damping amplitude", real=True)
    axes = {damp: np.linspace(0.1, 0.5, 16), amp: np.linspace(0.1, 1.0, 16)}

    def run(steps, **kw):
        return odesolve_2D([v, -damp * v - x - 0.2 * x ** 3 + amp * sp.cos(1.3 * t)], [x, v],
                           [1.0, 0.0], np.linspace(0.0, 20.0, steps + 1), var_arrays=axes, **kw)
    lim = dict(rms_atol=1e-4, max_atol=1e-3, sample_stride=2)
    extra = {"observable": dict(observable=x ** 2 + v ** 2), "mean": dict(output_mode="mean"),
             "trace": dict(return_time_trace=True, time_trace_every=8), "final": {}}[output]
    core._TIMESTEP_CALIBRATION_CACHE.clear()
    if solver == "tsit5a" and output == "mean":
        # The tsit5a mean is the trapezoid rule over the solver's own steps (reference
        # gqis/cuda_solvers.py:731-908): for a smooth solution its quadrature error is not set
        # by the step tolerance, so the exact time average is not the target.  What calibration
        # promises is its Richardson test (order 1 in the tolerance): tightening the chosen
        # tolerance tenfold moves the mean by at most 0.9 x the limits (MARGIN: the search
        # checks every 2nd point only), and the final run is an ordinary run at that tolerance.
        got, info = run(64, solver=solver, calibrate_timesteps=lim, return_timing_info=True,
                        **extra)
        rep = info["timestep_calibration"]
        assert rep["details"]["status"] == "converged"
        chosen = {"step_rtol": rep["step_rtol"], "step_atol": rep["step_atol"]}
        plain = run(64, solver=solver, adaptive_options=chosen, **extra)
        np.testing.assert_array_equal(got, plain)
        tight = run(64, solver=solver, adaptive_options={"max_substeps": 100000, **{
            k: val / 10.0 for k, val in chosen.items()}}, **extra)
        _within(got, tight, 0.9 * MARGIN * lim["rms_atol"], 0.9 * MARGIN * lim["max_atol"])
        return
    got = run(64, solver=solver, calibrate_timesteps=lim, **extra)
    if output == "trace":
        _, tr, times = got
        _, fine_tr, fine_times = run(16384, return_time_trace=True)
        picked = np.searchsorted(fine_times, times - 1e-6)
        np.testing.assert_allclose(fine_times[picked], times, rtol=1e-5)
        _within(tr, fine_tr[:, :, picked], lim["rms_atol"], lim["max_atol"])
    elif output == "mean":
        # the exact time average of a fine trace (the initial state is the first sample)
        _, fine_tr, fine_times = run(16384, return_time_trace=True)
        states = np.concatenate([np.broadcast_to([1.0, 0.0], fine_tr.shape[:2] + (1, 2)),
                                 fine_tr], axis=2)
        ref = np.trapezoid(states, np.concatenate([[0.0], fine_times]), axis=2) / 20.0
        _within(got, ref, lim["rms_atol"], lim["max_atol"])
    else:
        _within(got, run(16384, **extra), lim["rms_atol"], lim["max_atol"])
