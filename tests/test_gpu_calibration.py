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
