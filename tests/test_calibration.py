"""calibrate_timesteps search and store on the CPU, with the sweep replaced by fakes whose
error law is known in closed form (the reference pins the same behaviour with the same kind of
fakes: tests/test_ode_api.py:175-335).  The expected step counts follow from the error laws:
e.g. error 3e4/n^4 with an RMS limit 1e-3 needs n >= 73.6, and the RK4 Richardson estimate
against the finest computed run accepts the coarsest bracket point above that."""
import inspect
import warnings
from unittest.mock import patch

import numpy as np
import pytest
import sympy as sp

import gqis.calibration as cal
import gqis.solver as core
from gqis import mesolve_2D, odesolve_2D

A = sp.Symbol("a")
REAL = core._solve_symbolic_system
SIGNATURE = inspect.signature(REAL)


def _shape(call):
    axes = list(call["var_arrays"].values())
    return (len(axes[0]), len(axes[1]) if len(axes) > 1 else 1)


def _run(value_of, calls=None, tlist=None, var=None, **kw):
    """Calibrated mesolve_2D call; uncalibrated sweeps return value_of(call) on every point."""
    def sweep(**call):
        if call.get("calibrate_timesteps"):
            return REAL(**call)
        SIGNATURE.bind(**call)          # the sweep receives only real solver arguments
        if calls is not None:
            calls.append(call)
        out = np.full(_shape(call), value_of(call), dtype=complex)
        return (out, {}) if call.get("return_timing_info") else out
    with patch.object(core, "_solve_symbolic_system", side_effect=sweep):
        return REAL(H=sp.eye(2), Drive=sp.Integer(0), Col_Ops=[], mean_operator=sp.eye(2),
                    tlist=np.linspace(0, 1, 17) if tlist is None else tlist,
                    var_arrays={A: np.ones(4)} if var is None else var, **kw)


def _steps(call):
    return len(call["tlist"]) - 1


@pytest.fixture(autouse=True)
def _clear_store():
    core._TIMESTEP_CALIBRATION_CACHE.clear()
    yield
    core._TIMESTEP_CALIBRATION_CACHE.clear()


def test_defaults_differ_for_lindblad_and_ode():
    lind, ode = cal.settings_for(True), cal.settings_for(True, ode=True)
    assert (lind["rms_atol"], lind["max_atol"]) == (1e-3, 1e-2)
    assert (ode["rms_atol"], ode["max_atol"]) == (1e-4, 1e-3)


def test_rejects_unknown_and_empty_settings():
    x, y, t = sp.symbols("x y t")
    for bad, msg in ((dict(rms_tol=1e-4), "Unknown calibrate_timesteps"),
                     (dict(rms_atol=0, max_atol=0), "at least one"),
                     (dict(on_fail="loud"), "on_fail"), (dict(bisection_steps=-1), "bisection"),
                     (dict(recalibrate_every=0), "recalibrate_every")):
        with pytest.raises(ValueError, match=msg):
            odesolve_2D([y, -A * x], [x, y], [1, 0], np.linspace(0, 1, 9),
                        var_arrays={A: np.array([1.0, 2.0])}, calibrate_timesteps=bad)
    with pytest.raises(TypeError):
        cal.settings_for(1e-3)


@pytest.mark.parametrize("start,bisection,expected", [
    (16, 0, 128),    # doubling only: 64 fails (RMS 1.2e-3), 128 passes
    (16, 2, 76),     # bracket [64, 128] -> 91 passes -> 76 passes
    (512, 2, 76),    # a passing start is halved until 64 fails, then bisected
    (512, 0, 512),   # without bisection the supplied grid is the coarsest used
])
def test_bisection_finds_coarsest_passing_step(start, bisection, expected):
    _, info = _run(lambda c: 0.3 + 3e4 / _steps(c) ** 4, tlist=np.linspace(0, 1, start + 1),
                   return_timing_info=True, output_mode="final",
                   calibrate_timesteps=dict(rms_atol=1e-3, max_atol=1e-2,
                                            bisection_steps=bisection))
    rep = info["timestep_calibration"]
    assert rep["steps"] == expected
    assert rep["dt"] == pytest.approx(1.0 / expected)
    assert rep["calibrated_this_call"] and rep["details"]["status"] == "converged"


def test_mean_uses_first_order_estimate():
    # a right-endpoint mean with error 1/n needs n = 1000 (first-order Richardson); the same
    # error treated with the RK4 order would accept 724 steps (true error 1.4e-3)
    lim = dict(rms_atol=1e-3, max_atol=1e-2)
    law = lambda c: 0.3 + 1.0 / _steps(c)              # noqa: E731
    _, mean = _run(law, return_timing_info=True, output_mode="mean", calibrate_timesteps=lim)
    core._TIMESTEP_CALIBRATION_CACHE.clear()
    _, final = _run(law, return_timing_info=True, output_mode="final", calibrate_timesteps=lim)
    assert mean["timestep_calibration"]["steps"] == 1024
    assert final["timestep_calibration"]["steps"] == 724


@pytest.mark.parametrize("start", [1e-5, 1e-2])
def test_adaptive_scales_both_tolerances(start):
    # error 7 tol: the limit 1e-3 is met for tol <= 1.43e-4; the search steps by 10
    def law(call):
        assert _steps(call) == 16                        # output times never change
        o = call["adaptive_options"]
        return 0.3 + 7.0 * max(o["step_rtol"], o["step_atol"])
    _, info = _run(law, solver="tsit5a", return_timing_info=True,
                   calibrate_timesteps=dict(rms_atol=1e-3, max_atol=1e-2),
                   adaptive_options=dict(step_rtol=start, step_atol=start / 5))
    rep = info["timestep_calibration"]
    assert rep["step_tol"] == pytest.approx(1e-4)
    assert rep["step_rtol"] == pytest.approx(1e-4)
    assert rep["step_atol"] == pytest.approx(2e-5)       # the rtol/atol ratio is kept


def test_dop853a_is_calibrated_like_tsit5a():
    law = lambda c: 0.3 + 7.0 * c["adaptive_options"]["step_rtol"]   # noqa: E731
    _, info = _run(law, solver="dop853a", return_timing_info=True,
                   calibrate_timesteps=dict(rms_atol=1e-3, max_atol=1e-2),
                   adaptive_options=dict(step_rtol=1e-2, step_atol=1e-2))
    assert info["timestep_calibration"]["step_tol"] == pytest.approx(1e-4)


def test_adaptive_stops_at_max_substeps():
    def law(call):   # tolerances below 1e-6 exceed the substep budget
        tol = call["adaptive_options"]["step_rtol"]
        return np.nan if tol < 0.99e-6 else 0.3 + 1e3 * tol
    with pytest.warns(RuntimeWarning, match="max_substeps"):
        _, info = _run(law, solver="tsit5a", return_timing_info=True,
                       calibrate_timesteps=dict(rms_atol=1e-3, max_atol=1e-2))
    rep = info["timestep_calibration"]
    assert rep["step_tol"] == pytest.approx(1e-6)
    assert rep["details"]["status"] == "max_substeps"


def test_unstable_runs_never_pass():
    def sweep(**call):   # fewer than 4 steps is non-finite
        return np.full((2, 2), np.nan) if _steps(call) < 4 else np.ones((2, 2))
    args = dict(var_arrays={A: np.arange(10.0)}, rho0_var_arrays=None, rho0_values=None,
                solver="rk4", fp64=False)
    st = cal.settings_for(dict(rms_atol=0, max_atol=0, rms_rtol=1e-3))
    with patch.object(core, "_solve_symbolic_system", sweep):
        info = cal.calibrate(args, st, base_steps=1, duration=1.0)
    assert info["steps"] == 4                            # inf <= rtol * inf must not pass


def test_precision_floor_and_limits_report():
    # the change stops shrinking from 64 steps on: precision floor, best record accepted
    law = lambda c: 0.3 + (3e4 / _steps(c) ** 4 if _steps(c) < 64 else 1e-2 * (-1) ** _steps(c).bit_length())  # noqa: E501,E731
    with pytest.warns(RuntimeWarning, match="does not converge"):
        _, info = _run(law, return_timing_info=True, output_mode="final",
                       calibrate_timesteps=dict(rms_atol=1e-6, max_atol=0))
    assert info["timestep_calibration"]["details"]["status"] == "precision_floor"
    with pytest.raises(RuntimeError, match="max_steps"):
        _run(lambda c: 0.3 + 1.0 / _steps(c), output_mode="final",
             calibrate_timesteps=dict(rms_atol=1e-9, max_atol=0, max_steps=64, on_fail="error"))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        _run(lambda c: 0.3 + 1.0 / _steps(c), output_mode="final",
             calibrate_timesteps=dict(rms_atol=1e-9, max_atol=0, max_steps=64, on_fail="ignore",
                                      reset=True))


def test_test_grid_keeps_axis_ends():
    args = dict(var_arrays={A: np.arange(10.0)}, rho0_var_arrays=None,
                rho0_values=np.zeros((10, 3)), solver="rk4", fp64=False)
    test, full = cal.undersampled(args, 4)
    np.testing.assert_array_equal(test["var_arrays"][A], [0, 4, 8, 9])
    assert test["rho0_values"].shape == (4, 3)
    assert not full
    assert cal.undersampled(args, 1)[1]
    b = sp.Symbol("b")
    two = dict(args, var_arrays={A: np.arange(10.0), b: list(range(5))},
               rho0_values=np.zeros((10, 5, 3)))
    test, full = cal.undersampled(two, 4)
    assert test["var_arrays"][b] == [0, 4] and test["rho0_values"].shape == (4, 2, 3)


def test_reuse_and_recalibration():
    law = lambda c: 0.3 + 1e-2 / _steps(c) ** 4        # noqa: E731
    with patch.object(cal, "calibrate", wraps=cal.calibrate) as spy:
        st = dict(rms_atol=1e-4, sample_stride=2)
        _run(law, calibrate_timesteps=st)
        _run(law, calibrate_timesteps=st)
        assert spy.call_count == 1                       # the stored step is reused
        _run(law, calibrate_timesteps=st, cse_batch_size=1, block_size=128)
        assert spy.call_count == 1                       # code generation does not recalibrate
        _run(law, calibrate_timesteps=dict(st, rms_atol=1e-5))
        assert spy.call_count == 2                       # tolerances do
        every = dict(st, recalibrate_every=2, max_atol=0)
        for _ in range(3):
            _run(law, calibrate_timesteps=every)
        assert spy.call_count == 4                       # calls 1 and 3 of this setting
        _run(law, calibrate_timesteps=dict(every, reset=True))
        assert spy.call_count == 5
        _run(law, calibrate_timesteps=st, const_values={sp.Symbol("w"): 1.0})
        assert spy.call_count == 6                       # a baked constant is part of the model


def test_recalibration_only_refines():
    st = dict(rms_atol=1e-3, max_atol=0, recalibrate_every=1, bisection_steps=0)
    calls = []
    _, first = _run(lambda c: 0.3 + 3e4 / _steps(c) ** 4, calls, return_timing_info=True,
                    output_mode="final", calibrate_timesteps=st)
    # the second calibration would accept the 16-step start, but the stored step stays finer
    _, second = _run(lambda c: 0.3, calls, return_timing_info=True, output_mode="final",
                     calibrate_timesteps=st)
    assert first["timestep_calibration"]["steps"] == 128
    assert second["timestep_calibration"]["steps"] == 128
    assert second["timestep_calibration"]["details"]["steps"] == 16


def test_wider_sweep_range_warns_once():
    law = lambda c: 0.3 + 1e-2 / _steps(c) ** 4        # noqa: E731
    st = dict(rms_atol=1e-4, sample_stride=2)
    _run(law, calibrate_timesteps=st, var={A: np.array([1.0, 2.0])})
    wider = {A: np.array([1.0, 3.0])}
    with pytest.warns(RuntimeWarning, match="beyond the range used for calibration"):
        _run(law, calibrate_timesteps=st, var=wider)
    with warnings.catch_warnings():
        warnings.simplefilter("error")                  # the second call must not warn again
        _run(law, calibrate_timesteps=st, var=wider)
        quiet = dict(st, range_warning=False, rms_atol=2e-4)
        _run(law, calibrate_timesteps=quiet, var={A: np.array([1.0, 2.0])})
        _run(law, calibrate_timesteps=quiet, var=wider)


@pytest.mark.parametrize("bisection", [0, 2])
def test_full_grid_result_is_reused(bisection):
    calls = []
    _run(lambda c: 0.3 + 1e-2 / _steps(c) ** 4, calls,
         calibrate_timesteps=dict(rms_atol=1e-4, sample_stride=1, bisection_steps=bisection))
    assert len(calls) == 2          # start and one refinement; the final run is not repeated


def test_full_grid_bisection_and_max_steps_key():
    calls = []
    law = lambda c: 0.3 + 1e-2 / _steps(c) ** 4        # noqa: E731
    _run(law, calls, calibrate_timesteps=dict(rms_atol=1e-4, sample_stride=1,
                                              full_grid_bisection=True))
    assert len(calls) > 2
    with patch.object(cal, "calibrate", wraps=cal.calibrate) as spy:
        _run(law, calibrate_timesteps=dict(rms_atol=1e-4, sample_stride=1))
        _run(law, calibrate_timesteps=dict(rms_atol=1e-4, sample_stride=1, max_steps=2 ** 10))
        assert spy.call_count == 2


def test_trace_stays_on_the_supplied_grid():
    calls = []
    _, info = _run(lambda c: 0.3 + 3e4 / _steps(c) ** 4, calls, return_timing_info=True,
                   output_mode="final", return_time_trace=True, time_trace_every=2,
                   calibrate_timesteps=dict(rms_atol=1e-3, max_atol=1e-2))
    final = calls[-1]
    assert info["timestep_calibration"]["steps"] == 128      # 16 x 8 covers the 76 needed
    assert _steps(final) == 128
    assert final["time_trace_every"] == 16 and final["_time_trace_offset"] == 7
    assert final["return_time_trace"] and final["calibrate_timesteps"] is None


def test_full_grid_instability_falls_back_to_doubling_result():
    def law(call):
        n, full = _steps(call), _shape(call)[0] == 10
        return np.nan if (full and n < 128) else 0.3 + 3e4 / n ** 4
    calls = []
    with pytest.warns(RuntimeWarning, match="non-finite values on the full grid"):
        _, info = _run(law, calls, var={A: np.arange(10.0)}, return_timing_info=True,
                       output_mode="final", calibrate_timesteps=dict(rms_atol=1e-3, max_atol=1e-2))
    assert info["timestep_calibration"]["steps"] == 128
    assert _steps(calls[-1]) == 128 and _steps(calls[-2]) == 76
    key_entry = next(iter(core._TIMESTEP_CALIBRATION_CACHE.values()))
    assert key_entry["value"] == pytest.approx(1.0 / 128)


def test_timings_line(capsys):
    _run(lambda c: 0.3 + 3e4 / _steps(c) ** 4, output_mode="final", timings=True,
         calibrate_timesteps=dict(rms_atol=1e-3, max_atol=1e-2))
    out = capsys.readouterr().out
    assert "calibrate_timesteps: 76 steps (converged), tested from 16 steps" in out
