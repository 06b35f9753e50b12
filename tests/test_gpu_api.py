"""API behaviour on the GPU (the reference's GPU checks of the public entry points,
tests/test_gpu_solver.py:74-128, tests/test_ode_api.py:40-92): a compiled RHS gives the same
result as the direct call, solvers share one symbolic RHS, forced unrolling is reported,
device outputs stay on the device, and a time offset shifts trace times without a new kernel."""
import numpy as np
import pytest
import sympy as sp

from gqis import compile_lindblad_rhs, compile_ode_rhs, mesolve_2D, odesolve_2D, workloads

pytestmark = pytest.mark.gpu


def _lzs(n=6, n_t=201, t_end=20.0):
    return workloads.lzs_two_level(n_eps=n, n_amp=n + 1, n_t=n_t, t_end=t_end)


def test_compiled_lindblad_rhs_equals_the_direct_call(gpu):
    w = _lzs()
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"])
    direct = mesolve_2D(*args, w["tlist"], var_arrays=w["var_arrays"], fp64=True)
    c = compile_lindblad_rhs(*args, list(w["var_arrays"]))
    via = mesolve_2D(c, tlist=w["tlist"], var_arrays=w["var_arrays"], fp64=True)
    np.testing.assert_array_equal(via, direct)


def test_compiled_ode_rhs_equals_the_direct_call(gpu):
    w = workloads.duffing(n_gamma=5, n_x0=4, t_end=5.0, n_out=51)
    direct = odesolve_2D(w["rhs"], w["state_symbols"], w["y0"], w["tlist"],
                         var_arrays=w["var_arrays"], fp64=True)
    c = compile_ode_rhs(w["rhs"], w["state_symbols"], w["y0"], list(w["var_arrays"]))
    via = odesolve_2D(c, tlist=w["tlist"], var_arrays=w["var_arrays"], fp64=True)
    np.testing.assert_array_equal(via, direct)


def test_solvers_share_one_symbolic_rhs_and_unroll_is_reported(gpu):
    w = _lzs()
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    _, a = mesolve_2D(*args, var_arrays=w["var_arrays"], return_timing_info=True)
    _, b = mesolve_2D(*args, var_arrays=w["var_arrays"], solver="dp5", return_timing_info=True)
    assert b["cached_rhs"] == "hit" and b["solver"] == "dp5"
    _, u = mesolve_2D(*args, var_arrays=w["var_arrays"], unroll=True, return_timing_info=True)
    assert u["unroll"] is True


def test_device_outputs_stay_on_the_device(gpu):
    import torch
    w = _lzs()
    res, tr, times = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                                w["tlist"], var_arrays=w["var_arrays"], return_device=True,
                                output_mode="final_rho", return_time_trace=True,
                                time_trace_every=50)
    assert isinstance(res, torch.Tensor) and res.is_cuda and tuple(res.shape) == (6, 7, 3)
    assert isinstance(tr, torch.Tensor) and tr.is_cuda
    # continuation from the device state, no host round trip
    nxt = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], return_device=True, output_mode="final_rho",
                     rho0_values=res, t_in=20.0)
    assert nxt.is_cuda and torch.isfinite(nxt).all()


def test_time_offset_shifts_traces_and_reuses_the_kernel(gpu):
    w = _lzs()
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    _, _, t0, i0 = mesolve_2D(*args, var_arrays=w["var_arrays"], return_time_trace=True,
                              time_trace_every=40, return_timing_info=True)
    _, _, t1, i1 = mesolve_2D(*args, var_arrays=w["var_arrays"], return_time_trace=True,
                              time_trace_every=40, return_timing_info=True, t_in=3.5)
    np.testing.assert_allclose(t1, t0 + 3.5, rtol=1e-6)
    assert i1["cached_kernel"] in ("hit", "disk")
