"""prepare_sweep: buffers resident, the launch chain replayed as one HIP graph.  Every replay
is checked against an eager mesolve_2D call with the same inputs, with other GPU work and
synchronisations between replays."""
import time

import numpy as np
import pytest
import sympy as sp

from gqis import mesolve_2D, prepare_sweep, workloads

pytestmark = pytest.mark.gpu

EPS, AMP, DELTA, DRV, T = sp.symbols("eps A Delta Drive t", real=True)
SZ = sp.Matrix([[1, 0], [0, -1]])
SX = sp.Matrix([[0, 1], [1, 0]])
SM = sp.Matrix([[0, 1], [0, 0]])


def _model():
    h = EPS / 2 * SZ + DELTA / 2 * SX + DRV * SZ / 2
    col = [sp.sqrt(sp.Rational(1, 100)) * SM, sp.sqrt(sp.Rational(5, 1000)) * SZ]
    return h, {DRV: AMP * sp.cos(T)}, col, SZ


def _busy():
    import torch
    a = torch.randn(512, 512, device="cuda")
    (a @ a).sum().item()                       # unrelated eager work and a synchronisation


@pytest.mark.parametrize("solver", ["rk4", "tsit5a"])
def test_replays_match_eager_calls(gpu, solver):
    h, drive, col, obs = _model()
    tl = np.linspace(0.0, 20.0, 401)
    var = {EPS: np.linspace(-5, 5, 48), AMP: np.linspace(0, 8, 40)}
    kw = dict(fp64=True, solver=solver, runtime_consts={DELTA: 1.0})
    prep = prepare_sweep(h, drive, col, obs, tl, var_arrays=var, **kw)
    assert prep.shape == (48, 40)
    cases = [({}, {}),
             ({EPS: np.linspace(-2, 7, 48), AMP: np.linspace(1, 3, 40)}, {}),
             ({}, {DELTA: 0.6}),
             ({EPS: np.linspace(-5, 5, 48), AMP: np.linspace(0, 8, 40)}, {DELTA: 1.3})]
    state_var, state_consts = dict(var), {DELTA: 1.0}
    for new_var, new_consts in cases + cases[:2]:
        got = prep.run(var_arrays=new_var or None, runtime_consts=new_consts or None)
        state_var.update(new_var)
        state_consts.update(new_consts)
        ref = mesolve_2D(h, drive, col, obs, tl, var_arrays=state_var, fp64=True,
                         solver=solver, runtime_consts=state_consts)
        np.testing.assert_array_equal(got, ref)
        _busy()
    prep.close()


def test_state_continuation_and_traces(gpu):
    h, drive, col, obs = _model()
    tl = np.linspace(0.0, 5.0, 101)
    var = {EPS: np.linspace(-3, 3, 9), AMP: np.linspace(0, 4, 7)}
    first = mesolve_2D(h, drive, col, obs, tl, var_arrays=var, fp64=True, output_mode="final_rho",
                       runtime_consts={DELTA: 1.0})
    prep = prepare_sweep(h, drive, col, obs, tl, var_arrays=var, fp64=True,
                         output_mode="final_rho", runtime_consts={DELTA: 1.0},
                         rho0_values=first, t_in=5.0, return_time_trace=True,
                         time_trace_every=10)
    state = first
    for _ in range(3):
        got, tr, times = prep.run(rho0_values=state)
        ref, rtr, rtimes = mesolve_2D(h, drive, col, obs, tl, var_arrays=var, fp64=True,
                                      output_mode="final_rho", runtime_consts={DELTA: 1.0},
                                      rho0_values=state, t_in=5.0, return_time_trace=True,
                                      time_trace_every=10)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(tr, rtr)
        np.testing.assert_array_equal(times, rtimes)
        state = got
        _busy()
    with pytest.raises(ValueError, match="prepared sweep has 9"):
        prep.run(var_arrays={EPS: np.zeros(3)})


def test_grouped_and_nonfinite_replays(gpu):
    w = workloads.chain(8, n_eps=16, n_amp=16, n_t=201, t_end=5.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    prep = prepare_sweep(*args, var_arrays=w["var_arrays"], fp64=True, lanes_per_point=4,
                         liouville="lanes")
    ref = mesolve_2D(*args, var_arrays=w["var_arrays"], fp64=True, lanes_per_point=4,
                     liouville="lanes")
    for _ in range(3):
        np.testing.assert_array_equal(prep.run(), ref)
        _busy()
    e, a = list(w["var_arrays"])
    bad = {e: np.full(16, 1e30), a: np.linspace(0, 1, 16)}
    with pytest.raises(RuntimeError, match="Non-finite"):
        prep.run(var_arrays=bad)
    # the status word is cleared inside the graph: a good replay after a bad one is clean
    np.testing.assert_array_equal(prep.run(var_arrays=w["var_arrays"]), ref)


def test_replay_host_time_is_below_the_full_call(gpu):
    h, drive, col, obs = _model()
    tl = np.linspace(0.0, 2.0, 41)
    var = {EPS: np.linspace(-5, 5, 16), AMP: np.linspace(0, 8, 16)}
    kw = dict(fp64=True, runtime_consts={DELTA: 1.0})
    prep = prepare_sweep(h, drive, col, obs, tl, var_arrays=var, **kw)
    mesolve_2D(h, drive, col, obs, tl, var_arrays=var, **kw)
    t0 = time.perf_counter()
    for _ in range(20):
        mesolve_2D(h, drive, col, obs, tl, var_arrays=var, **kw)
    full = (time.perf_counter() - t0) / 20
    prep.run()
    t0 = time.perf_counter()
    for _ in range(20):
        prep.run()
    replay = (time.perf_counter() - t0) / 20
    print(f"per call: mesolve_2D {full * 1e3:.3f} ms, replay {replay * 1e3:.3f} ms", flush=True)
    assert replay < full


@pytest.mark.parametrize("n,fp64", [(8, True), (16, False)])
def test_matrix_core_sweep_replays(gpu, n, fp64):
    # the Hilbert-space kernels (FP64 matrix cores at N = 8, FP16 hi/lo products at N = 16)
    # behind prepare_sweep: new axes, a new drive frequency (a runtime constant inside the
    # tabulated time term) and explicit initial states, each replay equal to an eager call
    omega = sp.Symbol("omega", real=True)
    w = workloads.random_n_level(n, n_eps=6, n_amp=5, n_t=201, t_end=2.0)
    drive = {k: v.subs(sp.Symbol("t", real=True), omega * sp.Symbol("t", real=True))
             for k, v in w["Drive"].items()}
    args = (w["H"], drive, w["Col_Ops"], w["mean_operator"], w["tlist"])
    e, a = list(w["var_arrays"])
    var = dict(w["var_arrays"])
    kw = dict(fp64=fp64, runtime_consts={omega: 1.0})
    prep = prepare_sweep(*args, var_arrays=var, **kw)
    assert prep.plan["info"]["kernel"].startswith("gq_hilbert_rk4")
    state_var, state_consts = dict(var), {omega: 1.0}
    cases = [({}, {}), ({e: np.linspace(-2, 2, 6), a: np.linspace(1, 2, 5)}, {}),
             ({}, {omega: 1.7}), ({}, {})]
    for new_var, new_consts in cases:
        got = prep.run(var_arrays=new_var or None, runtime_consts=new_consts or None)
        state_var.update(new_var)
        state_consts.update(new_consts)
        ref = mesolve_2D(*args, var_arrays=state_var, fp64=fp64, runtime_consts=state_consts)
        np.testing.assert_array_equal(got, ref)
        _busy()
    prep.close()
    rho = mesolve_2D(*args, var_arrays=state_var, fp64=fp64, runtime_consts=state_consts,
                     output_mode="final_rho")
    prep = prepare_sweep(*args, var_arrays=state_var, output_mode="final_rho",
                         rho0_values=rho, **dict(kw, runtime_consts=state_consts))
    got = prep.run()
    ref = mesolve_2D(*args, var_arrays=state_var, fp64=fp64, runtime_consts=state_consts,
                     output_mode="final_rho", rho0_values=rho)
    np.testing.assert_array_equal(got, ref)
    got = prep.run(rho0_values=ref)
    ref2 = mesolve_2D(*args, var_arrays=state_var, fp64=fp64, runtime_consts=state_consts,
                      output_mode="final_rho", rho0_values=ref)
    np.testing.assert_array_equal(got, ref2)
    prep.close()
