"""Lanes-per-point (grouped) kernels on the GPU (RK4, the other fixed-step methods, Tsit5A and
DOP853A): the same results as the one-lane-per-point kernels for every output mode, FP64 parity with the oracle's golden fixture, continuation from
a saved state, ragged point counts, and the autotuner's choice among both forms."""
import os

import numpy as np
import pytest

from gqis import autotune, mesolve_2D, workloads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _chain8(n_eps=12, n_amp=10, n_t=401, t_end=20.0):
    w = workloads.chain(8, n_eps=n_eps, n_amp=n_amp, n_t=n_t, t_end=t_end)
    return (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"]), w["var_arrays"]


def _rel(got, ref):
    return float(np.max(np.abs(got - ref)) / max(1e-300, float(np.max(np.abs(ref)))))


@pytest.mark.parametrize("mode", ["mean", "final", "final_rho"])
@pytest.mark.parametrize("groups", [2, 4, 8, 16])
def test_grouped_matches_per_lane_fp64(gpu, mode, groups):
    args, var = _chain8()
    ref = mesolve_2D(*args, var_arrays=var, fp64=True, output_mode=mode, liouville="lanes",
                     lanes_per_point=1)
    got = mesolve_2D(*args, var_arrays=var, fp64=True, output_mode=mode, liouville="lanes",
                     lanes_per_point=groups)
    print(mode, groups, "bitwise" if np.array_equal(got, ref) else _rel(got, ref), flush=True)
    assert got.shape == ref.shape and got.dtype == ref.dtype
    assert _rel(got, ref) <= 1e-13


def test_grouped_traces_and_state_mean(gpu):
    args, var = _chain8(n_eps=9, n_amp=7)
    for kw in ({"output_mode": "mean", "return_time_trace": True, "time_trace_every": 7},
               {"output_mode": "final_rho", "return_time_trace": True, "time_trace_every": 50}):
        ref = mesolve_2D(*args, var_arrays=var, fp64=True, liouville="lanes", lanes_per_point=1,
                         **kw)
        got = mesolve_2D(*args, var_arrays=var, fp64=True, liouville="lanes",
                         lanes_per_point=4, **kw)
        for a, b in zip(got[:2], ref[:2]):
            assert a.shape == b.shape and _rel(a, b) <= 1e-13
        np.testing.assert_array_equal(got[2], ref[2])


def test_grouped_fp32_matches_per_lane(gpu):
    args, var = _chain8()
    ref = mesolve_2D(*args, var_arrays=var, fp64=False, liouville="lanes", lanes_per_point=1)
    got = mesolve_2D(*args, var_arrays=var, fp64=False, liouville="lanes", lanes_per_point=4)
    assert _rel(got, ref) <= 1e-5


@pytest.mark.parametrize("groups", [2, 4, 8, 16])
def test_grouped_random_8_level_golden_fixture(gpu, groups):
    g = np.load(os.path.join(GOLDEN, "c3_random_8_level_8x8_subset_rk4_mean_fp64.npz"))
    w = workloads.random_n_level(8, seed=int(g["seed"]), n_eps=8, n_amp=8)
    e, a = list(w["var_arrays"])
    got = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays={e: g["eps"], a: g["amp"]}, fp64=True, liouville="lanes",
                     lanes_per_point=groups)
    assert _rel(got, g["result"]) <= 1e-10


def test_grouped_continuation_and_ragged_grid(gpu):
    args, var = _chain8(n_eps=7, n_amp=9, n_t=201, t_end=10.0)    # 63 points: a ragged group
    e, a = list(var)
    whole = mesolve_2D(*args, var_arrays=var, fp64=True, output_mode="final_rho",
                       liouville="lanes", lanes_per_point=4)
    half = (args[0], args[1], args[2], args[3], np.linspace(0.0, 5.0, 101))
    first = mesolve_2D(*half, var_arrays=var, fp64=True, output_mode="final_rho",
                       liouville="lanes", lanes_per_point=4)
    second = mesolve_2D(*half, var_arrays=var, fp64=True, output_mode="final_rho",
                        liouville="lanes", lanes_per_point=4, rho0_values=first, t_in=5.0)
    assert whole.shape == (7, 9, 63)
    assert _rel(second, whole) <= 1e-12


def test_autotuner_considers_grouped_kernels(gpu):
    autotune.clear_memory()
    args, var = _chain8(n_eps=128, n_amp=128, n_t=201, t_end=10.0)
    ref = mesolve_2D(*args, var_arrays=var, fp64=True, liouville="lanes", lanes_per_point=1)
    got, info = mesolve_2D(*args, var_arrays=var, fp64=True, liouville="lanes",
                           block_size="auto", return_timing_info=True)
    print("autotune", info["lanes_per_point"], info["block_size"], info.get("autotune"),
          flush=True)
    assert {c[0] for c in info["autotune"]} == {1, 2, 4, 8, 16}
    assert _rel(got, ref) <= 1e-13


ADAPT = {"step_rtol": 1e-8, "step_atol": 1e-10, "max_substeps": 2000, "count_steps": True}


@pytest.mark.parametrize("solver", ["tsit5a", "dop853a"])
@pytest.mark.parametrize("groups", [2, 4, 8])
def test_grouped_adaptive_matches_per_lane(gpu, solver, groups):
    # The error norm is summed per wave and then over the waves, so a borderline
    # accept/reject decision can differ from the one-lane kernel's sequential sum; results
    # then differ at the level of the step tolerance.
    args, var = _chain8(n_eps=8, n_amp=8, n_t=101, t_end=5.0)
    kw = dict(var_arrays=var, fp64=True, solver=solver, liouville="lanes",
              adaptive_options=ADAPT, return_timing_info=True)
    for mode in ("mean", "final_rho"):
        ref, ri = mesolve_2D(*args, output_mode=mode, lanes_per_point=1, **kw)
        got, gi = mesolve_2D(*args, output_mode=mode, lanes_per_point=groups, **kw)
        same = np.mean(ri["adaptive_steps"]["accepted"] == gi["adaptive_steps"]["accepted"])
        print(solver, groups, mode, _rel(got, ref), "same counts", same, flush=True)
        assert _rel(got, ref) <= 1e-7 and same >= 0.95


def test_grouped_adaptive_golden_duffing(gpu):
    g = np.load(os.path.join(GOLDEN, "c4_duffing_8x8_subset_t20_tsit5a_fp64.npz"))
    w = workloads.duffing(t_end=float(g["t_end"]), n_out=int(g["n_t"]))
    gs, xs = list(w["var_arrays"])
    from gqis import odesolve_2D
    res, info = odesolve_2D(w["rhs"], w["state_symbols"], w["y0"], w["tlist"],
                            var_arrays={gs: g["gamma"], xs: g["x0"]}, solver="tsit5a",
                            fp64=True, return_timing_info=True, lanes_per_point=2,
                            adaptive_options={"step_rtol": float(g["rtol"]),
                                              "step_atol": float(g["atol"]),
                                              "max_substeps": 1000, "count_steps": True})
    same = np.mean(info["adaptive_steps"]["accepted"] == g["accepted"])
    print("duffing grouped", _rel(res, g["result"]), same, flush=True)
    assert same >= 0.9 and _rel(res, g["result"]) <= 1e-7


def test_grouped_adaptive_traces_and_overflow(gpu):
    args, var = _chain8(n_eps=5, n_amp=6, n_t=41, t_end=2.0)
    kw = dict(var_arrays=var, fp64=True, solver="tsit5a", liouville="lanes",
              adaptive_options=ADAPT)
    for mode, every in (("mean", 3), ("final_rho", 10)):
        ref = mesolve_2D(*args, output_mode=mode, return_time_trace=True,
                         time_trace_every=every, lanes_per_point=1, **kw)
        got = mesolve_2D(*args, output_mode=mode, return_time_trace=True,
                         time_trace_every=every, lanes_per_point=4, **kw)
        for a, b in zip(got[:2], ref[:2]):
            assert a.shape == b.shape and _rel(a, b) <= 1e-7
    tight = dict(kw, adaptive_options={"step_rtol": 1e-14, "step_atol": 1e-14,
                                       "max_substeps": 2})
    with pytest.raises(RuntimeError, match="Non-finite"):
        mesolve_2D(*args, lanes_per_point=4, **tight)
    res = mesolve_2D(*args, lanes_per_point=4, ignore_non_finite_output=True, **tight)
    assert np.all(np.isnan(res))


def test_autotuner_considers_grouped_adaptive_kernels(gpu):
    autotune.clear_memory()
    args, var = _chain8(n_eps=128, n_amp=128, n_t=101, t_end=5.0)
    kw = dict(var_arrays=var, fp64=True, solver="tsit5a", liouville="lanes",
              adaptive_options={"step_rtol": 1e-8, "step_atol": 1e-10, "max_substeps": 2000})
    ref = mesolve_2D(*args, lanes_per_point=1, **kw)
    got, info = mesolve_2D(*args, block_size="auto", return_timing_info=True, **kw)
    print("autotune adaptive", info["lanes_per_point"], info["block_size"],
          info.get("autotune"), flush=True)
    assert {c[0] for c in info["autotune"]} == {1, 2, 4, 8, 16}
    assert _rel(got, ref) <= 1e-7


def test_auto_lanes_per_point_default(gpu):
    # 63 states of the tridiagonal chain (5 coefficients per row): on a small grid the default
    # spreads each point over 4 lanes (16 FP64 rows per wavefront); on a grid that gives every
    # SIMD a wavefront, FP32 stays on one lane.  15 states stay on one lane; a dense 8-level H
    # in FP64 takes the matrix-core path
    args, var = _chain8(n_eps=8, n_amp=8, n_t=101, t_end=5.0)
    _, info = mesolve_2D(*args, var_arrays=var, fp64=True, return_timing_info=True)
    assert info["lanes_per_point"] == 4 and info["kernel"] == "gq_rk4_grouped_f64"
    args, var = _chain8(n_eps=256, n_amp=256, n_t=11, t_end=0.25)
    _, info = mesolve_2D(*args, var_arrays=var, return_timing_info=True)
    assert info["lanes_per_point"] == 1 and info["kernel"] == "gq_rk4_f32"
    w = workloads.random_n_level(8, n_eps=8, n_amp=8, n_t=101, t_end=1.0)
    _, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                         var_arrays=w["var_arrays"], fp64=True, return_timing_info=True)
    assert info["kernel"] == "gq_hilbert_rk4_dmfma_f64"
    w = workloads.four_level(n_eps=8, n_amp=8, n_t=101, t_end=5.0)
    _, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                         var_arrays=w["var_arrays"], fp64=True, return_timing_info=True)
    assert info["lanes_per_point"] == 1


TABLEAU = ("dp5", "tsit5", "dop853", "alshina6", "anas5", "lserk4", "lserk4f", "ab5")


@pytest.mark.parametrize("solver", TABLEAU)
def test_grouped_tableau_matches_per_lane(gpu, solver):
    # 63 points (a ragged group of 64) at 63 states; the grouped kernel evaluates the same
    # expressions in the same order as the one-lane kernel.  AB5 is unstable for the chain's
    # nearly imaginary eigenvalues at dt = 0.025 (rounding then grows without bound), so it
    # takes a five times smaller step.
    args, var = _chain8(n_eps=9, n_amp=7, n_t=2001 if solver == "ab5" else 401, t_end=10.0)
    for mode, groups in (("mean", 8), ("final_rho", 4), ("final", 2)):
        ref = mesolve_2D(*args, var_arrays=var, fp64=True, solver=solver, output_mode=mode,
                         liouville="lanes", lanes_per_point=1)
        got = mesolve_2D(*args, var_arrays=var, fp64=True, solver=solver, output_mode=mode,
                         liouville="lanes", lanes_per_point=groups)
        print(solver, mode, groups, "bitwise" if np.array_equal(got, ref) else _rel(got, ref),
              flush=True)
        assert got.shape == ref.shape and np.max(np.abs(ref)) < 10.0     # a stable case
        assert _rel(got, ref) <= 1e-13


@pytest.mark.parametrize("solver", ["dp5", "lserk4", "ab5"])
def test_grouped_tableau_traces_and_lane_time(gpu, solver):
    # observable and state traces; stage times evaluated per lane instead of from the table
    args, var = _chain8(n_eps=5, n_amp=6, n_t=1001 if solver == "ab5" else 201, t_end=5.0)
    for kw in ({"output_mode": "mean", "return_time_trace": True, "time_trace_every": 7},
               {"output_mode": "final_rho", "return_time_trace": True, "time_trace_every": 25},
               {"output_mode": "final_rho", "time_table": False}):
        ref = mesolve_2D(*args, var_arrays=var, fp64=True, solver=solver, liouville="lanes",
                         lanes_per_point=1, **kw)
        got = mesolve_2D(*args, var_arrays=var, fp64=True, solver=solver, liouville="lanes",
                         lanes_per_point=4, **kw)
        if not isinstance(ref, tuple):
            ref, got = (ref,), (got,)
        for a, b in zip(got[:2], ref[:2]):
            assert a.shape == b.shape and np.max(np.abs(b)) < 10.0 and _rel(a, b) <= 1e-13
        if len(ref) > 2:
            np.testing.assert_array_equal(got[2], ref[2])


def test_grouped_tableau_fp32(gpu):
    args, var = _chain8(n_eps=8, n_amp=8, n_t=201, t_end=5.0)
    ref = mesolve_2D(*args, var_arrays=var, fp64=False, solver="dp5", liouville="lanes",
                     lanes_per_point=1)
    got = mesolve_2D(*args, var_arrays=var, fp64=False, solver="dp5", liouville="lanes",
                     lanes_per_point=4)
    assert _rel(got, ref) <= 1e-5
