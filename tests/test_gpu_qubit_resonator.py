"""The reference's qubit-resonator benchmark model (drive inside a square root and a reciprocal:
per-point stage values at every new stage time) on the GPU against the FP64 oracle: RK4 in FP64
and FP32, adaptive Tsit5 with its accepted / rejected counts in FP64 and, in FP32, with the
integration error's size, two and three Fock states."""
import numpy as np
import pytest

from gqis import mesolve_2D, workloads
from oracle import integrate, models

pytestmark = pytest.mark.gpu


def _case(n_photon, periods=4, n_eps=5, n_amp=4):
    w = workloads.qubit_resonator(n_photon, n_eps=n_eps, n_amp=n_amp, periods=periods)
    eps, amp = w["var_arrays"].values()
    m = models.qubit_resonator(n_photon, eps, amp, w["params"])
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    return w, m, args


def _rel(got, ref):
    return float(np.max(np.abs(np.ravel(got) - np.ravel(ref))) / np.max(np.abs(ref)))


@pytest.mark.parametrize("n_photon", [2, 3])
def test_rk4_matches_oracle(gpu, n_photon):
    # 4 drive periods of 146 steps (the reference's published step density), mean of a
    w, m, args = _case(n_photon)
    steps = len(w["tlist"]) - 1
    dt = float(w["tlist"][1] - w["tlist"][0])
    ref, _ = integrate.rk4(m.f, m.ground_state(), dt, steps, obs=m.obs)
    got64 = mesolve_2D(*args, var_arrays=w["var_arrays"], fp64=True)
    assert got64.shape == (5, 4)
    assert _rel(got64, ref) <= 1e-10
    got32 = mesolve_2D(*args, var_arrays=w["var_arrays"])
    assert _rel(got32, ref) <= 1e-4


def test_tsit5a_matches_oracle(gpu):
    # four output points per period (the reference's adaptive benchmark), FP64 controller
    w, m, args = _case(2, periods=2)
    tl = np.linspace(0.0, float(w["tlist"][-1]), 9)
    opts = {"step_rtol": 1e-8, "step_atol": 1e-8, "max_substeps": 1000, "count_steps": True}
    got, info = mesolve_2D(*args[:4], tl, var_arrays=w["var_arrays"], fp64=True,
                           solver="tsit5a", adaptive_options=opts, return_timing_info=True)
    ref, _, acc, rej = integrate.tsit5a(m.f, m.ground_state(), float(tl[1]), 8, rtol=1e-8,
                                        atol=1e-8, max_substeps=1000, obs=m.obs, mode="mean")
    np.testing.assert_array_equal(info["adaptive_steps"]["accepted"].reshape(-1), acc)
    np.testing.assert_array_equal(info["adaptive_steps"]["rejected"].reshape(-1), rej)
    assert _rel(got, ref) <= 1e-10


@pytest.mark.parametrize("solver", ["rk4", "dp5", "tsit5a"])
def test_single_lds_copy_matches_oracle(gpu, solver):
    # 14 levels (195 states) in FP64: two LDS copies of the state of 64 points (2 x 99.8 KB) do
    # not fit a compute unit, so the 8-lanes-per-point kernels keep one copy and put a barrier
    # before each write
    from gqis.solver import grouped_buffers
    from gqis.solvers import get_solver_spec
    assert grouped_buffers(get_solver_spec(solver, step_size=0.1), 195, True, 8, 2) == 1
    w, m, args = _case(7, periods=1, n_eps=3, n_amp=2)
    steps = len(w["tlist"]) - 1
    dt = float(w["tlist"][1] - w["tlist"][0])
    if solver == "tsit5a":
        tl = np.linspace(0.0, float(w["tlist"][-1]), 5)
        opts = {"step_rtol": 1e-8, "step_atol": 1e-8, "max_substeps": 1000, "count_steps": True}
        got, info = mesolve_2D(*args[:4], tl, var_arrays=w["var_arrays"], fp64=True,
                               solver="tsit5a", adaptive_options=opts, lanes_per_point=8,
                               return_timing_info=True)
        ref, _, acc, rej = integrate.tsit5a(m.f, m.ground_state(), float(tl[1]), 4, rtol=1e-8,
                                            atol=1e-8, max_substeps=1000, obs=m.obs, mode="mean")
        np.testing.assert_array_equal(info["adaptive_steps"]["accepted"].reshape(-1), acc)
        np.testing.assert_array_equal(info["adaptive_steps"]["rejected"].reshape(-1), rej)
    else:
        got = mesolve_2D(*args, var_arrays=w["var_arrays"], fp64=True, solver=solver,
                         lanes_per_point=8)
        fn = (integrate.rk4 if solver == "rk4" else
              lambda *a, **k: integrate.fixed_step("dp5", *a, **k))
        ref, _ = fn(m.f, m.ground_state(), dt, steps, obs=m.obs)
    assert _rel(got, ref) <= 1e-10


def test_tsit5a_fp32_matches_oracle(gpu):
    # the FP32 kernel (FP32 right-hand side with the reciprocal of E, stage-time cos by FP32
    # polynomials, FP32 error sums) against the FP64 oracle at the same tolerance: its own
    # step sequence, so the check is the integration error's size (1e-4 of the map's scale for
    # rtol = atol = 1e-6) and the mean accepted steps (within 5% of the oracle's)
    w, m, args = _case(2, periods=2)
    tl = np.linspace(0.0, float(w["tlist"][-1]), 9)
    opts = {"step_rtol": 1e-6, "step_atol": 1e-6, "max_substeps": 1000, "count_steps": True}
    va = {k: v.astype(np.float32) for k, v in w["var_arrays"].items()}
    got, info = mesolve_2D(*args[:4], tl, var_arrays=va, solver="tsit5a", adaptive_options=opts,
                           return_timing_info=True)
    ref, _, acc, rej = integrate.tsit5a(m.f, m.ground_state(), float(tl[1]), 8, rtol=1e-6,
                                        atol=1e-6, max_substeps=1000, obs=m.obs, mode="mean")
    assert info["kernel"].endswith("f32")
    assert _rel(got, ref) <= 1e-4
    mean_acc = float(np.mean(info["adaptive_steps"]["accepted"]))
    assert abs(mean_acc - float(np.mean(acc))) <= 0.05 * float(np.mean(acc))
