"""Edge cases of the newer kernel forms on the GPU: a second axis of length one, a single point (a
work-group of 64 with 63 idle lanes), a single time step -- each against the one-lane-per-point
kernel or the oracle."""
import numpy as np
import pytest

from gqis import mesolve_2D, workloads
from oracle import integrate, models

pytestmark = pytest.mark.gpu


def _rel(got, ref):
    return float(np.max(np.abs(got - ref)) / max(1e-300, float(np.max(np.abs(ref)))))


def test_matrix_core_forms_with_a_second_axis_of_length_one(gpu):
    # the Hilbert-space kernels (FP64 matrix cores, FP32 and FP16 hi/lo products) on a 5 x 1
    # grid (two points per wavefront at N = 8 in FP64: a ragged last wavefront), values equal
    # to the per-point FP64 kernel
    w = workloads.random_n_level(8, n_eps=5, n_amp=1, n_t=101, t_end=1.0)
    e, a = list(w["var_arrays"])
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    var = {e: np.linspace(-2.0, 2.0, 5), a: np.array([1.5])}
    ref = mesolve_2D(*args, var_arrays=var, fp64=True, liouville="lanes")
    for lv, fp64, tol in (("mfma", True, 1e-12), ("mfma", False, 2e-5), ("mfma_f16", False, 2e-5)):
        got = mesolve_2D(*args, var_arrays=var, fp64=fp64, liouville=lv)
        assert got.shape == (5, 1)
        assert _rel(got, ref) <= tol


@pytest.mark.parametrize("solver,groups", [("rk4", 8), ("dp5", 4), ("tsit5a", 2)])
def test_grouped_kernels_with_a_single_point(gpu, solver, groups):
    w = workloads.chain(8, n_eps=1, n_amp=1, n_t=201, t_end=5.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    var = {k: np.array([v[0] + 0.3]) for k, v in w["var_arrays"].items()}
    kw = dict(var_arrays=var, fp64=True, solver=solver, liouville="lanes",
              output_mode="final_rho")
    ref = mesolve_2D(*args, lanes_per_point=1, **kw)
    got = mesolve_2D(*args, lanes_per_point=groups, **kw)
    assert got.shape == (1, 1, 63)
    assert _rel(got, ref) <= (1e-7 if solver == "tsit5a" else 1e-13)


@pytest.mark.parametrize("liouville,fp64", [("mfma", True), ("lanes", True)])
def test_a_single_time_step(gpu, liouville, fp64):
    # tlist of two samples: one RK4 step, the mean over that one right endpoint
    n = 8
    w = workloads.random_n_level(n, seed=5, n_eps=3, n_amp=2, n_t=2, t_end=0.01)
    eps_v, amp_v = w["var_arrays"].values()
    got = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], fp64=fp64, liouville=liouville)
    m = models.random_n_level(n, 5, eps_v, amp_v)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.01, 1, obs=m.obs)
    assert np.max(np.abs(got.ravel() - ref)) <= 1e-12


@pytest.mark.parametrize("solver,groups", [("dp5", 4), ("ab5", 2), ("rk4", 8), ("tsit5a", 4)])
def test_ode_api_with_lanes_per_point(gpu, solver, groups):
    # a 32-state ODE (a ring of driven, damped oscillators) through odesolve_2D: the grouped
    # kernels split its rows over several wavefronts and give the one-lane result
    import sympy as sp
    from gqis import odesolve_2D
    n = 16
    xs = sp.symbols(f"x0:{n}", real=True)
    vs = sp.symbols(f"v0:{n}", real=True)
    k, g, t = sp.symbols("k g t", real=True)
    rhs = list(vs) + [-(2 + k) * xs[i] + xs[(i - 1) % n] + xs[(i + 1) % n] - g * vs[i]
                      + (sp.cos(t) if i == 0 else 0) for i in range(n)]
    y0 = [1.0 if i == 3 else 0.0 for i in range(n)] + [0.0] * n
    axes = {k: np.linspace(0.5, 1.5, 5), g: np.linspace(0.05, 0.2, 3)}
    kw = dict(var_arrays=axes, solver=solver, fp64=True)
    if solver == "tsit5a":
        kw["adaptive_options"] = {"step_rtol": 1e-9, "step_atol": 1e-11, "max_substeps": 4000}
    tl = np.linspace(0.0, 4.0, 401)
    ref = odesolve_2D(rhs, list(xs) + list(vs), y0, tl, lanes_per_point=1, **kw)
    got = odesolve_2D(rhs, list(xs) + list(vs), y0, tl, lanes_per_point=groups, **kw)
    assert got.shape == ref.shape == (5, 3, 2 * n)
    assert _rel(got, ref) <= (1e-8 if solver == "tsit5a" else 1e-13)
