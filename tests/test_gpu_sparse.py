"""The wavefront-per-point sparse form (liouville="sparse", kernels/hilbert_sparse.hip) on the
reference's multi-qubit chain (Benchmark 04): FP64 against the FP64 oracle to 1e-10 of the
map's scale, FP32 to 2e-5, every output mode, a time trace, initial states, ragged grids."""
import numpy as np
import pytest

from gqis import mesolve_2D, workloads
from oracle import integrate, models

pytestmark = pytest.mark.gpu


def _model(nq, steps, n_eps=2, n_amp=2):
    w = workloads.multi_qubit(nq, n_eps=n_eps, n_amp=n_amp, n_t=steps + 1, t_end=0.01 * steps)
    eps_v, amp_v = w["var_arrays"].values()
    return w, models.multi_qubit(nq, eps_v, amp_v)


def _args(w):
    return (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])


@pytest.mark.parametrize("nq,steps", [(3, 200), (4, 100), (5, 40)])
def test_sparse_form_matches_the_oracle(gpu, nq, steps):
    n = 2 ** nq
    w, m = _model(nq, steps)
    print(f"sparse form: {nq} qubits, {steps} steps", flush=True)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, obs=m.obs)
    fin, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, mode="final_rho")
    scale = max(1.0, float(np.max(np.abs(ref))))
    for fp64, tol in ((True, 1e-10), (False, 2e-5)):
        got, info = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse",
                               fp64=fp64, return_timing_info=True)
        assert info["kernel"].startswith("gq_hilbert_sparse")
        err = float(np.max(np.abs(got.ravel() - ref)))
        print(f"  fp64={fp64}: mean max abs error {err:.2e}", flush=True)
        assert err <= tol * scale
        rho = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse", fp64=fp64,
                         output_mode="final_rho")
        assert rho.shape == (2, 2, n * n - 1)
        assert np.max(np.abs(rho.reshape(4, -1) - fin)) <= tol
        last = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse", fp64=fp64,
                          output_mode="final")
        ref_last, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, obs=m.obs, mode="final")
        assert np.max(np.abs(last.ravel() - ref_last)) <= tol * scale


def test_sparse_form_trace_warmup_and_initial_states(gpu):
    # 4 qubits on a ragged 3 x 5 grid: observable trace every 7th step, warmup, and a run
    # continued from the final states of a first one equals one run over both intervals
    nq, steps = 4, 60
    w, m = _model(nq, steps, n_eps=3, n_amp=5)
    res, tr, times = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse",
                                fp64=True, return_time_trace=True, time_trace_every=7,
                                warmup_time=0.25)
    ref, ref_tr = integrate.rk4(m.f, m.ground_state(), 0.01, steps, obs=m.obs, warmup_steps=15,
                                trace_stride=7, n_trace=9)
    assert res.shape == (3, 5) and tr.shape == (3, 5, 9)
    assert np.max(np.abs(res.ravel() - ref)) <= 1e-10
    assert np.max(np.abs(tr.reshape(15, 9) - ref_tr)) <= 1e-10
    half = workloads.multi_qubit(nq, n_eps=3, n_amp=5, n_t=31, t_end=0.3)
    mid = mesolve_2D(*_args(half), var_arrays=half["var_arrays"], liouville="sparse", fp64=True,
                     output_mode="final_rho")
    end = mesolve_2D(*_args(half), var_arrays=half["var_arrays"], liouville="sparse", fp64=True,
                     output_mode="final_rho", rho0_values=mid, t_in=0.3)
    whole = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse", fp64=True,
                       output_mode="final_rho")
    assert np.max(np.abs(end - whole)) <= 1e-12


def test_sparse_and_matrix_core_forms_agree(gpu):
    # 5 qubits, FP64: the ELL form and the dense FP64 matrix-core form of the same sweep
    w, _ = _model(5, 30, n_eps=3, n_amp=2)
    a = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse", fp64=True)
    b = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="mfma", fp64=True)
    assert np.max(np.abs(a - b)) <= 1e-12


def test_sparse_table_form_and_a_dense_model(gpu, monkeypatch):
    # the ELL / jump-table form (no XOR structure): the multi-qubit chain with the XOR form
    # switched off, and a random dense 8-level model with single-entry jumps, against the oracle
    from gqis import hilbert
    monkeypatch.setattr(hilbert, "SPARSE_XOR", False)
    w, m = _model(4, 50)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.01, 50, obs=m.obs)
    got = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse", fp64=True)
    assert np.max(np.abs(got.ravel() - ref)) <= 1e-10
    w = workloads.random_n_level(8, seed=2024, n_eps=3, n_amp=2, n_t=101, t_end=1.0)
    eps_v, amp_v = w["var_arrays"].values()
    rm = models.random_n_level(8, 2024, eps_v, amp_v)
    ref, _ = integrate.rk4(rm.f, rm.ground_state(), 0.01, 100, obs=rm.obs)
    got = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], liouville="sparse", fp64=True)
    assert np.max(np.abs(got.ravel() - ref)) <= 1e-10 * max(1.0, float(np.max(np.abs(ref))))


FIXED_OTHER = ("tsit5", "dop853", "lserk4", "lserk4f", "dp5", "alshina6", "anas5", "ab5")


@pytest.mark.parametrize("solver", FIXED_OTHER)
def test_sparse_form_other_fixed_step_methods(gpu, solver):
    # every fixed-step method on the sparse form against the FP64 oracle's same method: the
    # 3-qubit chain in FP64 (1e-10) and FP32 (2e-5), 60 steps with a warmup, the mean and
    # the final density matrix
    nq, steps = 3, 60
    w, m = _model(nq, steps, n_eps=3, n_amp=2)
    kw = dict(var_arrays=w["var_arrays"], liouville="sparse", solver=solver,
              solver_frequency=1.14)
    ref, _ = integrate.fixed_step(solver, m.f, m.ground_state(), 0.01, steps, obs=m.obs,
                                  warmup_steps=15, frequency=1.14)
    fin, _ = integrate.fixed_step(solver, m.f, m.ground_state(), 0.01, steps, mode="final_rho",
                                  frequency=1.14)
    print(f"sparse form: {solver}, {nq} qubits, {steps} steps", flush=True)
    for fp64, tol in ((True, 1e-10), (False, 2e-5)):
        got, info = mesolve_2D(*_args(w), fp64=fp64, warmup_time=0.25, return_timing_info=True,
                               **kw)
        assert info["kernel"].startswith("gq_hilbert_sparse")
        assert np.max(np.abs(got.ravel() - ref)) <= tol
        rho = mesolve_2D(*_args(w), fp64=fp64, output_mode="final_rho", **kw)
        assert np.max(np.abs(rho.reshape(6, -1) - fin)) <= tol


@pytest.mark.parametrize("solver", ("tsit5", "dop853", "lserk4"))
def test_sparse_form_five_qubits_other_methods(gpu, solver):
    # the published 5-qubit rows' methods at their full state (1023 reduced states), FP64 and
    # FP32 against the oracle over 20 steps on a 2 x 2 grid; auto picks the sparse form
    nq, steps = 5, 20
    w, m = _model(nq, steps)
    ref, _ = integrate.fixed_step(solver, m.f, m.ground_state(), 0.01, steps, obs=m.obs)
    scale = max(1.0, float(np.max(np.abs(ref))))
    print(f"sparse form: {solver}, {nq} qubits, {steps} steps", flush=True)
    for fp64, tol in ((True, 1e-10), (False, 2e-5)):
        got, info = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], solver=solver, fp64=fp64,
                               return_timing_info=True)
        assert info["kernel"].startswith("gq_hilbert_sparse")
        assert np.max(np.abs(got.ravel() - ref)) <= tol * scale


@pytest.mark.parametrize("solver", ["tsit5a", "dop853a"])
def test_sparse_form_adaptive_methods(gpu, solver):
    # adaptive Tsit5 / DOP853 on the sparse form, 3-qubit chain on a 3 x 2 grid, 40 output
    # intervals of 0.1: FP64 against the FP64 oracle within 1e-10 with the oracle's accepted and
    # rejected step counts on every point; FP32 at the reference's default tolerances within
    # 2e-5 (a CPU emulation with the RHS rounded to FP32 stays within 2e-7)
    w = workloads.multi_qubit(3, n_eps=3, n_amp=2, n_t=41, t_end=4.0)
    eps_v, amp_v = w["var_arrays"].values()
    m = models.multi_qubit(3, eps_v, amp_v)
    opts = {"step_rtol": 1e-6, "step_atol": 1e-8, "max_substeps": 1000, "count_steps": True}
    kw = dict(var_arrays=w["var_arrays"], liouville="sparse", solver=solver)
    fn = integrate.tsit5a if solver == "tsit5a" else integrate.dop853a
    print(f"sparse form: {solver}, 3 qubits, 40 output intervals", flush=True)
    if solver == "tsit5a":
        ref, _, acc, rej = fn(m.f, m.ground_state(), 0.1, 40, rtol=1e-6, atol=1e-8,
                              max_substeps=1000, obs=m.obs, mode="mean", warmup_steps=10)
        got, info = mesolve_2D(*_args(w), fp64=True, warmup_time=0.25, adaptive_options=opts,
                               return_timing_info=True, **kw)
    else:
        ref, _, acc, rej = fn(m.f, m.ground_state(), 0.1, 40, rtol=1e-6, atol=1e-8,
                              max_substeps=1000, obs=m.obs, mode="final")
        got, info = mesolve_2D(*_args(w), fp64=True, output_mode="final", adaptive_options=opts,
                               return_timing_info=True, **kw)
    assert info["kernel"].startswith("gq_hilbert_sparse")
    np.testing.assert_array_equal(info["adaptive_steps"]["accepted"].reshape(-1), acc)
    np.testing.assert_array_equal(info["adaptive_steps"]["rejected"].reshape(-1), rej)
    assert np.max(np.abs(got.ravel() - ref)) <= 1e-10
    fin, _, _, _ = fn(m.f, m.ground_state(), 0.1, 40, rtol=1e-6, atol=1e-8, max_substeps=1000,
                      mode="final_rho")
    rho = mesolve_2D(*_args(w), fp64=True, output_mode="final_rho", adaptive_options=opts, **kw)
    assert np.max(np.abs(rho.reshape(6, -1) - fin)) <= 1e-10
    ref32, _, _, _ = fn(m.f, m.ground_state(), 0.1, 40, rtol=1e-5, atol=1e-5,
                        max_substeps=1000, mode="final_rho")
    rho32 = mesolve_2D(*_args(w), fp64=False, output_mode="final_rho",
                       adaptive_options={"step_rtol": 1e-5, "step_atol": 1e-5}, **kw)
    err = float(np.max(np.abs(rho32.reshape(6, -1) - ref32)))
    print(f"  FP32 final state: max abs error {err:.2e}", flush=True)
    assert err <= 2e-5


def test_sparse_form_adaptive_five_qubits_trace_and_overflow(gpu):
    # 5 qubits (1023 reduced states), automatic form: Tsit5A FP64 with a trace every 2nd
    # output point against the oracle; too few substeps make every point NaN and raise
    w = workloads.multi_qubit(5, n_eps=2, n_amp=2, n_t=11, t_end=1.0)
    eps_v, amp_v = w["var_arrays"].values()
    m = models.multi_qubit(5, eps_v, amp_v)
    opts = {"step_rtol": 1e-6, "step_atol": 1e-8, "max_substeps": 1000, "count_steps": True}
    ref, ref_tr, acc, rej = integrate.tsit5a(m.f, m.ground_state(), 0.1, 10, rtol=1e-6,
                                             atol=1e-8, max_substeps=1000, obs=m.obs,
                                             mode="final", trace_stride=2, n_trace=5)
    print("sparse form: tsit5a, 5 qubits, 10 output intervals", flush=True)
    got, tr, times, info = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], solver="tsit5a",
                                      fp64=True, output_mode="final", adaptive_options=opts,
                                      return_time_trace=True, time_trace_every=2,
                                      return_timing_info=True)
    assert info["kernel"].startswith("gq_hilbert_sparse")
    np.testing.assert_array_equal(info["adaptive_steps"]["accepted"].reshape(-1), acc)
    assert np.max(np.abs(got.ravel() - ref)) <= 1e-10
    assert np.max(np.abs(tr.reshape(4, 5) - ref_tr)) <= 1e-10
    tight = {"step_rtol": 1e-14, "step_atol": 1e-14, "max_substeps": 2}
    with pytest.raises(RuntimeError, match="Non-finite"):
        mesolve_2D(*_args(w), var_arrays=w["var_arrays"], solver="tsit5a", fp64=True,
                   adaptive_options=tight)
    res = mesolve_2D(*_args(w), var_arrays=w["var_arrays"], solver="tsit5a", fp64=True,
                     adaptive_options=tight, ignore_non_finite_output=True)
    assert np.all(np.isnan(res))
