"""Drop-in API behaviour on the CPU: shapes, dtypes, trace times, caches, option groups and
validation (mirrors the reference's CPU-safe tests, tests/test_ode_api.py and
tests/test_package.py).  The device call is replaced by a recording fake that returns zero
arrays of the right layout -- test infrastructure only; the product has no CPU path."""
import numpy as np
import pytest
import sympy as sp

import gqis.device as device
import gqis.solver as core
from gqis import compile_lindblad_rhs, compile_ode_rhs, mesolve_2D, odesolve_2D


@pytest.fixture
def fake(monkeypatch):
    import torch
    sources = []

    def run_sweep(*, source, compiled, mode, var_arrays, num_trace, state_trace, fp64, spec,
                  **kw):
        sources.append(source)
        axes = list(var_arrays.values())
        nx = len(np.atleast_1d(axes[0]))
        ny = len(np.atleast_1d(axes[1])) if len(axes) > 1 else 1
        d = compiled.state_size
        rdt = torch.float64 if fp64 else torch.float32
        cdt = torch.complex128 if fp64 else torch.complex64
        res = (torch.zeros((nx, ny, d), dtype=rdt) if mode in ("final_rho", "mean_state")
               else torch.zeros((nx, ny), dtype=cdt))
        tr = None
        if num_trace:
            tr = (torch.zeros((nx, ny, num_trace, d), dtype=rdt) if state_trace
                  else torch.zeros((nx, ny, num_trace), dtype=cdt))
        steps = np.zeros((nx, ny, 2), np.int32) if spec.options.get("count_steps") else None
        return {"result": res, "trace": tr, "status": 0, "gpu_s": 0.0, "build_s": 0.0,
                "cached_kernel": len(sources) > 1 and sources[-1] in sources[:-1],
                "setup_s": 0.0, "adaptive_steps": steps}
    monkeypatch.setattr(device, "run_sweep", run_sweep)
    for cache in (core._RHS_CODE_CACHE, core._KERNEL_CACHE):
        cache.clear()
    yield sources
    for cache in (core._RHS_CODE_CACHE, core._KERNEL_CACHE):
        cache.clear()


X, Y, T, A = sp.symbols("x y t a", real=True)


def solve(**kw):
    return odesolve_2D([Y, -A * X + sp.sin(T)], [X, Y], [1, 0], np.linspace(0, 1, 9),
                       var_arrays={A: np.array([1.0, 2.0])}, **kw)


def test_shapes_and_trace_times(fake):
    final, trace, times = solve(return_time_trace=True, time_trace_every=3)
    assert final.shape == (2, 1, 2) and trace.shape == (2, 1, 3, 2)
    np.testing.assert_allclose(times, [0.125, 0.5, 0.875])
    code = fake[-1]
    assert "#define GQ_D 2" in code
    rhs = code.split("void gq_rhs(")[1].split("}")[0]
    assert "sin(" not in rhs          # time terms are evaluated per stage, not per RHS


def test_time_offset_shifts_trace_and_reuses_kernel(fake):
    _, _, times = solve(t_in=3.0, return_time_trace=True)
    np.testing.assert_allclose(times, 3 + np.arange(1, 9) / 8)
    solve(t_in=5.0, return_time_trace=True)
    assert fake[0] == fake[1]


def test_observable_fp64_dtypes(fake):
    res, trace, times = solve(observable=(X + Y) ** 2 + sp.I * (X + Y) ** 3 + T, fp64=True,
                              return_time_trace=True)
    assert res.dtype == np.complex128 and trace.shape == (2, 1, 8)
    assert times.dtype == np.float64
    res32, _, times32 = solve(observable=X, return_time_trace=True)
    assert res32.dtype == np.complex64 and times32.dtype == np.float32


def test_mean_state_and_lindblad_shapes(fake):
    assert solve(output_mode="mean").shape == (2, 1, 2)
    res = mesolve_2D(sp.zeros(2), {}, [], sp.diag(1, 0), [0, 0.1], var_arrays={A: np.array([0.0])})
    assert res.shape == (1, 1) and res.dtype == np.complex64
    rho = mesolve_2D(sp.zeros(3), {}, [], sp.eye(3), [0, 0.1], var_arrays={A: np.zeros(4)},
                     output_mode="final_rho")
    assert rho.shape == (4, 1, 8)


def test_static_precomputation_and_drive_stage(fake):
    odesolve_2D([sp.exp(A) * X], [X], [1], [0, 0.1], var_arrays={A: np.array([1.0])})
    statics = fake[-1].split("void gq_statics(")[1].split("}")[0]
    assert "exp(px)" in statics


def test_compiled_rhs_equivalence(fake):
    compiled = compile_ode_rhs([Y, -A * X + sp.sin(T)], [X, Y], [1, 0], [A])
    assert set(compiled.timings) == {"ode_prepare_s", "ode_optimize_s"}
    assert fake == []
    via = odesolve_2D(compiled, tlist=np.linspace(0, 1, 9), var_arrays={A: np.array([1.0, 2.0])})
    assert via.shape == (2, 1, 2)
    solve()
    assert fake[0] == fake[1]
    with pytest.raises(ValueError, match="var_arrays must sweep"):
        odesolve_2D(compiled, tlist=[0, 1], var_arrays={})
    with pytest.raises(ValueError, match="fixed when the RHS is compiled"):
        odesolve_2D(compiled, tlist=[0, 1], var_arrays={A: np.array([1.0])}, observable=X)
    with pytest.raises(ValueError, match="mesolve_2D"):
        odesolve_2D(compile_lindblad_rhs(sp.zeros(2), {}, [], sp.diag(1, 0), [A]),
                    tlist=[0, 1], var_arrays={A: np.array([0.0])})


def test_compiled_lindblad_matches_direct(fake):
    model = (sp.Matrix([[A, 1], [1, -A]]), {}, [sp.Matrix([[0, 1], [0, 0]])], sp.diag(1, 0))
    compiled = compile_lindblad_rhs(*model, [A])
    _, info = mesolve_2D(compiled, tlist=[0, 0.1], var_arrays={A: np.array([0.0])},
                         return_timing_info=True)
    mesolve_2D(*model, [0, 0.1], var_arrays={A: np.array([0.0])}, RHSreuse=False)
    assert fake[0] == fake[1]
    assert info["cached_rhs"] == "provided"
    assert info["ode_prepare_s"] == info["ode_optimize_s"] == 0.0
    with pytest.raises(ValueError, match="order given when the RHS was compiled"):
        mesolve_2D(compiled, tlist=[0, 0.1], var_arrays={X: np.array([0.0])})
    with pytest.raises(ValueError, match="do not pass the model"):
        mesolve_2D(compiled, {}, tlist=[0, 0.1], var_arrays={A: np.array([0.0])})


def test_timing_info_keys(fake):
    _, info = solve(return_timing_info=True)
    keys = {"rhs_stage_s", "gpu_kernel_s", "rhs_codegen_s", "total_s", "ode_prepare_s",
            "ode_optimize_s", "code_emit_s", "kernel_build_s", "cached_rhs", "cached_kernel",
            "solver", "solver_frequency", "unroll"}
    assert keys <= set(info)
    assert info["cached_rhs"] == "miss" and info["solver"] == "rk4"
    _, info2 = solve(solver="tsit5a", return_timing_info=True)
    assert info2["cached_rhs"] == "hit" and info2["rhs_codegen_s"] == 0.0


def test_grouped_options_match_flat(fake):
    flat = solve(solver="tsit5a", fp64=True, return_time_trace=True, time_trace_every=3)
    grouped = solve(solver_options={"solver": "tsit5a", "fp64": True},
                    output={"return_time_trace": True, "time_trace_every": 3})
    assert fake[0] == fake[1]
    for a, b in zip(flat, grouped):
        np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError, match="both directly"):
        solve(fp64=True, solver_options={"fp64": False})
    with pytest.raises(ValueError, match="Unknown codegen option"):
        solve(codegen={"fp64": True})
    with pytest.raises(ValueError, match="does not accept"):
        mesolve_2D(sp.zeros(2), {}, [], sp.diag(1, 0), [0, 0.1], var_arrays={A: np.array([0.0])},
                   initial={"y0_values": None})


def _minimal():
    s = sp.Symbol("sweep", real=True)
    return sp.zeros(2), sp.Integer(0), [], sp.eye(2), {s: np.array([0.0], dtype=np.float32)}


@pytest.mark.parametrize("tlist,kw,match", [
    ([0.0, 1.0, 2.5], {}, "uniformly spaced"),
    ([1.0, 2.0], {}, "begin at zero"),
    ([0.0, 1.0], {"solver": "not_a_solver"}, "Unsupported solver"),
    ([0.0, 1.0], {"unroll": "False"}, "unroll must be True or False"),
    ([0.0], {}, "at least two"),
    ([0.0, 1.0], {"warmup_time": 1.5}, "warmup_time"),
    ([0.0, 1.0], {"fp64": "half"}, "fp64 must be"),
    ([0.0, 1.0], {"fp64": "mixed", "solver": "tsit5a"}, "one-lane RK4"),
    ([0.0, 1.0], {"fp64": "mixed", "time_table": False}, "one-lane RK4"),
    ([0.0, 1.0], {"fp64": "mixed", "lanes_per_point": 2}, "one-lane RK4"),
])
def test_validation_before_launch(fake, tlist, kw, match):
    # reference tests/test_package.py:88-116
    h, d, c, o, sweep = _minimal()
    with pytest.raises(ValueError, match=match):
        mesolve_2D(h, d, c, o, tlist, var_arrays=sweep, **kw)
    assert fake == []


def test_mixed_form_emits_the_mixed_kernel(fake):
    h, d, c, o, sweep = _minimal()
    mesolve_2D(h, d, c, o, [0.0, 0.5, 1.0], var_arrays=sweep, fp64="mixed")
    src = fake[-1]
    assert "gq_rk4_mixed" in src and "#define GQ_TABLE_F64 1" in src
    assert "gq_stage_narrow" in src and "typedef float real;" in src


def test_invalid_ode_input(fake):
    with pytest.raises(ValueError, match="Unassigned"):
        odesolve_2D([A * X], [X], [1], [0, 1])
    with pytest.raises(ValueError, match="real"):
        odesolve_2D([sp.I * X], [X], [1], [0, 1])


def test_package_import_is_lazy():
    import subprocess
    import sys
    code = ("import sys, builtins\n"
            "orig = builtins.__import__\n"
            "def imp(name, *a, **k):\n"
            "    if name == 'torch' or name.startswith('torch.'):\n"
            "        raise ModuleNotFoundError('blocked', name='torch')\n"
            "    return orig(name, *a, **k)\n"
            "builtins.__import__ = imp\n"
            "import gqis\n"
            "assert 'gqis.solver' not in sys.modules\n")
    import os
    import gqis
    env = dict(os.environ)
    env["PYTHONPATH"] = os.path.dirname(os.path.dirname(gqis.__file__))
    proc = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env)
    assert proc.returncode == 0, proc.stderr


def test_prepare_sweep_argument_checks():
    from gqis import prepare_sweep, workloads
    w = workloads.lzs_two_level(n_eps=4, n_amp=4, n_t=11, t_end=1.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    with pytest.raises(ValueError, match="calibrate_timesteps cannot be combined"):
        prepare_sweep(*args, var_arrays=w["var_arrays"], calibrate_timesteps=True)
    with pytest.raises(ValueError, match="not available for a prepared sweep"):
        prepare_sweep(*args, var_arrays=w["var_arrays"], return_timing_info=True)
    with pytest.raises(ValueError, match="does not apply: N = 2"):
        prepare_sweep(*args, var_arrays=w["var_arrays"], liouville="mfma")
    with pytest.raises(ValueError, match="output_mode"):
        prepare_sweep(*args, var_arrays=w["var_arrays"], output_mode="mean_state")
