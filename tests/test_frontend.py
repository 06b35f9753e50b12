"""Symbolic front end and code generation, on the CPU (no device)."""
import re

import numpy as np
import pytest
import sympy as sp

from gqis import codegen, symbolic, workloads
from gqis.solvers import available_solvers, get_solver_spec, normalize_solver_name
from oracle import lindblad as lb


def test_independent_density_matrix_layout():
    # reference tests/test_package.py:55-60
    rho, meta = symbolic.build_independent_rho(3)
    assert rho.shape == (3, 3)
    assert len(meta["rho_syms"]) == 8
    assert meta["num_coherences"] == 3
    assert sp.simplify(sp.trace(rho) - 1) == 0
    assert rho == rho.H.applyfunc(lambda e: e)  # Hermitian


def test_reduced_lindblad_rhs_matches_two_level_bloch_equations():
    # reference tests/test_package.py:63-85
    delta, drive, g1, g2 = sp.symbols("delta drive gamma1 gamma2", real=True, nonnegative=True)
    sx, sz, sm = sp.Matrix([[0, 1], [1, 0]]), sp.Matrix([[1, 0], [0, -1]]), sp.Matrix([[0, 1], [0, 0]])
    eqs, mre, mim, meta = symbolic.build_reduced_lindblad_rhs(
        2, delta * sx / 2 + drive * sz / 2, [sp.sqrt(g1) * sm, sp.sqrt(g2) * sz], sp.diag(0, 1))
    r00, rre, rim = meta["rho_syms"]
    cd = g1 / 2 + 2 * g2
    expected = [g1 * (1 - r00) - delta * rim, drive * rim - cd * rre,
                delta * (r00 - sp.Rational(1, 2)) - drive * rre - cd * rim]
    assert all(sp.simplify(a - b) == 0 for a, b in zip(eqs, expected))
    assert sp.simplify(mre - (1 - r00)) == 0 and mim == 0


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (4, 2)])
def test_symbolic_reduction_matches_numeric_generator(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = (a + a.conj().T) / 2
    ops = [rng.standard_normal((n, n)) * 0.3 for _ in range(2)]
    obs = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))

    def sym(m):
        return sp.Matrix(n, n, lambda i, j: sp.nsimplify(complex(m[i, j]).real, rational=False)
                         + sp.I * sp.Float(complex(m[i, j]).imag))
    eqs, mre, mim, meta = symbolic.build_reduced_lindblad_rhs(n, sym(h), [sym(o) for o in ops],
                                                              sym(obs))
    states = meta["rho_syms"]
    m_num, b_num = lb.affine_generator(h, ops)
    y = rng.standard_normal(len(states))
    subs = dict(zip(states, y))
    got = np.array([float(e.subs(subs)) for e in eqs])
    np.testing.assert_allclose(got, m_num @ y + b_num, atol=1e-10)
    o0, o = lb.observable_form(obs)
    val = complex(o0 + o @ y)
    assert abs(float(mre.subs(subs)) - val.real) < 1e-10
    assert abs(float(mim.subs(subs)) - val.imag) < 1e-10


def _eval_plan(plan, px, py, t, y, consts=()):
    """Evaluate a plan numerically with SymPy (statics, atoms, stage, RHS)."""
    env = {symbolic.PARX: px, symbolic.PARY: py, symbolic.TIME: t}
    env.update(dict(zip(plan.const_syms, consts)))
    env.update(dict(zip(plan.states, y)))
    for sym, e in plan.statics:
        env[sym] = sp.sympify(e).xreplace(env)
    for sym, e in plan.time_atoms:
        env[sym] = sp.sympify(e).xreplace(env)
    for sym, e in plan.stage:
        env[sym] = sp.sympify(e).xreplace(env)
    for sym, e in plan.rhs_temps:
        env[sym] = sp.sympify(e).xreplace(env)
    out = [complex(sp.N(sp.sympify(e).xreplace(env))).real for e in plan.rhs_out]
    for sym, e in plan.obs_temps:
        env[sym] = sp.sympify(e).xreplace(env)
    obs = [float(sp.N(sp.sympify(e).xreplace(env))) for e in plan.obs_out]
    return np.array(out), obs


def test_plan_reproduces_the_equations():
    w = workloads.four_level(n_eps=2, n_amp=2)
    from gqis.solver import _compile_rhs
    c, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                        sweep_symbols=tuple(w["var_arrays"]))
    plan = c.plan
    assert plan.affine and plan.nnz == 66
    assert len(plan.time_atoms) == 1
    rng = np.random.default_rng(5)
    y = rng.standard_normal(15)
    for px, py, t in ((3.0, 7.0, 1.3), (-12.5, 29.0, 77.7)):
        got, obs = _eval_plan(plan, px, py, t, y)
        m = __import__("oracle.models", fromlist=["x"]).four_level_chain([px], [py])
        ref = m.f(t, y[:, None])[:, 0]
        np.testing.assert_allclose(got, ref, atol=1e-10)
        assert abs(obs[0] - y[0]) < 1e-12 and obs[1] == 0


def test_nonlinear_ode_plan_hoists_time_terms():
    w = workloads.duffing(n_gamma=2, n_x0=2)
    from gqis.gqis_ode_api import ode_model
    from gqis.solver import _compile_rhs
    model = ode_model(w["rhs"], w["state_symbols"], w["y0"], w["var_arrays"])
    c, _ = _compile_rhs(model.equations, {}, [], model.observable, sweep_symbols=tuple(model.sweeps),
                        const_values=model.constants, runtime_consts=model.runtime,
                        keep_symbolic_consts=model.keep, rho0=model.initial,
                        ode_size=model.size)
    assert not c.plan.affine
    assert [str(e) for _, e in c.plan.time_atoms] == ["cos(1.2*t)"]
    got, _ = _eval_plan(c.plan, 0.4, 1.5, 3.1, [0.7, -0.2])
    x, v = 0.7, -0.2
    np.testing.assert_allclose(got, [v, -0.2 * v + x - x ** 3 + 0.4 * np.cos(1.2 * 3.1)],
                               atol=1e-12)


def test_generated_source_prints_literals_in_kernel_precision():
    x, y, t, a = sp.symbols("x y t a", real=True)
    from gqis.gqis_ode_api import ode_model
    from gqis.solver import _compile_rhs, _emit_source
    model = ode_model([y, -sp.sqrt(2) * x - sp.pi * a * y + sp.E * sp.cos(2 * sp.pi * t)],
                      [x, y], [1, 0], {a: np.ones(1)})
    c, _ = _compile_rhs(model.equations, {}, [], model.observable,
                        sweep_symbols=tuple(model.sweeps), rho0=model.initial, ode_size=2)
    src = _emit_source(c, get_solver_spec("rk4"), "final_rho", False, False, False, False, True,
                       256, False)
    assert not re.search(r"\bM_\w+", src)            # no C macros such as M_SQRT2
    assert "R(1.4142135623730951)" in src
    assert "pow(" not in src.split("gq_rhs(")[1].split("}")[0]


def test_solver_names_and_aliases():
    # reference tests/test_cuda_solvers.py:10-18
    assert available_solvers()[:10] == ("rk4", "lserk4", "lserk4f", "dp5", "ab5", "anas5",
                                        "tsit5", "alshina6", "dop853", "tsit5a")
    assert normalize_solver_name("RK4") == "rk4"
    assert normalize_solver_name("RK45") == "dp5"
    assert normalize_solver_name("Anastassi-Simos") == "anas5"
    assert normalize_solver_name("Alshina") == "alshina6"
    assert normalize_solver_name("DOP-853") == "dop853"
    with pytest.raises(ValueError, match="Unsupported solver"):
        normalize_solver_name("not_a_solver")


def test_reported_rhs_counts_and_adaptive_options():
    # reference tests/test_cuda_solvers.py:35-42,62-69
    assert get_solver_spec("rk4").rhs_evaluations == 4
    assert get_solver_spec("lserk4").rhs_evaluations == 5
    for name in ("dp5", "anas5", "tsit5"):
        assert get_solver_spec(name, step_size=0.1).rhs_evaluations == 6
    assert get_solver_spec("ab5").rhs_evaluations == 1
    assert get_solver_spec("alshina6").rhs_evaluations == 7
    assert get_solver_spec("dop853").rhs_evaluations == 12
    spec = get_solver_spec("tsit5a", step_size=0.1, adaptive_options={"step_rtol": 1e-6})
    assert spec.parameters[:3] == (1e-6, 1e-5, 64.0)
    with pytest.raises(ValueError, match="only to solver='tsit5a'"):
        get_solver_spec("tsit5", step_size=0.1, adaptive_options={"step_rtol": 1e-6})
    with pytest.raises(ValueError, match="Unknown adaptive_options"):
        get_solver_spec("tsit5a", adaptive_options={"rtol": 1e-6})
    anas = get_solver_spec("anas5", frequency=1.14, step_size=0.01)
    assert len(anas.parameters) == 1


def test_hiprtc_compiles_generated_kernels_for_gfx950():
    from gqis import runtime
    from gqis.solver import _compile_rhs, _emit_source
    w = workloads.lzs_two_level(n_eps=2, n_amp=2)
    c, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                        sweep_symbols=tuple(w["var_arrays"]))
    for name, mode in (("rk4", "mean"), ("tsit5a", "mean"), ("dop853a", "final_rho")):
        for fp64 in (False, True):
            src = _emit_source(c, get_solver_spec(name, step_size=0.1), mode, False, True, False,
                               fp64, True, 256, False)
            assert runtime.compile_check(src, codegen.HIPRTC_BASE_OPTIONS) > 1000


def test_bare_time_in_an_equation_becomes_a_time_atom():
    x, t = sp.symbols("x t", real=True)
    from gqis.gqis_ode_api import ode_model
    from gqis.solver import _compile_rhs
    model = ode_model([t], [x], [0], None)
    c, _ = _compile_rhs(model.equations, {}, [], model.observable, sweep_symbols=tuple(model.sweeps),
                        rho0=model.initial, ode_size=1)
    assert [str(e) for _, e in c.plan.time_atoms] == ["t"]
    got, _ = _eval_plan(c.plan, 0.0, 0.0, 2.25, [0.3])
    assert got[0] == 2.25


_A, _DRIVE, _DUMMY = sp.symbols("A Drive dummy", real=True)
_RABI = (_DRIVE / 2 * sp.Matrix([[0, 1], [1, 0]]), _A, [], sp.diag(0, 1))


def test_installation_symbolic_and_code_generation_stages():
    # the reference's installation self-test (gqis/check_environment.py:284-301): three real
    # equations, Im(rho_01)' = A/2 at |0><0| for A = 2, three generated RHS lines
    from gqis import compile_lindblad_rhs
    from gqis.solver import _emit_source
    rhs = compile_lindblad_rhs(*_RABI, [_A, _DUMMY])
    at_ground = dict(zip(rhs.state_symbols, (1, 0, 0)))
    drive = rhs.drive_map[_DRIVE] if isinstance(rhs.drive_map, dict) else rhs.drive_map
    values = [float(sp.sympify(e).subs(_DRIVE, drive).subs(at_ground)
                    .subs({sp.Symbol("ParX", real=True): 2.0, _A: 2.0}))
              for e in rhs.equations]
    assert rhs.state_size == 3
    np.testing.assert_allclose(values, [0.0, 0.0, 1.0], rtol=0, atol=1e-12)
    src = _emit_source(rhs, get_solver_spec("rk4", step_size=0.01), "final", False, False,
                       False, False, True, 256, False)
    body = src[src.index("void gq_rhs("):]
    body = body[:body.index("\n}")]
    assert body.count("dy[") == 3
