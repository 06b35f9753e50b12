"""``odesolve_2D`` / ``compile_ode_rhs``: general real ODE sweeps on the same kernels.

Behaviour of the reference ODE front end (``gqis/gqis_ode_api.py:17-255``): states are renamed
to ``rho[i]``, parameters to ``ode_param_i``, the time symbol to the kernel time; ``final`` maps
to the full state, ``mean`` to the state mean; without an observable the result is the state.
"""
import dataclasses

import numpy as np
import sympy as sp

from ._options import OPTION_GROUPS, apply_option_groups
from .symbolic import TIME

_ODE_CODEGEN_DEFAULTS = {"pre_expand": False, "collect_rho": False, "cse_simplify": True}
_ODE_CODEGEN_OPTIONS = ("pre_expand", "collect_rho", "factor_terms", "cse_batch_size",
                        "cse_simplify", "hoist_rho_independent", "RHSreuse")
_DUMMY_SWEEP = sp.Symbol("ode_dummy", real=True)


def odesolve_2D(rhs, state_symbols=None, y0=None, tlist=None, *, var_arrays=None,
                const_values=None, time_symbol=None, observable=None, output_mode="final",
                y0_values=None, runtime_consts=None, keep_symbolic_consts=None, solver="rk4",
                fp64=False, return_time_trace=False, time_trace_every=None, t_in=0.0,
                return_device=False, calibrate_timesteps=None, consts=None, initial=None,
                output=None, solver_options=None, codegen=None, diagnostics=None, **option_set):
    """Real first-order ODEs over zero, one or two sweep axes (``gqis/gqis_ode_api.py:17-67``).

    Results: ``(nx, ny, nstate)`` real without an observable, ``(nx, ny)`` complex with one;
    traces add a sample axis before the state axis.  ``rhs`` may be a ``CompiledRHS`` from
    ``compile_ode_rhs``.  Other ``mesolve_2D`` options are forwarded through ``option_set``.
    """
    args = dict(locals())
    extra = args.pop("option_set")
    groups = {name: args.pop(name) for name in OPTION_GROUPS}
    args, extra = apply_option_groups(odesolve_2D, args, groups, extra)
    return _solve_ode_system(**args, **extra)


def compile_ode_rhs(rhs, state_symbols, y0=None, sweep_symbols=(), *, const_values=None,
                    runtime_consts=None, keep_symbolic_consts=None, time_symbol=None,
                    observable=None, **codegen_options):
    """Prepare and optimise an ODE RHS without the GPU (``gqis/gqis_ode_api.py:70-112``)."""
    unknown = set(codegen_options) - set(_ODE_CODEGEN_OPTIONS)
    if unknown:
        raise TypeError(f"Unknown compile_ode_rhs option(s): {sorted(unknown)}")
    sweep_symbols = tuple(sweep_symbols)
    if len(sweep_symbols) > 2:
        raise ValueError("At most two sweep symbols are supported.")
    model = _prepare_ode_model(rhs, state_symbols, y0, dict.fromkeys(sweep_symbols), const_values,
                               runtime_consts, keep_symbolic_consts, time_symbol, observable)
    from .solver import _compile_rhs
    options = {**_ODE_CODEGEN_DEFAULTS, **codegen_options}
    reuse = options.pop("RHSreuse", True)
    compiled, _ = _compile_rhs(
        model["H"], {}, [], model["mean"], sweep_symbols=tuple(model["sweeps"]),
        const_values=model["constants"], runtime_consts=model["runtime"],
        keep_symbolic_consts=model["keep"], rho0=model["initial"], ode_size=model["size"],
        reuse=reuse, **options)
    return dataclasses.replace(compiled, ode={"symbol_map": model["replacements"],
                                              "sweep_symbols": sweep_symbols,
                                              "has_observable": observable is not None})


def _prepare_ode_model(rhs, state_symbols, y0, var_arrays, const_values, runtime_consts,
                       keep_symbolic_consts, time_symbol, observable):
    """Validate an ODE model and rename it to kernel symbols
    (``gqis/gqis_ode_api.py:115-181``)."""
    states = tuple(state_symbols)
    if (not states or any(not isinstance(s, sp.Symbol) for s in states)
            or len(set(states)) != len(states)):
        raise ValueError("state_symbols must contain distinct SymPy Symbols.")
    if any(s.is_real is False for s in states):
        raise ValueError("State symbols must represent real variables.")
    exprs = [sp.sympify(e) for e in rhs]
    if len(exprs) != len(states):
        raise ValueError("Provide one RHS expression per state symbol.")
    init = None if y0 is None else [sp.sympify(e) for e in y0]
    if init is not None and len(init) != len(states):
        raise ValueError("y0 must contain one initial value per state symbol.")
    obs = sp.sympify(observable) if observable is not None else sp.Integer(0)
    all_syms = set().union(*(e.free_symbols for e in [*exprs, *(init or []), obs]))
    candidates = [s for s in all_syms if s.name == "t" and s not in states]
    if time_symbol is None:
        if len(candidates) > 1:
            raise ValueError("Specify time_symbol when multiple different symbols are named t.")
        time_symbol = candidates[0] if candidates else sp.Symbol("t", real=True)
    if not isinstance(time_symbol, sp.Symbol) or time_symbol in states:
        raise ValueError("time_symbol must be a Symbol distinct from state_symbols.")
    for i, e in enumerate(exprs):
        if isinstance(e, sp.Equality):
            if e.lhs != sp.Derivative(states[i], time_symbol, evaluate=False):
                raise ValueError("Each equation must explicitly define its state derivative.")
            exprs[i] = e.rhs
    if any(e.has(sp.Derivative) for e in exprs):
        raise ValueError("Supply explicit first-order RHS expressions, without derivatives.")
    sweeps = {} if var_arrays is None else dict(var_arrays)
    constants = {} if const_values is None else dict(const_values)
    runtime = {} if runtime_consts is None else dict(runtime_consts)
    params = set(sweeps) | set(constants) | set(runtime)
    if (any(not isinstance(s, sp.Symbol) for s in params)
            or params & (set(states) | {time_symbol})):
        raise ValueError("Parameter keys must be Symbols distinct from the state and time symbols.")
    if set(sweeps) & (set(constants) | set(runtime)):
        raise ValueError("A swept parameter cannot also be a constant.")
    unknown = all_syms - set(states) - {time_symbol} - params
    if unknown:
        raise ValueError(f"Unassigned ODE parameters: {sorted(map(str, unknown))}")
    if init is not None and any(e.free_symbols & (set(states) | {time_symbol}) for e in init):
        raise ValueError("Initial values may depend on parameters, not on state or time.")
    repl = {s: sp.Symbol(f"rho[{i}]", real=True) for i, s in enumerate(states)}
    repl[time_symbol] = TIME
    repl.update({s: sp.Symbol(f"ode_param_{i}", real=True)
                 for i, s in enumerate(sorted(params, key=sp.srepr))})
    exprs = [e.xreplace(repl) for e in exprs]
    init = None if init is None else [e.xreplace(repl) for e in init]
    if any(sp.simplify(sp.im(e)) != 0 for e in [*exprs, *(init or [])]):
        raise ValueError("ODE states and RHS must be real; split complex states into "
                         "real/imaginary parts.")
    obs = obs.xreplace(repl)

    def remap(mapping):
        return {repl[s]: v for s, v in mapping.items()}
    sweeps = remap(sweeps) or {_DUMMY_SWEEP: np.zeros(1)}
    if keep_symbolic_consts is not None and not isinstance(keep_symbolic_consts, str):
        keep_symbolic_consts = [repl[s] for s in keep_symbolic_consts]
    return {"H": sp.Matrix(exprs), "mean": sp.Matrix([obs]), "initial": init, "sweeps": sweeps,
            "constants": remap(constants), "runtime": remap(runtime), "keep": keep_symbolic_consts,
            "size": len(states), "replacements": repl}


def _solve_ode_system(rhs, state_symbols, y0, tlist, *, var_arrays, const_values, time_symbol,
                      observable, output_mode, y0_values, runtime_consts, keep_symbolic_consts,
                      solver, fp64, return_time_trace, time_trace_every, t_in, return_device,
                      calibrate_timesteps, **option_set):
    from .solver import CompiledRHS, _solve_symbolic_system
    if tlist is None:
        raise TypeError("odesolve_2D needs tlist.")
    if output_mode not in {"mean", "final"}:
        raise ValueError("output_mode must be 'mean' or 'final'.")
    if y0 is not None and y0_values is not None:
        raise ValueError("Use either y0 or y0_values, not both.")
    if y0_values is not None and not hasattr(y0_values, "data_ptr") and np.iscomplexobj(y0_values):
        raise ValueError("y0_values must be real; split complex states into real/imaginary parts.")
    if any(k.startswith("_") or k in {"rho0", "rho0_values", "rho0_var_arrays", "compiled_rhs"}
           for k in option_set):
        raise ValueError("Use y0/y0_values and var_arrays for ODE initial conditions.")
    if isinstance(rhs, CompiledRHS):
        compiled = rhs
        if compiled.kind != "ode" or compiled.ode is None:
            raise ValueError("Pass a CompiledRHS from compile_ode_rhs (compile_lindblad_rhs "
                             "results go to mesolve_2D).")
        if any(v is not None for v in (state_symbols, y0, const_values, keep_symbolic_consts,
                                       time_symbol, observable)):
            raise ValueError("state_symbols, y0, const_values, keep_symbolic_consts, time_symbol "
                             "and observable are fixed when the RHS is compiled.")
        if y0_values is None and compiled.rho0_exprs is None:
            raise ValueError("Provide y0_values: no y0 was given when the RHS was compiled.")
        if tuple(var_arrays or {}) != compiled.ode["sweep_symbols"]:
            raise ValueError(f"var_arrays must sweep {compiled.ode['sweep_symbols']}, in the "
                             "order given when the RHS was compiled.")
        smap = compiled.ode["symbol_map"]

        def remap(mapping, name):
            bad = [s for s in mapping if s not in smap]
            if bad:
                raise ValueError(f"{name} contains symbols unknown to the compiled RHS: {bad}")
            return {smap[s]: v for s, v in mapping.items()}
        sweeps = remap(var_arrays or {}, "var_arrays") or {_DUMMY_SWEEP: np.zeros(1)}
        runtime = remap(runtime_consts or {}, "runtime_consts")
        has_obs = compiled.ode["has_observable"]
        mode = output_mode if has_obs else {"final": "final_rho", "mean": "mean_state"}[output_mode]
        return _solve_symbolic_system(
            None, None, None, None, tlist, var_arrays=sweeps, runtime_consts=runtime,
            rho0_values=y0_values, output_mode=mode, solver=solver, fp64=fp64,
            return_time_trace=return_time_trace, time_trace_every=time_trace_every, t_in=t_in,
            return_device=return_device, calibrate_timesteps=calibrate_timesteps,
            compiled_rhs=compiled, _ode_size=compiled.state_size, _state_trace=not has_obs,
            **option_set)
    if state_symbols is None:
        raise TypeError("odesolve_2D needs state_symbols (or a CompiledRHS as rhs).")
    if y0 is None and y0_values is None:
        raise ValueError("Provide y0 or y0_values.")
    model = _prepare_ode_model(rhs, state_symbols, y0, var_arrays, const_values, runtime_consts,
                               keep_symbolic_consts, time_symbol, observable)
    for k, v in _ODE_CODEGEN_DEFAULTS.items():
        option_set.setdefault(k, v)
    mode = output_mode if observable is not None else {"final": "final_rho",
                                                       "mean": "mean_state"}[output_mode]
    return _solve_symbolic_system(
        model["H"], {}, [], model["mean"], tlist, var_arrays=model["sweeps"],
        const_values=model["constants"], runtime_consts=model["runtime"],
        keep_symbolic_consts=model["keep"], rho0=model["initial"], rho0_values=y0_values,
        output_mode=mode, solver=solver, fp64=fp64, return_time_trace=return_time_trace,
        time_trace_every=time_trace_every, t_in=t_in, return_device=return_device,
        calibrate_timesteps=calibrate_timesteps, _ode_size=model["size"],
        _state_trace=observable is None, **option_set)
