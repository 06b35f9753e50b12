"""``odesolve_2D`` / ``compile_ode_rhs``: general real ODE sweeps on the same kernels.

The public behaviour follows the reference ODE front end (``gqis/gqis_ode_api.py:17-255``):
the same signatures, defaults, option names, error messages and result shapes.  Internally a
user model becomes an ``OdeModel`` in kernel names -- states ``rho[i]``, the time symbol the
kernel time, other parameters ``ode_param_i`` -- and runs through the Lindblad engine as an
``ode_size``-state system whose "Hamiltonian" slot holds the RHS column.  ``final`` output
maps to the full state and ``mean`` to the state mean when no observable is given.
"""
import dataclasses

import numpy as np
import sympy as sp

from ._options import fold_option_groups, split_groups
from .symbolic import TIME

# ODE right-hand sides are not simplified on input, so CSE runs with simplification, and the
# Lindblad-specific rewrites (expansion, collection by state) are off by default.
_ODE_CODEGEN_DEFAULTS = {"pre_expand": False, "collect_rho": False, "cse_simplify": True}
_ODE_CODEGEN_OPTIONS = ("pre_expand", "collect_rho", "factor_terms", "cse_batch_size",
                        "cse_simplify", "hoist_rho_independent", "RHSreuse")
# A sweep axis of length one stands in for a model without sweep parameters.
_NO_SWEEP = sp.Symbol("ode_dummy", real=True)
# Keywords of mesolve_2D that have an ODE counterpart (or none) and are refused here.
_LINDBLAD_ONLY = frozenset({"rho0", "rho0_values", "rho0_var_arrays", "compiled_rhs"})
# Fixed when the RHS is compiled; a call with a CompiledRHS must leave them unset.
_COMPILE_TIME_ARGS = ("state_symbols", "y0", "const_values", "keep_symbolic_consts",
                      "time_symbol", "observable")


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
    call = dict(locals())
    forwarded = call.pop("option_set")
    call, forwarded = fold_option_groups(odesolve_2D, call, split_groups(call), forwarded)
    return _run(call, forwarded)


def compile_ode_rhs(rhs, state_symbols, y0=None, sweep_symbols=(), *, const_values=None,
                    runtime_consts=None, keep_symbolic_consts=None, time_symbol=None,
                    observable=None, **codegen_options):
    """Prepare and optimise an ODE RHS without the GPU (``gqis/gqis_ode_api.py:70-112``)."""
    stray = sorted(set(codegen_options).difference(_ODE_CODEGEN_OPTIONS))
    if stray:
        raise TypeError(f"Unknown compile_ode_rhs option(s): {stray}")
    sweep_symbols = tuple(sweep_symbols)
    if len(sweep_symbols) > 2:
        raise ValueError("At most two sweep symbols are supported.")
    model = ode_model(rhs, state_symbols, y0, dict.fromkeys(sweep_symbols),
                      const_values=const_values, runtime_consts=runtime_consts,
                      keep_symbolic_consts=keep_symbolic_consts, time_symbol=time_symbol,
                      observable=observable)
    options = dict(_ODE_CODEGEN_DEFAULTS, **codegen_options)
    from .solver import _compile_rhs
    compiled, _ = _compile_rhs(model.equations, {}, [], model.observable,
                               sweep_symbols=tuple(model.sweeps), reuse=options.pop("RHSreuse", True),
                               **model.symbol_kwargs(), **options)
    # the name map belongs to this call: one cached internal model can serve several user
    # spellings of the same equations
    return dataclasses.replace(compiled, ode={"symbol_map": model.names,
                                              "sweep_symbols": sweep_symbols,
                                              "has_observable": observable is not None})


@dataclasses.dataclass
class OdeModel:
    """A validated ODE model in kernel names."""
    equations: sp.Matrix          # RHS column, one row per state
    observable: sp.Matrix         # 1 x 1 (zero without an observable)
    initial: list                 # y0 expressions, or None
    sweeps: dict                  # kernel symbol -> values (a dummy axis when none is swept)
    constants: dict
    runtime: dict
    keep: object                  # keep_symbolic_consts in kernel names (or its string form)
    names: dict                   # user symbol -> kernel symbol

    @property
    def size(self):
        return self.equations.shape[0]

    def symbol_kwargs(self):
        """Keywords shared by the engine's compile and solve entry points."""
        return {"const_values": self.constants, "runtime_consts": self.runtime,
                "keep_symbolic_consts": self.keep, "rho0": self.initial, "ode_size": self.size}


def _checked_states(state_symbols):
    states = tuple(state_symbols)
    distinct = len(set(states)) == len(states)
    if not states or not distinct or not all(isinstance(s, sp.Symbol) for s in states):
        raise ValueError("state_symbols must contain distinct SymPy Symbols.")
    if any(s.is_real is False for s in states):
        raise ValueError("State symbols must represent real variables.")
    return states


def _choose_time_symbol(time_symbol, used, states):
    """The given time symbol, else the one free symbol named ``t`` that is not a state (a fresh
    real ``t`` when there is none)."""
    if time_symbol is None:
        named_t = {s for s in used if s.name == "t"}.difference(states)
        if len(named_t) > 1:
            raise ValueError("Specify time_symbol when multiple different symbols are named t.")
        time_symbol = named_t.pop() if named_t else sp.Symbol("t", real=True)
    if not isinstance(time_symbol, sp.Symbol) or time_symbol in states:
        raise ValueError("time_symbol must be a Symbol distinct from state_symbols.")
    return time_symbol


def _explicit_rhs(exprs, states, time_symbol):
    """Right-hand sides of ``Eq(Derivative(state, t), rhs)`` entries; plain entries as given."""
    out = []
    for state, e in zip(states, exprs):
        if isinstance(e, sp.Equality):
            if e.lhs != sp.Derivative(state, time_symbol, evaluate=False):
                raise ValueError("Each equation must explicitly define its state derivative.")
            e = e.rhs
        out.append(e)
    if any(e.has(sp.Derivative) for e in out):
        raise ValueError("Supply explicit first-order RHS expressions, without derivatives.")
    return out


def _kernel_names(states, time_symbol, parameters):
    names = {s: sp.Symbol(f"rho[{i}]", real=True) for i, s in enumerate(states)}
    names[time_symbol] = TIME
    ordered = sorted(parameters, key=lambda s: (s.name, sp.srepr(s)))
    names.update((s, sp.Symbol(f"ode_param_{i}", real=True)) for i, s in enumerate(ordered))
    return names


def ode_model(rhs, state_symbols, y0=None, var_arrays=None, *, const_values=None,
              runtime_consts=None, keep_symbolic_consts=None, time_symbol=None, observable=None):
    """Validate a user ODE model and rename it to kernel symbols (the checks and messages of
    ``gqis/gqis_ode_api.py:115-181``)."""
    states = _checked_states(state_symbols)
    exprs = list(map(sp.sympify, rhs))
    if len(exprs) != len(states):
        raise ValueError("Provide one RHS expression per state symbol.")
    init = None if y0 is None else list(map(sp.sympify, y0))
    if init is not None and len(init) != len(states):
        raise ValueError("y0 must contain one initial value per state symbol.")
    obs = sp.Integer(0) if observable is None else sp.sympify(observable)
    used = set()
    for e in (*exprs, *(init or ()), obs):
        used |= e.free_symbols
    time_symbol = _choose_time_symbol(time_symbol, used, states)
    exprs = _explicit_rhs(exprs, states, time_symbol)

    sweeps, constants, runtime = (dict(m or {}) for m in (var_arrays, const_values,
                                                          runtime_consts))
    parameters = set(sweeps) | set(constants) | set(runtime)
    reserved = set(states) | {time_symbol}
    if not all(isinstance(s, sp.Symbol) for s in parameters) or parameters & reserved:
        raise ValueError("Parameter keys must be Symbols distinct from the state and time symbols.")
    if set(sweeps) & (set(constants) | set(runtime)):
        raise ValueError("A swept parameter cannot also be a constant.")
    unassigned = used - reserved - parameters
    if unassigned:
        raise ValueError(f"Unassigned ODE parameters: {sorted(map(str, unassigned))}")
    if init is not None and any(e.free_symbols & reserved for e in init):
        raise ValueError("Initial values may depend on parameters, not on state or time.")

    names = _kernel_names(states, time_symbol, parameters)
    exprs = [e.xreplace(names) for e in exprs]
    init = None if init is None else [e.xreplace(names) for e in init]
    if not all(sp.simplify(sp.im(e)) == 0 for e in (*exprs, *(init or ()))):
        raise ValueError("ODE states and RHS must be real; split complex states into "
                         "real/imaginary parts.")
    keep = keep_symbolic_consts
    if keep is not None and not isinstance(keep, str):
        keep = [names[s] for s in keep]
    return OdeModel(equations=sp.Matrix(exprs), observable=sp.Matrix([obs.xreplace(names)]),
                    initial=init,
                    sweeps={names[s]: v for s, v in sweeps.items()} or {_NO_SWEEP: np.zeros(1)},
                    constants={names[s]: v for s, v in constants.items()},
                    runtime={names[s]: v for s, v in runtime.items()}, keep=keep, names=names)


def _check_call(call, forwarded):
    if call["tlist"] is None:
        raise TypeError("odesolve_2D needs tlist.")
    if call["output_mode"] not in ("mean", "final"):
        raise ValueError("output_mode must be 'mean' or 'final'.")
    y0_values = call["y0_values"]
    if call["y0"] is not None and y0_values is not None:
        raise ValueError("Use either y0 or y0_values, not both.")
    # device tensors are taken as they are (their dtype is checked where they are uploaded)
    if y0_values is not None and not hasattr(y0_values, "data_ptr") \
            and np.iscomplexobj(y0_values):
        raise ValueError("y0_values must be real; split complex states into real/imaginary parts.")
    if any(k.startswith("_") or k in _LINDBLAD_ONLY for k in forwarded):
        raise ValueError("Use y0/y0_values and var_arrays for ODE initial conditions.")


def _compiled_model_args(call):
    """Engine keywords for a ``CompiledRHS`` from ``compile_ode_rhs``."""
    compiled = call["rhs"]
    if compiled.kind != "ode" or compiled.ode is None:
        raise ValueError("Pass a CompiledRHS from compile_ode_rhs (compile_lindblad_rhs "
                         "results go to mesolve_2D).")
    if any(call[name] is not None for name in _COMPILE_TIME_ARGS):
        raise ValueError("state_symbols, y0, const_values, keep_symbolic_consts, time_symbol "
                         "and observable are fixed when the RHS is compiled.")
    if call["y0_values"] is None and compiled.rho0_exprs is None:
        raise ValueError("Provide y0_values: no y0 was given when the RHS was compiled.")
    info = compiled.ode
    if tuple(call["var_arrays"] or {}) != info["sweep_symbols"]:
        raise ValueError(f"var_arrays must sweep {info['sweep_symbols']}, in the order given "
                         "when the RHS was compiled.")

    def to_kernel_names(mapping, what):
        foreign = [s for s in mapping if s not in info["symbol_map"]]
        if foreign:
            raise ValueError(f"{what} contains symbols unknown to the compiled RHS: {foreign}")
        return {info["symbol_map"][s]: v for s, v in mapping.items()}
    sweeps = to_kernel_names(call["var_arrays"] or {}, "var_arrays")
    engine = {"compiled_rhs": compiled, "var_arrays": sweeps or {_NO_SWEEP: np.zeros(1)},
              "runtime_consts": to_kernel_names(call["runtime_consts"] or {}, "runtime_consts"),
              "_ode_size": compiled.state_size}
    return (None, None, None, None), engine, info["has_observable"], {}


def _symbolic_model_args(call):
    """Engine keywords for a model given as expressions."""
    if call["state_symbols"] is None:
        raise TypeError("odesolve_2D needs state_symbols (or a CompiledRHS as rhs).")
    if call["y0"] is None and call["y0_values"] is None:
        raise ValueError("Provide y0 or y0_values.")
    model = ode_model(call["rhs"], call["state_symbols"], call["y0"], call["var_arrays"],
                      const_values=call["const_values"], runtime_consts=call["runtime_consts"],
                      keep_symbolic_consts=call["keep_symbolic_consts"],
                      time_symbol=call["time_symbol"], observable=call["observable"])
    kw = model.symbol_kwargs()
    engine = {"var_arrays": model.sweeps, "const_values": kw["const_values"],
              "runtime_consts": kw["runtime_consts"], "keep_symbolic_consts": kw["keep_symbolic_consts"],
              "rho0": kw["rho0"], "_ode_size": model.size}
    return ((model.equations, {}, [], model.observable), engine, call["observable"] is not None,
            _ODE_CODEGEN_DEFAULTS)


def _run(call, forwarded):
    """Validate an ``odesolve_2D`` call and run it on the shared sweep engine."""
    from .solver import CompiledRHS, _solve_symbolic_system
    _check_call(call, forwarded)
    build = _compiled_model_args if isinstance(call["rhs"], CompiledRHS) else _symbolic_model_args
    model_args, engine, has_observable, codegen_defaults = build(call)
    # without an observable the result is the state: final -> final_rho, mean -> mean_state
    mode = call["output_mode"]
    if not has_observable:
        mode = "final_rho" if mode == "final" else "mean_state"
    passed = ("solver", "fp64", "return_time_trace", "time_trace_every", "t_in", "return_device",
              "calibrate_timesteps")
    engine.update({name: call[name] for name in passed})
    engine.update(output_mode=mode, rho0_values=call["y0_values"], _state_trace=not has_observable)
    options = {**codegen_defaults, **forwarded}
    return _solve_symbolic_system(*model_args, call["tlist"], **engine, **options)

