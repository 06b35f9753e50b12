"""``mesolve_2D`` and the shared sweep driver on the MI355X.

Driver semantics follow the reference ``_solve_symbolic_system`` (``gqis/solver.py:1739-2391``):
validation of ``tlist`` and options, warmup and trace bookkeeping, the process-local RHS and
kernel caches, the finite check of the outputs and the timing dictionary.  The device work is
HIP: the kernel source is emitted from the symbolic plan, compiled with hiprtc and launched
through ``libgqis_hip.so``; torch provides the device arrays and the stream.
"""
from collections import ChainMap
import dataclasses
from dataclasses import dataclass, field
import os
import time

import numpy as np
import sympy as sp

from ._options import fold_option_groups, split_groups
from . import codegen, symbolic
from .solvers import get_solver_spec, normalize_solver_name  # noqa: F401 (re-export)
from .symbolic import (PARX, PARY, TIME, build_independent_rho,  # noqa: F401 (re-export)
                       build_reduced_lindblad_rhs, rho_matrix_to_independent_exprs)

_RHS_CODE_CACHE = {}
_KERNEL_CACHE = {}
_TIMESTEP_CALIBRATION_CACHE = {}


@dataclass(eq=False)
class CompiledRHS:
    """Prepared and optimised right-hand side, independent of the device code
    (the reference's ``CompiledRHS``, ``gqis/solver.py:1344-1371``)."""
    kind: str
    state_size: int
    sweep_symbols: tuple
    state_symbols: tuple
    equations: tuple
    mean_re: object
    mean_im: object
    drive_map: object
    rho0_exprs: object
    runtime_const_syms: tuple
    base_runtime_syms: tuple
    derived_const_exprs: tuple
    const_values: dict
    plan: object
    cache_key: tuple
    timings: dict
    ode: object = None
    emitted: dict = field(default_factory=dict, repr=False)
    source_model: dict = field(default=None, repr=False)   # prepared operators (Lindblad)


def _rename_time(expr):
    """Any symbol named ``t`` is the time (the reference prints it as the kernel's ``t``)."""
    expr = sp.sympify(expr)
    subs = {s: TIME for s in expr.free_symbols if s.name == "t" and s != TIME}
    return expr.xreplace(subs) if subs else expr


def _prepare_model(H, Drive, Col_Ops, mean_operator, *, sweep_symbols, const_values,
                   runtime_consts, keep_symbolic_consts, auto_runtime_consts, rho0, ode_size):
    """Validate the model; map sweep symbols to ParX/ParY; fold or keep constants
    (behaviour of ``gqis/solver.py:1374-1521``)."""
    const_values = {} if const_values is None else dict(const_values)
    runtime_consts = runtime_consts or {}
    if not isinstance(runtime_consts, dict):
        raise ValueError("runtime_consts must be a dict of {Symbol: value}")
    n = H.shape[0]
    state_size = n * n - 1 if ode_size is None else ode_size
    rho0_exprs = None
    if rho0 is not None:
        rho0_exprs = (rho_matrix_to_independent_exprs(rho0) if ode_size is None
                      else [sp.sympify(v) for v in rho0])
        if len(rho0_exprs) != state_size:
            raise ValueError("rho0 does not match Hamiltonian size.")
    for op in Col_Ops:
        if op.shape != (n, n):
            raise ValueError(f"Collapse operator dimension {op.shape} does not match H {n, n}")
    if ode_size is None and mean_operator.shape != (n, n):
        raise ValueError(f"Mean operator dimension {mean_operator.shape} does not match H {n, n}")

    sweep_map = {s: p for s, p in zip(sweep_symbols, (PARX, PARY))}
    if keep_symbolic_consts is None:
        keep = set(const_values) if auto_runtime_consts else set()
    elif isinstance(keep_symbolic_consts, str):
        if keep_symbolic_consts not in ("all", "auto", "const_values"):
            raise ValueError("keep_symbolic_consts must be an iterable of Symbols, 'all', "
                             "'auto', 'const_values' or None")
        keep = set(const_values)
    else:
        keep = set(keep_symbolic_consts)
    keep |= set(runtime_consts)
    for s in keep:
        if not isinstance(s, sp.Symbol):
            raise ValueError("keep_symbolic_consts/runtime_consts keys must be SymPy Symbols")
    folded = {k: sp.Float(float(v)) for k, v in const_values.items() if k not in keep}

    def prep(e):
        return _rename_time(sp.sympify(e).xreplace(sweep_map).xreplace(folded))

    H = H.applyfunc(prep)
    if isinstance(Drive, dict):
        Drive = {k: prep(v) for k, v in Drive.items()}
    else:
        Drive = prep(Drive)
    Col_Ops = [op.applyfunc(prep) for op in Col_Ops]
    mean_operator = mean_operator.applyfunc(prep)
    if rho0_exprs is not None:
        rho0_exprs = [prep(e) for e in rho0_exprs]
    used = set()
    for e in list(H) + (list(Drive.values()) if isinstance(Drive, dict) else [Drive]) \
            + [x for op in Col_Ops for x in op] + list(mean_operator) + list(rho0_exprs or []):
        used |= sp.sympify(e).free_symbols
    runtime_syms = tuple(sorted((s for s in keep if s in used), key=lambda s: s.name))
    defaults = {s: v for s, v in {**const_values, **runtime_consts}.items() if s in runtime_syms}
    return {"N": n, "state_size": state_size, "H": H, "Drive": Drive, "Col_Ops": Col_Ops,
            "mean_operator": mean_operator, "rho0_exprs": rho0_exprs,
            "runtime_const_syms": runtime_syms, "folded": folded, "defaults": defaults}


def _drive_map(drive, default_symbol=sp.Symbol("Drive", real=True)):
    if isinstance(drive, dict):
        return dict(drive)
    if drive is None:
        return {}
    return {default_symbol: drive}


def _compile_rhs(H, Drive, Col_Ops, mean_operator, *, sweep_symbols, const_values=None,
                 runtime_consts=None, keep_symbolic_consts=None, auto_runtime_consts=False,
                 rho0=None, ode_size=None, pre_expand=True, collect_rho=False, factor_terms=False,
                 cse_batch_size=None, cse_simplify=False, hoist_rho_independent=True, reuse=True,
                 symbolic_workers="auto", max_symbolic_workers=None):
    """Symbolic ODE and hoisting plan, cached per model; returns ``(CompiledRHS, status)``."""
    start = time.time()
    model = _prepare_model(H, Drive, Col_Ops, mean_operator, sweep_symbols=sweep_symbols,
                           const_values=const_values, runtime_consts=runtime_consts,
                           keep_symbolic_consts=keep_symbolic_consts,
                           auto_runtime_consts=auto_runtime_consts, rho0=rho0, ode_size=ode_size)
    drive = _drive_map(model["Drive"])
    key = ("compiled_rhs", model["N"], ode_size, sp.srepr(model["H"]),
           tuple((str(k), sp.srepr(v)) for k, v in drive.items()),
           tuple(sp.srepr(op) for op in model["Col_Ops"]), sp.srepr(model["mean_operator"]),
           tuple(sorted((str(k), float(v)) for k, v in model["folded"].items())),
           tuple(s.name for s in model["runtime_const_syms"]),
           None if model["rho0_exprs"] is None else tuple(sp.srepr(e) for e in model["rho0_exprs"]),
           pre_expand, collect_rho, factor_terms, cse_batch_size, cse_simplify,
           hoist_rho_independent)
    if reuse and key in _RHS_CODE_CACHE:
        return _RHS_CODE_CACHE[key], "hit"
    if ode_size is None:
        eqs, mre, mim, meta = build_reduced_lindblad_rhs(
            model["N"], model["H"], model["Col_Ops"], model["mean_operator"],
            pre_expand=pre_expand, collect_rho=collect_rho, factor_terms=factor_terms,
            symbolic_workers=symbolic_workers, max_symbolic_workers=max_symbolic_workers)
        states = meta["rho_syms"]
    else:
        states = symbolic.state_symbols(ode_size)
        eqs = list(model["H"])
        if pre_expand:
            eqs = [sp.expand(e) for e in eqs]
        if factor_terms:
            eqs = [sp.factor_terms(e) for e in eqs]
        if collect_rho:
            eqs = [sp.collect(e, states) for e in eqs]
        obs = sp.sympify(model["mean_operator"][0])
        mre, mim = sp.re(obs), sp.im(obs)
    prepared = time.time()
    plan = symbolic.make_plan(eqs, mre, mim, states, drive, model["runtime_const_syms"],
                              cse_simplify=cse_simplify)
    compiled = CompiledRHS(
        kind="lindblad" if ode_size is None else "ode", state_size=model["state_size"],
        sweep_symbols=tuple(sweep_symbols), state_symbols=tuple(states), equations=tuple(eqs),
        mean_re=mre, mean_im=mim, drive_map=model["Drive"],
        rho0_exprs=None if model["rho0_exprs"] is None else tuple(model["rho0_exprs"]),
        runtime_const_syms=tuple(model["runtime_const_syms"]),
        base_runtime_syms=tuple(model["runtime_const_syms"]), derived_const_exprs=(),
        const_values=model["defaults"], plan=plan, cache_key=key,
        timings={"ode_prepare_s": prepared - start, "ode_optimize_s": time.time() - prepared},
        source_model=model if ode_size is None else None)
    if reuse:
        _RHS_CODE_CACHE[key] = compiled
    return compiled, "miss"


def compile_lindblad_rhs(H, Drive, Col_Ops, mean_operator, sweep_symbols, *, const_values=None,
                         runtime_consts=None, keep_symbolic_consts=None, auto_runtime_consts=False,
                         rho0=None, pre_expand=True, collect_rho=False, factor_terms=False,
                         cse_batch_size=None, cse_simplify=False, hoist_rho_independent=True,
                         RHSreuse=True, symbolic_workers="auto", max_symbolic_workers=None):
    """Prepare and optimise a Lindblad RHS without the GPU (``gqis/solver.py:1592-1627``)."""
    sweep_symbols = tuple(sweep_symbols)
    if not 1 <= len(sweep_symbols) <= 2:
        raise ValueError("sweep_symbols must contain one or two symbols.")
    compiled, _ = _compile_rhs(
        H, Drive, Col_Ops, mean_operator, sweep_symbols=sweep_symbols, const_values=const_values,
        runtime_consts=runtime_consts, keep_symbolic_consts=keep_symbolic_consts,
        auto_runtime_consts=auto_runtime_consts, rho0=rho0, pre_expand=pre_expand,
        collect_rho=collect_rho, factor_terms=factor_terms, cse_batch_size=cse_batch_size,
        cse_simplify=cse_simplify, hoist_rho_independent=hoist_rho_independent, reuse=RHSreuse,
        symbolic_workers=symbolic_workers, max_symbolic_workers=max_symbolic_workers)
    return compiled


def _runtime_const_values(compiled, runtime_consts, const_values, use_defaults):
    """Kernel-argument values of the compiled model's runtime constants, in argument order.
    A call's ``runtime_consts`` take precedence over its ``const_values``, and both over the
    values given when the RHS was compiled (consulted only with ``use_defaults``)."""
    layers = [runtime_consts or {}, const_values or {}]
    if use_defaults:
        layers.append(compiled.const_values)
    lookup = ChainMap(*layers)
    absent = [s for s in compiled.runtime_const_syms if s not in lookup]
    if absent:
        raise ValueError(f"Missing value for runtime constant symbol: {absent[0]}")
    return [float(lookup[s]) for s in compiled.runtime_const_syms]


def mesolve_2D(H, Drive=None, Col_Ops=None, mean_operator=None, tlist=None,
               var_arrays=None, const_values=None, kernel_template_file=None, *,
               RHSreuse=True, runtime_consts=None, keep_symbolic_consts=None,
               auto_runtime_consts=False, output_mode="mean", solver="rk4",
               solver_frequency=None, adaptive_options=None, unroll=False, fp64=False,
               pre_expand=True, collect_rho=False, factor_terms=False,
               cse_batch_size=None, cse_simplify=False, hoist_rho_independent=True,
               symbolic_workers="auto", max_symbolic_workers=None,
               nvrtc_options=(), timings=False, return_timing_info=False, warmup_time=0.0,
               rho0=None, rho0_var_arrays=None, rho0_values=None,
               return_time_trace=False, time_trace_every=None,
               time_trace_samples_per_period=None, solver_samples_per_period=None,
               Actual_Kernel_Save=False, beep_on_error=False, ignore_non_finite_output=False,
               t_in=0.0, return_device=False, calibrate_timesteps=None, consts=None, initial=None,
               output=None, solver_options=None, codegen=None, diagnostics=None,
               hiprtc_options=(), time_table=True, block_size=256, liouville="auto",
               lanes_per_point="auto"):
    """Lindblad master-equation sweep over one or two parameter axes on the GPU.

    Same signature, defaults, return shapes and errors as the reference ``mesolve_2D``
    (``gqis/solver.py:1712-1736``, ``GQIS_API.md:505-528``), plus three code-generation options:
    ``hiprtc_options`` (extra compiler flags), ``time_table`` (FP64 table of the time-only drive
    terms for fixed-step solvers; False evaluates them per point as the reference does) and
    ``block_size`` (threads per work-group, or ``"auto"``: measured once per kernel and device)
    ``lanes_per_point`` (1, 2, 4, 8 or 16 lanes, one per wavefront, share each point's state rows;
    "auto": 8 for states of 48 or more reals where the grouped kernel applies, else 1;
    ``block_size="auto"`` also measures 2, 4, 8 for models of 24 or more states) and
    ``liouville`` ("auto", "lanes", "mfma", or "mfma_f16" / "mfma_bf16": the commutator product
    as three FP16 / BF16 matrix-core products of hi/lo parts with FP32 accumulation; the
    Hilbert-space path for N = 8, 16, 32 on the FP32 or FP64 matrix cores, used automatically
    in RK4 sweeps whose operators fit it: at N = 32, at N = 16 for a dense H_eff, at N = 8 for a
    dense H_eff in FP64).  ``return_device=True`` returns torch tensors on
    the GPU.  ``H`` may be a ``CompiledRHS`` from ``compile_lindblad_rhs``.

    Returns ``(num_X, num_Y)`` complex for ``"mean"``/``"final"``, ``(num_X, num_Y, N*N-1)``
    real for ``"final_rho"``; with ``return_time_trace`` also the ``(num_X, num_Y, num_trace)``
    trace and its times; with ``return_timing_info`` the timing dictionary last.
    """
    return _solve_symbolic_system(**_lindblad_call(dict(locals())))


def _lindblad_call(call):
    """Engine keywords of a ``mesolve_2D`` call: option groups folded into the flat names, the
    output mode checked, a ``CompiledRHS`` in place of ``H`` moved to ``compiled_rhs``."""
    call, _ = fold_option_groups(mesolve_2D, call, split_groups(call))
    if call["output_mode"] not in {"mean", "final", "final_rho"}:
        raise ValueError("output_mode must be 'mean', 'final', or 'final_rho'.")
    model = call["H"]
    if isinstance(model, CompiledRHS):
        if model.kind != "lindblad":
            raise ValueError("This CompiledRHS comes from compile_ode_rhs; pass it to odesolve_2D.")
        call["H"], call["compiled_rhs"] = None, model
    return call


def set_kernel_cache_dir(path, device=None):
    """Keep compiled sweep kernels in ``path`` across processes (``None`` turns it off).  The
    environment variable ``GQIS_KERNEL_CACHE_DIR`` sets the same at the first sweep.  The
    reference's kernel cache is per process only (``gqis/solver.py:2054-2060``)."""
    from .runtime import Context
    Context.get(device).set_cache_dir(path)


def prepare_sweep(*args, **kwargs):
    """A sweep prepared once and replayed many times (``PreparedSweep.run``): same arguments
    as ``mesolve_2D``; the per-point kernels (any fixed-step or adaptive solver, lanes per
    point) with their buffers resident and the launch chain recorded as a HIP graph.  For
    fitting loops that re-run one sweep with new sweep values, runtime constants or initial
    states.  No counterpart in the reference, which redoes the whole call each time."""
    import inspect
    bound = inspect.signature(mesolve_2D).bind(*args, **kwargs)
    bound.apply_defaults()
    call = _lindblad_call(dict(bound.arguments))
    for name in ("return_timing_info", "timings"):
        if call[name]:
            raise ValueError(f"{name} is not available for a prepared sweep (use run(timed=True) "
                             "and last_gpu_s).")
    return _solve_symbolic_system(**call, _prepare=True)


def _check_tlist(tlist):
    """(num_steps, dt) of a uniform grid starting at zero (``gqis/solver.py:1951-1979``)."""
    if tlist is None:
        raise ValueError("tlist is required for fixed-step integration.")
    n_t = int(len(tlist))
    if n_t < 2:
        raise ValueError("tlist must contain at least two time samples.")
    t_in_arr = np.asarray(tlist)
    t = np.asarray(t_in_arr, dtype=np.float64)
    if t.ndim != 1 or not np.all(np.isfinite(t)):
        raise ValueError("tlist must be a finite one-dimensional array.")
    diffs = np.diff(t)
    if np.any(diffs <= 0.0):
        raise ValueError("tlist must be strictly increasing.")
    ftype = t_in_arr.dtype if np.issubdtype(t_in_arr.dtype, np.floating) else np.dtype(np.float64)
    tol = 8.0 * np.finfo(ftype).eps * max(1.0, float(np.max(np.abs(t))))
    dt = float((t[-1] - t[0]) / (n_t - 1))
    return n_t - 1, dt, t, diffs, tol


def _solve_symbolic_system(H, Drive, Col_Ops, mean_operator, tlist, var_arrays=None,
                           const_values=None, kernel_template_file=None, *, RHSreuse=True,
                           runtime_consts=None, keep_symbolic_consts=None,
                           auto_runtime_consts=False, output_mode="mean", solver="rk4",
                           solver_frequency=None, adaptive_options=None, unroll=False, fp64=False,
                           pre_expand=True, collect_rho=False, factor_terms=False,
                           cse_batch_size=None, cse_simplify=False, hoist_rho_independent=True,
                           symbolic_workers="auto", max_symbolic_workers=None, nvrtc_options=(),
                           timings=False, return_timing_info=False, warmup_time=0.0, rho0=None,
                           rho0_var_arrays=None, rho0_values=None, return_time_trace=False,
                           time_trace_every=None, time_trace_samples_per_period=None,
                           solver_samples_per_period=None, Actual_Kernel_Save=False,
                           beep_on_error=False, ignore_non_finite_output=False, t_in=0.0,
                           return_device=False, calibrate_timesteps=None, compiled_rhs=None,
                           hiprtc_options=(), time_table=True, block_size=256,
                           liouville="auto", lanes_per_point="auto", _ode_size=None,
                           _state_trace=False, _prepare=False,
                           _time_trace_offset=0):
    if calibrate_timesteps is not None and calibrate_timesteps is not False:
        if _prepare:
            raise ValueError("calibrate_timesteps cannot be combined with prepare_sweep: "
                             "calibrate once with mesolve_2D, then prepare on the chosen tlist.")
        call = dict(locals())
        from .calibration import solve_with_timestep_calibration
        return solve_with_timestep_calibration(call)
    if output_mode not in {"mean", "final", "final_rho", "mean_state"}:
        raise ValueError(f"Unsupported output_mode '{output_mode}'. "
                         f"Use one of: {sorted(['final', 'final_rho', 'mean', 'mean_state'])}")
    if not isinstance(unroll, (bool, np.bool_)):
        raise ValueError("unroll must be True or False.")
    unroll = bool(unroll)
    if kernel_template_file is not None:
        raise ValueError("kernel_template_file: the HIP kernels are built from the package's "
                         "solver templates; custom CUDA templates are not supported.")
    mixed = isinstance(fp64, str)
    if mixed:
        if fp64 != "mixed":
            raise ValueError("fp64 must be True, False or 'mixed'.")
        if (normalize_solver_name(solver) != "rk4" or not time_table
                or lanes_per_point not in ("auto", 1) or liouville not in ("auto", "lanes")
                or return_time_trace):
            raise ValueError("fp64='mixed' (FP64 state, parameters and time terms with an FP32 "
                             "right-hand side) is a form of the one-lane RK4 kernel with a time "
                             "table: solver 'rk4', time_table=True, lanes_per_point 1, no "
                             "matrix-core path and no time trace.")
        fp64, liouville, lanes_per_point = False, "lanes", 1
    t_in = float(t_in)
    if not np.isfinite(t_in):
        raise ValueError("t_in must be finite.")
    if block_size != "auto":
        block_size = int(block_size)
        if block_size not in (64, 128, 256, 512, 1024):
            raise ValueError("block_size must be 64, 128, 256, 512, 1024 or 'auto'.")
    if lanes_per_point != "auto":
        lanes_per_point = int(lanes_per_point)
        if lanes_per_point not in (1, 2, 4, 8, 16):
            raise ValueError("lanes_per_point must be 1, 2, 4, 8, 16 or 'auto'.")

    var_arrays = {} if var_arrays is None else dict(var_arrays)
    for sym, arr in ({} if rho0_var_arrays is None else dict(rho0_var_arrays)).items():
        if sym in var_arrays:
            raise ValueError(f"Symbol {sym} is present in both var_arrays and rho0_var_arrays.")
        var_arrays[sym] = arr
    if rho0 is not None and rho0_values is not None:
        raise ValueError("Use either rho0 or rho0_values, not both.")
    uses_rho0_values = rho0_values is not None
    if not var_arrays:
        raise ValueError("var_arrays or rho0_var_arrays must contain at least one sweep array "
                         "(ParX).")
    if len(var_arrays) > 2:
        raise ValueError("Only 2 total sweep arrays are supported across var_arrays and "
                         "rho0_var_arrays.")
    if compiled_rhs is None:
        if H is None or Drive is None or Col_Ops is None or mean_operator is None:
            raise TypeError("Give H, Drive, Col_Ops and mean_operator, or a CompiledRHS.")
    else:
        if any(x is not None for x in (H, Drive, Col_Ops, mean_operator)):
            raise ValueError("With a CompiledRHS, do not pass the model (H, Drive, Col_Ops, "
                             "mean_operator) again.")
        if const_values or keep_symbolic_consts is not None or auto_runtime_consts or rho0 is not None:
            raise ValueError("const_values, keep_symbolic_consts, auto_runtime_consts and rho0 are "
                             "fixed when the RHS is compiled; give values of constants kept "
                             "symbolic through runtime_consts.")
        if tuple(var_arrays) != compiled_rhs.sweep_symbols:
            raise ValueError(f"The sweep arrays must be for {compiled_rhs.sweep_symbols}, in the "
                             "order given when the RHS was compiled.")

    num_steps, dt, t_host, diffs, tol = _check_tlist(tlist)
    spec = get_solver_spec(solver, frequency=solver_frequency, step_size=dt,
                           adaptive_options=adaptive_options, fp64=fp64)
    if mixed:
        spec = dataclasses.replace(spec, template="fixed_rk4_mixed.hip", kernel="gq_rk4_mixed")
    if not np.isclose(t_host[0], 0.0, rtol=0.0, atol=tol):
        raise ValueError("tlist must begin at zero; the kernel derives stage times from the "
                         "step index.")
    if not np.allclose(diffs, dt, rtol=1.0e-6, atol=tol):
        raise ValueError("tlist must be uniformly spaced because the GPU solver uses one "
                         "fixed dt.")
    warmup_time = float(warmup_time)
    if warmup_time < 0.0 or warmup_time > 1.0:
        raise ValueError("warmup_time must be within [0, 1].")
    warmup_steps = num_steps if warmup_time >= 1.0 else int(np.floor(warmup_time * num_steps))
    warmup_steps = max(0, min(num_steps, warmup_steps))
    return_time_trace = bool(return_time_trace)
    if time_trace_every is not None and time_trace_samples_per_period is not None:
        raise ValueError("Use either time_trace_every or time_trace_samples_per_period, not both.")
    if time_trace_samples_per_period is not None:
        if solver_samples_per_period is None:
            raise ValueError("solver_samples_per_period is required when "
                             "time_trace_samples_per_period is used.")
        sspp, tspp = int(solver_samples_per_period), int(time_trace_samples_per_period)
        if sspp <= 0 or tspp <= 0:
            raise ValueError("solver_samples_per_period and time_trace_samples_per_period must "
                             "be positive.")
        trace_stride = max(1, int(round(sspp / tspp)))
    elif time_trace_every is not None:
        trace_stride = int(time_trace_every)
        if trace_stride <= 0:
            raise ValueError("time_trace_every must be a positive integer.")
    else:
        trace_stride = 1
    trace_offset = int(_time_trace_offset)
    num_trace = (int(np.ceil((num_steps - trace_offset) / trace_stride))
                 if return_time_trace else 0)
    mode = output_mode
    if output_mode == "mean" and warmup_steps >= num_steps:
        mode = "final"
    if output_mode == "mean_state" and warmup_steps >= num_steps:
        mode = "final_rho"
    if timings not in (False, None, True, "compact", "full"):
        raise ValueError("timings must be False, True, 'compact' or 'full'.")
    collect = bool(timings or return_timing_info)
    total_start = time.time()

    if liouville not in ("auto", "lanes", "mfma", "mfma_f16", "mfma_bf16", "sparse"):
        raise ValueError("liouville must be 'auto', 'lanes', 'mfma', 'mfma_f16', 'mfma_bf16' or "
                         "'sparse'.")
    if liouville in ("mfma_f16", "mfma_bf16") and fp64:
        raise ValueError(f"liouville='{liouville}' (split-precision matrix-core products) is an "
                         "FP32-mode option.")
    hilbert_out = None
    if (compiled_rhs is None and _ode_size is None and liouville != "lanes"
            and not (lanes_per_point != "auto" and lanes_per_point > 1
                     and liouville == "auto")):
        hilbert_out = _try_hilbert(H, Drive, Col_Ops, mean_operator, var_arrays, const_values,
                                   runtime_consts, keep_symbolic_consts, auto_runtime_consts,
                                   rho0, rho0_values, spec, mode, fp64, return_time_trace,
                                   liouville, dt, t_in, num_steps, warmup_steps, block_size,
                                   collect, tuple(hiprtc_options) + _as_tuple(nvrtc_options),
                                   prepare=_prepare, num_trace=num_trace,
                                   trace_stride=trace_stride, trace_offset=trace_offset)
    if hilbert_out is not None and _prepare:
        from .prepared import PreparedSweep
        plan, stage_s, rhs_status = hilbert_out

        def finish(outs, *, return_device, ignore_non_finite_output, total_start):
            return _finish(outs, stage_s, rhs_status, spec, solver_frequency, unroll, False,
                           False, return_device, return_time_trace, False,
                           ignore_non_finite_output, None, num_trace, trace_stride, trace_offset,
                           fp64, t_in, dt, total_start)
        return PreparedSweep(plan, compiled=None, spec=spec, const_values=const_values,
                             runtime_consts=runtime_consts, use_defaults=False,
                             sweep_symbols=tuple(var_arrays),
                             uses_rho0_values=uses_rho0_values, finish=finish)
    if hilbert_out is not None:
        outputs, stage_s, rhs_status = hilbert_out
        return _finish(outputs, stage_s, rhs_status, spec, solver_frequency, unroll, collect,
                       timings, return_device, return_time_trace, return_timing_info,
                       ignore_non_finite_output, _ode_size, num_trace, trace_stride, trace_offset,
                       fp64, t_in, dt, total_start)

    # 1) symbolic ODE and plan
    if compiled_rhs is None:
        compiled, rhs_status = _compile_rhs(
            H, Drive, Col_Ops, mean_operator, sweep_symbols=tuple(var_arrays),
            const_values=const_values, runtime_consts=runtime_consts,
            keep_symbolic_consts=keep_symbolic_consts, auto_runtime_consts=auto_runtime_consts,
            rho0=rho0, ode_size=_ode_size, pre_expand=pre_expand, collect_rho=collect_rho,
            factor_terms=factor_terms, cse_batch_size=cse_batch_size, cse_simplify=cse_simplify,
            hoist_rho_independent=hoist_rho_independent, reuse=RHSreuse,
            symbolic_workers=symbolic_workers, max_symbolic_workers=max_symbolic_workers)
    else:
        compiled, rhs_status = compiled_rhs, "provided"
    stage_s = (dict(compiled.timings) if rhs_status == "miss"
               else {"ode_prepare_s": 0.0, "ode_optimize_s": 0.0})
    const_vals = _runtime_const_values(compiled, runtime_consts, const_values,
                                       use_defaults=compiled_rhs is not None)

    # 2) HIP source of this solver / output variant
    variant = _kernel_variant(compiled, spec, mode, uses_rho0_values, return_time_trace,
                              _state_trace, fp64, time_table, block_size, unroll,
                              tuple(hiprtc_options) + _as_tuple(nvrtc_options))
    emit_start = time.time()
    def emit_for(groups):
        key = ("source", groups) + variant
        src = _KERNEL_CACHE.get(key) if RHSreuse else None
        if src is None:
            src = _emit_source(compiled, spec, mode, uses_rho0_values, return_time_trace,
                               _state_trace, fp64, time_table, block_size, unroll, groups)
            if RHSreuse:
                _KERNEL_CACHE[key] = src
        return src

    if lanes_per_point == "auto":
        # with block_size="auto" the tuner measures every grouping instead
        points = 1
        for v in var_arrays.values():
            points *= int(v.numel()) if hasattr(v, "numel") else int(np.asarray(v).size)
        lanes_per_point = (1 if block_size == "auto" else
                           auto_lanes_per_point(spec, compiled.state_size, fp64,
                                                compiled.plan.nnz if compiled.plan.affine
                                                else None, points,
                                                len(compiled.plan.time_atoms)))
    if lanes_per_point > 1 and not grouped_applies(spec, compiled.state_size, fp64,
                                                   lanes_per_point,
                                                   len(compiled.plan.time_atoms)):
        raise ValueError(f"lanes_per_point={lanes_per_point} needs a solver with a grouped "
                         f"kernel ({', '.join(sorted(GROUPED_TEMPLATES))}) and a state "
                         f"that splits into {lanes_per_point} row blocks whose LDS copy fits "
                         f"(state size {compiled.state_size}).")
    cached_source = RHSreuse and ("source", lanes_per_point) + variant in _KERNEL_CACHE
    source = emit_for(lanes_per_point)
    emit_s = 0.0 if cached_source else time.time() - emit_start
    if rhs_status != "miss":
        emit_s = 0.0
    stage_s["emit_s"] = emit_s
    if Actual_Kernel_Save:
        path = (Actual_Kernel_Save if isinstance(Actual_Kernel_Save, str)
                else os.path.splitext(os.path.abspath(__import__("sys").argv[0] or "gqis"))[0]
                + "_Kernel.hip")
        with open(path, "w") as fh:
            fh.write(source)
        print(f"Generated kernel saved to: {path}")

    # 3) device work
    from . import device
    outputs = device.run_sweep(
        source=source, emit=None if mixed else emit_for, lanes_per_point=lanes_per_point,
        variant=variant, spec=spec, compiled=compiled, mode=mode,
        var_arrays=var_arrays, rho0_values=rho0_values, const_vals=const_vals, fp64=fp64,
        dt=dt, t_in=t_in, num_steps=num_steps, warmup_steps=warmup_steps,
        num_trace=num_trace, trace_stride=trace_stride, trace_offset=trace_offset,
        state_trace=_state_trace, time_table=time_table, block_size=block_size,
        options=tuple(hiprtc_options) + _as_tuple(nvrtc_options), reuse=RHSreuse,
        collect=collect, prepare=_prepare)
    if _prepare:
        from .prepared import PreparedSweep

        def finish(outs, *, return_device, ignore_non_finite_output, total_start):
            return _finish(outs, stage_s, rhs_status, spec, solver_frequency, unroll, False,
                           False, return_device, return_time_trace, False,
                           ignore_non_finite_output, _ode_size, num_trace, trace_stride,
                           trace_offset, fp64, t_in, dt, total_start)
        return PreparedSweep(outputs, compiled=compiled, spec=spec, const_values=const_values,
                             runtime_consts=runtime_consts,
                             use_defaults=compiled_rhs is not None,
                             sweep_symbols=tuple(var_arrays),
                             uses_rho0_values=uses_rho0_values, finish=finish)
    return _finish(outputs, stage_s, rhs_status, spec, solver_frequency, unroll, collect, timings,
                   return_device, return_time_trace, return_timing_info, ignore_non_finite_output,
                   _ode_size, num_trace, trace_stride, trace_offset, fp64, t_in, dt, total_start)


def _finish(outputs, stage_s, rhs_status, spec, solver_frequency, unroll, collect, timings,
            return_device, return_time_trace, return_timing_info, ignore_non_finite_output,
            _ode_size, num_trace, trace_stride, trace_offset, fp64, t_in, dt, total_start):
    """Finite checks, host copies, timing dictionary and return tuple
    (``gqis/solver.py:2347-2391``)."""
    flags = outputs["status"]
    if flags & 1 and not ignore_non_finite_output:
        raise RuntimeError(f"Non-finite values detected in "
                           f"{'mesolve_2D' if _ode_size is None else 'odesolve_2D'} output. "
                           "Try reducing dt (increase time samples) or check model "
                           "parameters/RHS.")
    if flags & 2 and not ignore_non_finite_output:
        raise RuntimeError(f"Non-finite values detected in "
                           f"{'mesolve_2D' if _ode_size is None else 'odesolve_2D'} time trace. "
                           "Try reducing dt (increase time samples) or check model "
                           "parameters/RHS.")
    result = outputs["result"]
    trace = outputs.get("trace")
    if not return_device:
        result = result.cpu().numpy()
        trace = trace.cpu().numpy() if trace is not None else None

    info = None
    if collect:
        prep = stage_s["ode_prepare_s"] + stage_s["ode_optimize_s"] + stage_s["emit_s"]
        info = {"rhs_stage_s": outputs["setup_s"] + prep, "gpu_kernel_s": outputs["gpu_s"],
                "rhs_codegen_s": prep, "total_s": time.time() - total_start,
                "ode_prepare_s": stage_s["ode_prepare_s"],
                "ode_optimize_s": stage_s["ode_optimize_s"], "code_emit_s": stage_s["emit_s"],
                "kernel_build_s": outputs["build_s"], "cached_rhs": rhs_status,
                "cached_kernel": {0: "miss", 1: "hit", 2: "disk"}[int(outputs["cached_kernel"])],
                "solver": spec.name, "solver_frequency": solver_frequency, "unroll": unroll,
                "block_size": outputs.get("block_size"),
                "lanes_per_point": outputs.get("lanes_per_point", 1),
                "autotune_s": outputs.get("autotune_s", 0.0),
                "kernel": outputs.get("kernel")}
        if outputs.get("autotune"):
            info["autotune"] = outputs["autotune"]
        if outputs.get("adaptive_steps") is not None:
            counts = outputs["adaptive_steps"]
            info["adaptive_steps"] = {"accepted": counts[..., 0], "rejected": counts[..., 1]}
        if timings:
            _print_timings("mesolve_2D" if _ode_size is None else "odesolve_2D", info, timings)
    if return_time_trace:
        steps = np.arange(num_trace, dtype=np.float64 if fp64 else np.float32) * trace_stride
        trace_t = t_in + (steps + 1.0 + trace_offset) * dt
        out = (result, trace, trace_t)
        return (*out, info) if return_timing_info else out
    if return_timing_info:
        return result, info
    return result


SPLIT_F16_MAX_ENTRY = 1.0e4     # largest |H_eff| entry routed to FP16 hi/lo products (max 65504)


def _try_hilbert(H, Drive, Col_Ops, mean_operator, var_arrays, const_values, runtime_consts,
                 keep_symbolic_consts, auto_runtime_consts, rho0, rho0_values, spec, mode, fp64,
                 trace, liouville, dt, t_in, num_steps, warmup_steps, block_size, collect, options,
                 prepare=False, num_trace=0, trace_stride=1, trace_offset=0):
    """Run the matrix-core (Hilbert-space) path when it applies; None otherwise (``prepare``:
    return its launch plan instead of running it)."""
    from . import hilbert
    n = H.shape[0] if hasattr(H, "shape") else 0
    forced = liouville.startswith("mfma") or liouville == "sparse"
    sparse = liouville == "sparse"
    split = {"mfma_f16": "f16", "mfma_bf16": "bf16"}.get(liouville)
    # measured on gfx950 (256^2, 2000 steps): for a dense H_eff (hilbert.dense_enough) the
    # matrix-core forms are fastest from 16 levels in both precisions and at 8 levels in FP64
    # (52 ms against 77 ms with 8 lanes per point); a sparse one (the 8-level chain, the
    # multi-qubit chain: 3-5 nonzeros per row) runs 3-6x faster on the per-point kernels below
    # 32 levels.  At 32 levels the per-point form needs 1023 symbolic equations and a very large
    # kernel, so the Hilbert-space form is taken whenever the model fits it: the dense
    # matrix-core product, or the sparse form for a sparse H_eff.
    wanted = forced or (liouville == "auto" and (n >= 16 or (fp64 and n >= 8)))
    reasons = []
    if n not in hilbert.SUPPORTED_N:
        reasons.append(f"N = {n} (supported: {hilbert.SUPPORTED_N})")
    if spec.name != "rk4" and not sparse and liouville != "auto":
        reasons.append(f"solver {spec.name} (RK4 only; the sparse form takes every method)")
    if spec.adaptive and spec.options.get("state_fp64") and not fp64:
        reasons.append("adaptive_options state_fp64 (the Hilbert-space forms keep the state in "
                       "the kernel's precision)")
    if mode not in ("mean", "final", "final_rho"):
        reasons.append(f"output mode {mode}")
    if not wanted or (liouville == "auto" and spec.name != "rk4" and n < 32):
        return None          # auto picks the sparse form (the one for other methods) at 32 levels
    if reasons:
        if forced:
            raise ValueError(f"liouville='{liouville}' does not apply: " + ", ".join(reasons))
        return None
    start = time.time()
    model = _prepare_model(H, Drive, Col_Ops, mean_operator, sweep_symbols=tuple(var_arrays),
                           const_values=const_values, runtime_consts=runtime_consts,
                           keep_symbolic_consts=keep_symbolic_consts,
                           auto_runtime_consts=auto_runtime_consts, rho0=rho0, ode_size=None)
    hm = hilbert.extract(model, _drive_map(model["Drive"]))
    if hm is None:
        if forced:
            raise ValueError(f"liouville='{liouville}' needs H = constant + diagonal parameter/time "
                             "terms, constant diagonal or single-entry collapse operators and a "
                             "diagonal observable.")
        return None
    if liouville == "auto" and not hilbert.dense_enough(hm):
        if n < 32:
            return None      # a sparse H_eff runs faster on the per-point kernels
        # at 32 levels the wavefront-per-point sparse form: the 5-qubit chain (5 nonzeros per
        # row) takes 29.7 ns per point-step in FP32 against 31.2 with FP16 hi/lo products and
        # 51 against 117 ns in FP64 (128^2, 200 RK4 steps)
        sparse = True
    if spec.name != "rk4" and not sparse:
        return None          # the dense matrix-core forms integrate with RK4 only
    values = dict(model["defaults"])
    values.update(runtime_consts or {})
    rt = [float(values[s]) for s in model["runtime_const_syms"]]
    if liouville == "auto" and not fp64 and np.max(np.abs(hm.h_eff)) <= SPLIT_F16_MAX_ENTRY \
            and not sparse:
        # FP32 mode: FP16 hi/lo products reach the FP32 product's error at 1.4x its speed;
        # the constant operator must stay well inside the FP16 range
        split = "f16"
    prep = time.time() - start
    outputs = hilbert.run(hm, mode=mode, var_arrays=var_arrays, rho0_values=rho0_values,
                          runtime_values=rt, dt=dt, t_in=t_in, num_steps=num_steps,
                          warmup_steps=warmup_steps, block_size=block_size, collect=collect,
                          options=options, fp64=fp64, split=split, prepare=prepare,
                          num_trace=num_trace if trace else 0, trace_stride=trace_stride,
                          trace_offset=trace_offset, sparse=sparse, spec=spec)
    if prepare:
        outputs["runtime_defaults"] = dict(model["defaults"])
    stage_s = {"ode_prepare_s": prep, "ode_optimize_s": 0.0, "emit_s": 0.0}
    return outputs, stage_s, "miss"


def _as_tuple(options):
    if options is None:
        return ()
    if isinstance(options, str):
        return (options,)
    return tuple(options)


def _kernel_variant(compiled, spec, mode, uses_rho0_values, trace, state_trace, fp64, time_table,
                    block_size, unroll, options):
    adaptive_key = tuple(sorted((k, v) for k, v in spec.options.items()
                                if k in ("count_steps", "state_fp64", "persistent")))
    return (compiled.cache_key, spec.name, spec.template, adaptive_key, mode,
            bool(uses_rho0_values),
            bool(trace), bool(state_trace), bool(fp64), bool(time_table),
            block_size if block_size == "auto" else int(block_size),
            bool(unroll), options)


LDS_BYTES_PER_CU = 160 * 1024      # gfx950


GROUPED_AUTO_MIN_STATES = 48

GROUPED_TEMPLATES = {"rk4": "grouped_rk4.hip", "tsit5a": "grouped_adaptive.hip",
                     "dop853a": "grouped_adaptive.hip"}
GROUPED_TEMPLATES.update({name: "grouped_tableau.hip" for name in
                          ("dp5", "tsit5", "dop853", "alshina6", "anas5", "lserk4", "lserk4f",
                           "ab5")})


def auto_lanes_per_point(spec, d, fp64, nnz, points, natoms=0):
    """Default wavefronts per point (1 = the one-lane kernel), from measurements on gfx950 of RK4
    and Tsit5A over chain, random and qubit-resonator models of 63 to 399 states:

    * rows with many coefficients (``nnz`` >= 16 d, e.g. a dense 8-level H) or a non-affine
      RHS (``nnz`` None): 8 lanes from 48 states (the one-lane kernel spills; 1.2x FP32, 1.8x
      FP64 at 63 states);
    * sparse rows, RK4 and the 2N-storage LSRK4 (three state vectors; at 63 FP32 states one
      lane takes 15.8 ms against 30.4 with 8): one lane while a point's state is at most 128
      words (FP32 words; FP64
      counts two) and the sweep gives every SIMD a wavefront (65536 points) -- the grouped
      kernels' LDS round trips cost more than they save there (1.4x at 63 FP32 states) --
      otherwise the fewest lanes that leave each wavefront at most 50 words of a vector and
      give a compute unit at least 8 wavefronts (work-groups per CU are limited by their LDS
      copies of the state): 4 lanes at 143 FP32 states, 8 from 195 to 399 (at 195, 4 lanes
      fit one work-group per CU and take 3.74 s against 2.81 s);
    * sparse rows, adaptive solvers: one lane up to 64 words on such a sweep (Tsit5A on
      1024 x 1024 points, 63 FP32 states: 0.347 s against 0.490-0.664 s with 2, 4 or 8 lanes
      on the 8-level resonator, 0.344 s against 0.389-0.568 s on the 3-qubit chain), else as
      the fixed-step tableaux;
    * sparse rows, other solvers (more stage vectors): 8 lanes from 48 states, 16 where 8 would
      leave a wavefront more than 36 words of a vector or, fixed-step, more than 160 words of
      stage vectors (the 4-qubit chain, 255 FP32 states: DOP853 6.7 against 35.4 ns per
      point-step with 8 lanes, which spill; Tsit5 1.17 against 1.59)."""
    words = 2 if fp64 else 1
    if d < GROUPED_AUTO_MIN_STATES:
        return 1
    sparse = nnz is not None and nnz < 16 * d
    if sparse and spec.adaptive and d * words <= 64 and points >= 65536:
        groups = 1
    elif not sparse or spec.name not in ("rk4", "lserk4", "lserk4f"):
        rw = -(-d // 8) * words
        many = not spec.adaptive and stage_vectors(spec) * rw > 160
        groups = 16 if sparse and (rw > 36 or many) else 8
    elif d * words <= 128 and points >= 65536:
        groups = 1
    else:
        def fits(g):
            nbuf = grouped_buffers(spec, d, fp64, g, natoms)
            per_cu = LDS_BYTES_PER_CU // max(1, nbuf * d * 64 * 4 * words)
            return nbuf and -(-d // g) * words <= 50 and g * per_cu >= 8
        groups = next((g for g in (2, 4, 8) if fits(g)), 16)
    while groups > 1 and not grouped_applies(spec, d, fp64, groups, natoms):
        groups //= 2
    return groups


def stage_vectors(spec):
    """State-sized vectors a fixed-step method keeps per lane: the stages and the state (the
    2N-storage LSRK4: state, register and stage; AB5: five past derivatives and the state)."""
    if spec.name in ("lserk4", "lserk4f"):
        return 3
    if spec.name == "ab5":
        return 6
    return spec.rhs_evaluations + 1


def grouped_buffers(spec, d, fp64, groups, natoms=0):
    """LDS copies of the state a grouped kernel keeps: 2 (alternating, one barrier per stage)
    when both fit one compute unit, 1 (a barrier before each write as well) when one does, 0
    when no grouped kernel applies -- the solver has no grouped template, the state does not
    split into ``groups`` non-empty row blocks, or one copy (d x 64 reals) does not fit.  The
    adaptive kernels also keep their error-norm partials, the FP64 time atoms (``natoms``) of
    up to 12 stages, double buffered, and two rows of one vote word per wavefront."""
    if spec.name not in GROUPED_TEMPLATES or groups < 2:
        return 0
    rw = -(-d // groups)
    if (groups - 1) * rw >= d:
        return 0
    copy = d * 64 * (8 if fp64 else 4)
    extra = ((groups * 64 * 8 + 2 * 12 * max(1, natoms) * 64 * 8 + 2 * groups * 4)
             if spec.adaptive else 0)
    for n in (2, 1):
        if n * copy + extra <= LDS_BYTES_PER_CU:
            return n
    return 0


def grouped_applies(spec, d, fp64, groups, natoms=0):
    """A grouped kernel exists for this solver and state (``grouped_buffers`` > 0)."""
    return grouped_buffers(spec, d, fp64, groups, natoms) > 0


def _emit_source(compiled, spec, mode, uses_rho0_values, trace, state_trace, fp64, time_table,
                 block_size, unroll, groups=1):
    defines = {
        "GQ_OUT_MEAN": int(mode == "mean"),
        "GQ_OUT_FINAL": int(mode == "final"),
        "GQ_OUT_FINAL_RHO": int(mode == "final_rho"),
        "GQ_OUT_MEAN_STATE": int(mode == "mean_state"),
        "GQ_TRACE": (2 if state_trace else 1) if trace else 0,
        "GQ_RHO0_VALUES": int(bool(uses_rho0_values)),
        "GQ_TIME_TABLE": int(bool(time_table) and not spec.adaptive),
    }
    if spec.template in ("fixed_rk4.hip", "fixed_rk4_mixed.hip"):
        templates = ["rk4_table.hip", spec.template]
    elif spec.template == "fixed_tableau.hip":
        templates = ["tableau_table.hip", "fixed_tableau.hip"]
    else:
        templates = ["step_control.hip", spec.template] if spec.adaptive else [spec.template]
    if groups > 1:
        # the fixed-step grouped kernels share their output and LDS helpers
        grouped = GROUPED_TEMPLATES[spec.name]
        templates = templates[:-1] + ([] if spec.adaptive else ["grouped_common.hip"]) + [grouped]
    from . import tableaux
    extra = ""
    if spec.adaptive:
        extra = tableaux.emit_arrays(tableaux.DOP853 if spec.name == "dop853a" else tableaux.TSIT5)
        defines["GQ_METHOD_DOP853"] = int(spec.name == "dop853a")
        defines["GQ_COUNT_STEPS"] = int(bool(spec.options.get("count_steps")))
        state64 = spec.options.get("state_fp64")
        defines["GQ_STATE_FP64"] = int(bool(fp64 or state64))
        defines["GQ_PERSISTENT"] = int(bool(spec.options.get("persistent", True)))
    elif spec.template == "fixed_tableau.hip":
        extra = tableaux.emit_fixed(spec.name)
    elif spec.template == "fixed_rk4_mixed.hip":
        extra = codegen.emit_model_f64_stage(compiled.plan)
        defines["GQ_TABLE_F64"] = 1
    if groups > 1:
        block_size = 64 * groups
        nbuf = grouped_buffers(spec, compiled.state_size, fp64, groups,
                               len(compiled.plan.time_atoms))
        if not nbuf:
            raise ValueError(f"no grouped {spec.name} kernel for {compiled.state_size} states in "
                             f"{groups} row blocks")
        defines["GQ_NBUF"] = nbuf
    return codegen.assemble(compiled.plan, templates, fp64=fp64, acc64=True,
                            block_size=256 if block_size == "auto" else block_size,
                            defines=defines,
                            rho0_exprs=compiled.rho0_exprs, extra=extra, groups=groups)


def _print_timings(api_name, info, mode):
    """Report a call's timing dictionary on stdout: one summary line, or with ``mode="full"``
    one line per stage as well."""
    rhs_origin = {"miss": "derived", "hit": "cached", "provided": "precompiled"}[info["cached_rhs"]]
    kernel_origin = "built" if info["cached_kernel"] == "miss" else "cached"

    def fmt(seconds):
        return f"{seconds:.2f} s" if seconds >= 1.0 else f"{1e3 * seconds:.1f} ms"
    lines = [f"{api_name} [{info['solver']}]: host {fmt(info['rhs_stage_s'])} (RHS {rhs_origin}, "
             f"kernel {kernel_origin}) | GPU {fmt(info['gpu_kernel_s'])} | "
             f"total {fmt(info['total_s'])}"]
    if mode == "full":
        stages = {"symbolic model": "ode_prepare_s", "CSE and hoisting": "ode_optimize_s",
                  "HIP source": "code_emit_s", "hiprtc build": "kernel_build_s",
                  "GPU kernel": "gpu_kernel_s"}
        lines += [f"    {label:.<20}{fmt(info[key]):>11}" for label, key in stages.items()]
    print("\n".join(lines))