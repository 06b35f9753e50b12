"""Grouped optional arguments of ``mesolve_2D`` / ``odesolve_2D``.

Mirrors the reference's option groups (``gqis/_options.py:6-57``): a group is a dict that uses
the flat keyword names; a name may be given flat or in its group, not both.  The groups here
are a superset of the reference's (new names are marked), so every reference call still works.
"""
from collections.abc import Mapping
import inspect

OPTION_GROUPS = {
    "consts": ("const_values", "runtime_consts", "keep_symbolic_consts", "auto_runtime_consts"),
    "initial": ("rho0", "rho0_var_arrays", "rho0_values", "y0_values", "t_in"),
    "output": ("warmup_time", "return_time_trace", "time_trace_every",
               "time_trace_samples_per_period", "solver_samples_per_period", "return_device"),
    "solver_options": ("solver", "fp64", "solver_frequency", "adaptive_options"),
    "codegen": ("RHSreuse", "kernel_template_file", "unroll", "pre_expand", "collect_rho",
                "factor_terms", "cse_batch_size", "cse_simplify", "hoist_rho_independent",
                "symbolic_workers", "max_symbolic_workers", "nvrtc_options",
                # new in this implementation:
                "hiprtc_options", "time_table", "block_size", "liouville", "lanes_per_point"),
    "diagnostics": ("timings", "return_timing_info", "Actual_Kernel_Save", "beep_on_error",
                    "ignore_non_finite_output"),
}


def _unchanged(value, default):
    """True when a flat argument still holds its signature default."""
    if value is default:
        return True
    simple = (bool, int, float, str, tuple)
    return isinstance(default, simple) and type(value) is type(default) and value == default


def apply_option_groups(func, args, groups, extra=None):
    """Fold ``groups`` (name -> dict or None) into the flat keyword dict ``args`` of ``func``.

    ``extra`` is the ``**kwargs`` dict of ``func`` when it forwards further options.  Raises
    ``TypeError`` for a group that is not a mapping and ``ValueError`` for an unknown name or an
    option given both flat and grouped (``gqis/_options.py:30-57``).
    """
    params = inspect.signature(func).parameters
    for group, values in groups.items():
        if values is None:
            continue
        if not isinstance(values, Mapping):
            raise TypeError(f"{group} must be a dict of option names and values.")
        allowed = OPTION_GROUPS[group]
        for name, value in values.items():
            if name not in allowed:
                raise ValueError(f"Unknown {group} option '{name}'. "
                                 f"Valid options: {', '.join(allowed)}.")
            if name in params:
                if not _unchanged(args[name], params[name].default):
                    raise ValueError(f"'{name}' is given both directly and in {group}; use one.")
                args[name] = value
            elif extra is not None:
                if name in extra:
                    raise ValueError(f"'{name}' is given both directly and in {group}; use one.")
                extra[name] = value
            else:
                raise ValueError(f"{func.__name__} does not accept the {group} option '{name}'.")
    return args, extra
