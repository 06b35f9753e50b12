"""Grouped optional arguments of ``mesolve_2D`` / ``odesolve_2D``.

The group names, their members and the error messages are the reference's interface
(``gqis/_options.py:6-57``): a group is a dict keyed by the flat keyword names, and a name may be
given flat or in its group, not both.  The groups here are a superset of the reference's (the
names added by this implementation are marked), so every reference call still works.
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
                # added by this implementation:
                "hiprtc_options", "time_table", "block_size", "liouville", "lanes_per_point"),
    "diagnostics": ("timings", "return_timing_info", "Actual_Kernel_Save", "beep_on_error",
                    "ignore_non_finite_output"),
}

# Every option belongs to exactly one group.
_GROUP_OF = {name: group for group, names in OPTION_GROUPS.items() for name in names}

# Defaults of these types are compared by value: a flat ``fp64=False`` counts as not given.
_VALUE_DEFAULTS = (bool, int, float, str, tuple)


def _given(value, default):
    """Whether a flat keyword was set to something other than its signature default."""
    if value is default:
        return False
    if isinstance(default, _VALUE_DEFAULTS) and type(value) is type(default):
        return value != default
    return True


def _grouped_options(groups):
    """Yield ``(group, name, value)`` for every entry of every group dict, checking that each
    group is a mapping and each name is one of that group's options."""
    for group, values in groups.items():
        if values is None:
            continue
        if not isinstance(values, Mapping):
            raise TypeError(f"{group} must be a dict of option names and values.")
        for name, value in values.items():
            if _GROUP_OF.get(name) != group:
                members = OPTION_GROUPS[group]
                raise ValueError(f"Unknown {group} option '{name}'. "
                                 f"Valid options: {', '.join(members)}.")
            yield group, name, value


def fold_option_groups(func, flat, groups, forwarded=None):
    """Move the grouped options into the flat keyword dict ``flat`` of ``func``.

    ``groups`` maps group name -> dict or None.  Names that are not parameters of ``func`` go
    to ``forwarded`` (its ``**kwargs`` dict), when it has one.  Returns ``(flat, forwarded)``.
    Raises ``TypeError`` for a group that is not a mapping and ``ValueError`` for an unknown
    name or a name given both flat and grouped (``gqis/_options.py:30-57``).
    """
    defaults = {name: p.default for name, p in inspect.signature(func).parameters.items()}
    for group, name, value in _grouped_options(groups):
        if name in defaults:
            target, clash = flat, _given(flat[name], defaults[name])
        elif forwarded is not None:
            target, clash = forwarded, name in forwarded
        else:
            raise ValueError(f"{func.__name__} does not accept the {group} option '{name}'.")
        if clash:
            raise ValueError(f"'{name}' is given both directly and in {group}; use one.")
        target[name] = value
    return flat, forwarded


def split_groups(call):
    """Remove the group arguments from a call's keyword dict and return them."""
    return {group: call.pop(group) for group in OPTION_GROUPS}
