"""GQIS sweep path on the AMD Instinct MI355X (gfx950).

Same public interface as the reference package (``gqis/__init__.py:5-29``): the solver modules
load lazily, so ``import gqis`` needs neither torch nor a GPU.
"""
from importlib import import_module

__all__ = ["build_independent_rho", "build_reduced_lindblad_rhs", "mesolve_2D", "odesolve_2D",
           "compile_lindblad_rhs", "compile_ode_rhs", "CompiledRHS", "set_kernel_cache_dir",
           "prepare_sweep"]
_ODE_API_NAMES = {"odesolve_2D", "compile_ode_rhs"}

__version__ = "0.3.0+mi355x"


def __getattr__(name):
    if name not in __all__:
        raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
    module = import_module(".gqis_ode_api" if name in _ODE_API_NAMES else ".solver", __name__)
    value = getattr(module, name)
    globals()[name] = value
    return value
