"""Solver registry: names, aliases and the metadata the API reports.

Same public names and aliases as the reference (``gqis/cuda_solvers.py:993-1080``), plus the
new adaptive ``dop853a``.  Each spec names the hand-written HIP template that implements it.
"""
from dataclasses import dataclass, field
from fractions import Fraction
import math
import struct


@dataclass(frozen=True)
class SolverSpec:
    name: str
    label: str
    order: int
    rhs_evaluations: int
    template: str                 # kernels/<template>
    kernel: str                   # __global__ name in the template
    adaptive: bool = False
    parameters: tuple = ()        # runtime scalars (solver_param[])
    options: dict = field(default_factory=dict)
    stage_times_per_step: int = 2   # rows of the time table per step (fixed-step only)


TSIT5A_DEFAULTS = {
    "step_rtol": 1e-5,      # per-step relative tolerance of the local error estimate
    "step_atol": 1e-5,      # per-step absolute tolerance
    "max_substeps": 64,     # accepted + rejected steps allowed per output interval
    "fp64_time": True,      # time in FP64 (the only mode of this implementation; kept for parity)
    "count_steps": False,   # return accepted/rejected step counts in the timing info
}
DOP853A_DEFAULTS = {
    "step_rtol": 1e-8,
    "step_atol": 1e-10,
    "max_substeps": 4096,
    "fp64_time": True,
    "count_steps": False,
}
ADAPTIVE_EXTRA_OPTIONS = {
    "state_fp64": None,     # FP32 mode: keep the state and stage sums in FP64 (RHS in FP32)
    "persistent": True,     # persistent work-stealing launch
}

_SPECS = {
    "rk4": SolverSpec("rk4", "RK4", 4, 4, "fixed_rk4.hip", "gq_rk4"),
    "lserk4": SolverSpec("lserk4", "LSRK4", 4, 5, "fixed_tableau.hip", "gq_fixed"),
    "lserk4f": SolverSpec("lserk4f", "LSRK4F", 4, 5, "fixed_tableau.hip", "gq_fixed"),
    "dp5": SolverSpec("dp5", "DP5", 5, 6, "fixed_tableau.hip", "gq_fixed"),
    "ab5": SolverSpec("ab5", "AB5", 5, 1, "fixed_tableau.hip", "gq_fixed"),
    "anas5": SolverSpec("anas5", "Anas5", 5, 6, "fixed_tableau.hip", "gq_fixed"),
    "tsit5": SolverSpec("tsit5", "Tsit5", 5, 6, "fixed_tableau.hip", "gq_fixed"),
    "alshina6": SolverSpec("alshina6", "Alshina6", 6, 7, "fixed_tableau.hip", "gq_fixed"),
    "dop853": SolverSpec("dop853", "DOP853", 8, 12, "fixed_tableau.hip", "gq_fixed"),
    "tsit5a": SolverSpec("tsit5a", "Tsit5A", 5, 6, "adaptive.hip", "gq_adaptive", adaptive=True),
    "dop853a": SolverSpec("dop853a", "DOP853A", 8, 12, "adaptive.hip", "gq_adaptive",
                          adaptive=True),
}

_ALIASES = {
    "classical_rk4": "rk4",
    "carpenter_kennedy": "lserk4",
    "carpenter_kennedy_2n54": "lserk4",
    "dormand_prince": "dp5",
    "dormand_prince_5": "dp5",
    "rk45": "dp5",
    "anastassi5": "anas5",
    "anastassi_simos": "anas5",
    "adams_bashforth_5": "ab5",
    "alshina": "alshina6",
    "dormand_prince_853": "dop853",
    "dop_853": "dop853",
    "adaptive_dop853": "dop853a",
}


def available_solvers():
    """Canonical solver names in display order."""
    return tuple(_SPECS)


def normalize_solver_name(solver):
    """Canonical lower-case name, or ValueError (``gqis/cuda_solvers.py:1031-1038``)."""
    name = str(solver).strip().lower().replace(" ", "_").replace("-", "_")
    name = _ALIASES.get(name, name)
    if name not in _SPECS:
        raise ValueError(f"Unsupported solver '{solver}'. Use one of: {', '.join(_SPECS)}")
    return name


def _float32(x):
    return struct.unpack("f", struct.pack("f", float(x)))[0]


def _exact(x):
    """The rational with a denominator below 10^5 that the double ``x`` rounds (the tableau
    entries are such rationals), as an mpmath number."""
    import mpmath as mp
    q = Fraction(x).limit_denominator(100000)
    return mp.mpf(q.numerator) / q.denominator


def _stability_value(tab, z, p):
    """R(z) = 1 + z b.k of an explicit tableau whose rows ``a1`` add p times their entries
    (mpmath, exact rational coefficients)."""
    import mpmath as mp
    rows = [[_exact(a) for a in r] for r in tab["a"]]
    for s, extra in tab["a1"].items():
        rows[s] = [a + p * _exact(e) for a, e in zip(rows[s], extra)]
    k = []
    for row in rows:
        k.append(1 + z * mp.fsum(a * kj for a, kj in zip(row, k)))
    return 1 + z * mp.fsum(_exact(b) * kj for b, kj in zip(tab["b"], k))


def anas5_coefficient(frequency, step_size):
    """Fitted coefficient p of the Anas5(w) tableau for v = w*dt (the reference's
    ``gqis/cuda_solvers.py:971-990``): the value that makes the method's phase lag vanish on
    y' = i w y, i.e. arg R(iv) = v.  R is affine in p (p enters only the last stage), so
    Im(R(iv) e^(-iv)) = 0 is one linear equation in p, solved here in 60-digit arithmetic.
    At v = 0 the unfitted value -800/1071 is returned."""
    import mpmath as mp
    from .tableaux import ANAS5
    if step_size is None:
        return ()
    if not math.isfinite(step_size) or step_size <= 0.0:
        raise ValueError("anas5 requires a positive finite step_size")
    frequency = 1.0 if frequency is None else float(frequency)
    if not math.isfinite(frequency) or frequency < 0.0:
        raise ValueError("solver_frequency must be a finite non-negative angular frequency")
    with mp.workdps(60):
        v = mp.mpf(str(frequency)) * mp.mpf(str(step_size))
        if v == 0:
            return (float(mp.mpf(-800) / 1071),)
        z, turn = mp.mpc(0, v), mp.expj(-v)
        r0 = _stability_value(ANAS5, z, 0)
        r1 = _stability_value(ANAS5, z, 1) - r0
        return (float(-mp.im(r0 * turn) / mp.im(r1 * turn)),)


def get_solver_spec(solver, *, frequency=None, step_size=None, adaptive_options=None,
                    fp64=False):
    """Spec with the runtime parameters of one call (``gqis/cuda_solvers.py:1041-1075``)."""
    spec = _SPECS[normalize_solver_name(solver)]
    if not spec.adaptive:
        if adaptive_options:
            raise ValueError("adaptive_options apply only to solver='tsit5a' or 'dop853a'.")
        params = anas5_coefficient(frequency, step_size) if spec.name == "anas5" else ()
        return SolverSpec(spec.name, spec.label, spec.order, spec.rhs_evaluations, spec.template,
                          spec.kernel, False, tuple(params), {})
    defaults = dict(TSIT5A_DEFAULTS if spec.name == "tsit5a" else DOP853A_DEFAULTS)
    allowed = set(defaults) | set(ADAPTIVE_EXTRA_OPTIONS)
    unknown = set(adaptive_options or {}) - allowed
    if unknown:
        raise ValueError(f"Unknown adaptive_options {sorted(unknown)}; use {sorted(allowed)}.")
    options = {**defaults, **ADAPTIVE_EXTRA_OPTIONS, **(adaptive_options or {})}
    rtol, atol = float(options["step_rtol"]), float(options["step_atol"])
    if not (rtol >= 0.0 and atol >= 0.0 and rtol + atol > 0.0 and math.isfinite(rtol + atol)):
        raise ValueError("step_rtol and step_atol must be finite, non-negative and not both zero.")
    max_substeps = int(options["max_substeps"])
    if max_substeps < 1:
        raise ValueError("max_substeps must be at least 1.")
    # The FP32 reference kernel received dt in FP32 and passed its rounding error separately
    # (gqis/cuda_solvers.py:1069-1072); this implementation passes dt in FP64, the last
    # parameter is kept for parity.
    dt_low = 0.0
    if step_size is not None and not fp64:
        dt_low = float(step_size) - _float32(step_size)
    return SolverSpec(spec.name, spec.label, spec.order, spec.rhs_evaluations, spec.template,
                      spec.kernel, True, (rtol, atol, float(max_substeps), dt_low), options)
