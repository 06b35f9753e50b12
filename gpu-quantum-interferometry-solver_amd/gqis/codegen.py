"""HIP source emission for gfx950 (hiprtc), from an ``RHSPlan``.

The generated source is: the ABI header (argument structs), a precision prelude, the model
functions printed from the plan (statics, time atoms, stage values, RHS, observable, initial
state) and one hand-written solver template from ``kernels/``.  This replaces the reference's
CUDA template assembly (``gqis/solver.py:2073-2233``, ``gqis/N_Level_Kernel.cu``); the
reference's FP64 mode is a textual promotion of its FP32 kernel (``gqis/solver.py:2219-2232``),
here the precision is a typedef.
"""
import os
import re

import sympy as sp
from sympy.printing.c import C99CodePrinter
from sympy.printing.precedence import precedence

_PKG = os.path.dirname(os.path.abspath(__file__))
KERNEL_DIR = os.path.join(_PKG, "kernels")


def _header_path():
    """The C-ABI header (include/gqis_hip.h): one struct definition for host and device."""
    cands = [os.path.join(_PKG, "..", "..", "include", "gqis_hip.h"),
             os.path.join(_PKG, "kernels", "gqis_hip.h")]
    for c in cands:
        if os.path.exists(c):
            return os.path.abspath(c)
    raise FileNotFoundError("gqis_hip.h not found next to the package")


def read_kernel_template(name):
    with open(os.path.join(KERNEL_DIR, name)) as fh:
        return fh.read()


class HipPrinter(C99CodePrinter):
    """C++ printer for device code: literals as ``R(x)`` (the kernel's real type), integer
    powers as products, symbol names mapped to kernel variables."""

    def __init__(self, names, literal="R", settings=None):
        super().__init__(settings or {})
        self._names = names
        self._lit = literal

    def doprint(self, expr, assign_to=None):
        """Fold exact irrational constants (sqrt(2), pi, E, log(2), ...) to literals first."""
        expr = sp.sympify(expr)
        consts = {}
        walk = sp.preorder_traversal(expr)
        for sub in walk:
            if sub.is_number and not (sub.is_Rational or sub.is_Float) and sub.is_real:
                consts[sub] = sp.Float(sub.evalf(20), 20)
                walk.skip()
        if consts:
            expr = expr.xreplace(consts)
        return super().doprint(expr, assign_to)

    def _lit_str(self, value):
        return f"{self._lit}({float(value)!r})"

    def _print_Symbol(self, expr):
        name = self._names.get(expr.name)
        if name is None:
            raise ValueError(f"symbol {expr.name!r} has no kernel binding")
        return name

    def _print_Float(self, expr):
        return self._lit_str(expr)

    def _print_Integer(self, expr):
        return self._lit_str(int(expr))

    def _print_Rational(self, expr):
        return self._lit_str(expr.p / expr.q)

    def _print_NumberSymbol(self, expr):
        return self._lit_str(expr.evalf(20))

    def _print_Exp1(self, expr):
        return self._lit_str(sp.E.evalf(20))

    def _print_Pi(self, expr):
        return self._lit_str(sp.pi.evalf(20))

    def _print_ImaginaryUnit(self, expr):
        raise ValueError("complex values cannot be printed into the real kernel")

    def _print_Pow(self, expr):
        base, exp = expr.base, expr.exp
        b = self.parenthesize(base, precedence(expr))
        if exp.is_Integer:
            n = int(exp)
            if n == 0:
                return self._lit_str(1)
            if 0 < n <= 8:
                return "(" + "*".join([b] * n) + ")"
            if -8 <= n < 0:
                return "gq_rcp(" + "*".join([b] * (-n)) + ")"
        if exp == sp.Rational(1, 2):
            return f"sqrt({self._print(base)})"
        if exp == -sp.Rational(1, 2):
            return f"({self._lit}(1.0)/sqrt({self._print(base)}))"
        return f"pow({self._print(base)}, {self._print(exp)})"

    def _print_Piecewise(self, expr):
        out = None
        for value, cond in reversed(expr.args):
            v = self._print(value)
            if cond == True:  # noqa: E712 (SymPy true)
                out = v
            else:
                out = f"(({self._print(cond)}) ? ({v}) : ({out if out is not None else 'R(0.0)'}))"
        return out

    def _print_re(self, expr):
        raise ValueError("re() left in a real kernel expression")

    def _print_im(self, expr):
        raise ValueError("im() left in a real kernel expression")

    # math functions: the C++ overloads pick float or double by argument type
    def _print_Function(self, expr):
        name = expr.func.__name__
        known = {"sin", "cos", "tan", "exp", "log", "sinh", "cosh", "tanh", "asin", "acos",
                 "atan", "atan2", "Abs", "sign", "floor", "ceiling", "Min", "Max"}
        if name not in known:
            return super()._print_Function(expr)
        args = ", ".join(self._print(a) for a in expr.args)
        mapping = {"Abs": "fabs", "ceiling": "ceil", "Min": "fmin", "Max": "fmax",
                   "sign": "gq_sign"}
        return f"{mapping.get(name, name)}({args})"

    _print_sin = _print_cos = _print_tan = _print_exp = _print_log = _print_Function
    _print_sinh = _print_cosh = _print_tanh = _print_asin = _print_acos = _print_Function
    _print_atan = _print_atan2 = _print_Abs = _print_sign = _print_floor = _print_Function
    _print_ceiling = _print_Min = _print_Max = _print_Function


def _names_for(plan):
    """Kernel variable names of every symbol the plan can use."""
    names = {"ParX": "px", "ParY": "py", "t": "t_now"}
    for i, s in enumerate(plan.states):
        names[s.name] = f"y[{i}]"
    for i, s in enumerate(plan.const_syms):
        names[s.name] = f"C[{i}]"
    for sym, _ in plan.statics:
        names[sym.name] = f"S.{sym.name}"
    for i, (sym, _) in enumerate(plan.time_atoms):
        names[sym.name] = f"TT[{i}]"
    for sym, _ in plan.stage:
        names[sym.name] = f"G.{sym.name}"
    for sym, _ in plan.rhs_temps:
        names[sym.name] = sym.name
    for sym, _ in plan.obs_temps:
        names[sym.name] = sym.name
    return names


def _atoms_stage_lines(plan, const_double, const_real):
    """``gq_atoms_stage``: the time atoms at an FP64 time, in the kernel's real type (the
    adaptive kernels' stage times).  Each sin / cos of an affine function of t is evaluated
    without the library function: FP64 kernels by ``gq_sincos64`` (Cody-Waite reduction to
    |r| <= pi/4 and Taylor polynomials, within 2 ulp of the FP64 function), FP32 kernels by
    ``gq_sincos32`` (the same reduction in FP64, the polynomials in FP32): FP32 accuracy, as the
    FP32 RHS, for a fraction of the FP64 evaluation.  Other atoms (t outside such a function)
    are evaluated in FP64."""
    lines = ["__device__ __forceinline__ void gq_atoms_stage(const double t_now, "
             "const real* __restrict__ C, real* __restrict__ TT) {",
             "    (void)t_now; (void)C; (void)TT;"]
    dnames = {"t": "t_now", **const_double}
    dpr = HipPrinter(dnames, literal="DBL")
    rnames = dict(const_real)
    reduced = {}
    for i, (_, e) in enumerate(plan.time_atoms):
        tsym = next((x for x in e.free_symbols if x.name == "t"), None)
        subs = {}
        for f in e.atoms(sp.sin, sp.cos):
            arg = f.args[0]
            if (tsym is not None and arg.is_polynomial(tsym)
                    and sp.degree(arg, tsym) <= 1):
                if f not in reduced:
                    name = f"q{len(reduced)}"
                    reduced[f] = sp.Symbol(name, real=True)
                    rnames[name] = name
                    fn = "cos" if isinstance(f, sp.cos) else "sin"
                    a = dpr.doprint(arg)
                    lines += ["#if GQ_FP64",
                              f"    const real {name} = gq_{fn}64({a});",
                              "#else",
                              f"    const real {name} = gq_{fn}32({a});",
                              "#endif"]
                subs[f] = reduced[f]
        rest = e.xreplace(subs)
        if subs and not any(x.name == "t" for x in rest.free_symbols):
            lines.append(f"    TT[{i}] = {HipPrinter(rnames).doprint(rest)};")
        else:
            lines.append(f"    TT[{i}] = (real)({dpr.doprint(e)});")
    lines += ["}"]
    return lines


def emit_model(plan, rho0_exprs=None):
    """Device functions of one model (the part the reference prints from SymPy,
    ``gqis/solver.py:633-700,1649-1692``)."""
    names = _names_for(plan)
    pr = HipPrinter(names)
    lines = []
    n_static = len(plan.statics)
    n_stage = len(plan.stage)
    n_atoms = len(plan.time_atoms)
    lines.append(f"#define GQ_D {plan.state_size}")
    lines.append(f"#define GQ_NATOMS {n_atoms}")
    lines.append(f"#define GQ_NATOMS_ALLOC {max(1, n_atoms)}")
    lines.append("struct gq_static_t {")
    lines += [f"    real {s.name};" for s, _ in plan.statics] or ["    real unused_;"]
    lines.append("};")
    lines.append("struct gq_stage_t {")
    lines += [f"    real {s.name};" for s, _ in plan.stage] or ["    real unused_;"]
    lines.append("};")
    # statics: once per parameter point
    lines.append("__device__ __forceinline__ void gq_statics(const real px, const real py, "
                 "const real* __restrict__ C, gq_static_t& S) {")
    lines.append("    (void)px; (void)py; (void)C;")
    lines += [f"    S.{s.name} = {pr.doprint(e)};" for s, e in plan.statics]
    lines.append("}")
    # time atoms in FP64 (tables, adaptive FP64 time) and in the kernel's real type
    for fname, rtype, lit in (("gq_atoms_f64", "double", "DBL"), ("gq_atoms_real", "real", "R")):
        anames = dict(names)
        anames["t"] = "t_now"
        for i, s in enumerate(plan.const_syms):
            anames[s.name] = f"({rtype})C[{i}]"
        apr = HipPrinter(anames, literal=lit)
        lines.append(f"__device__ __forceinline__ void {fname}(const {rtype} t_now, "
                     f"const real* __restrict__ C, {rtype}* __restrict__ TT) {{")
        lines.append("    (void)t_now; (void)C; (void)TT;")
        lines += [f"    TT[{i}] = {apr.doprint(e)};" for i, (_, e) in enumerate(plan.time_atoms)]
        lines.append("}")
    lines += _atoms_stage_lines(plan, {s.name: f"(double)C[{i}]" for i, s in
                                       enumerate(plan.const_syms)},
                                {s.name: f"C[{i}]" for i, s in enumerate(plan.const_syms)})
    # stage values: once per distinct stage time
    lines.append("__device__ __forceinline__ void gq_stage(const real px, const real py, "
                 "const real* __restrict__ C, const gq_static_t& S, const real* __restrict__ TT, "
                 "gq_stage_t& G) {")
    lines.append("    (void)px; (void)py; (void)C; (void)S; (void)TT;")
    lines += [f"    G.{s.name} = {pr.doprint(e)};" for s, e in plan.stage]
    lines.append("}")
    # right-hand side
    lines.append("__device__ __forceinline__ void gq_rhs(const real* __restrict__ y, const real px, "
                 "const real py, const real* __restrict__ C, const gq_static_t& S, "
                 "const gq_stage_t& G, real* __restrict__ dy) {")
    lines.append("    (void)px; (void)py; (void)C; (void)S; (void)G;")
    lines += [f"    const real {s.name} = {pr.doprint(e)};" for s, e in plan.rhs_temps]
    lines += [f"    dy[{i}] = {pr.doprint(e)};" for i, e in enumerate(plan.rhs_out)]
    lines.append("}")
    # observable
    lines.append("__device__ __forceinline__ void gq_obs(const real* __restrict__ y, const real px, "
                 "const real py, const real* __restrict__ C, const gq_static_t& S, "
                 "const gq_stage_t& G, real& o_re, real& o_im) {")
    lines.append("    (void)y; (void)px; (void)py; (void)C; (void)S; (void)G;")
    lines += [f"    const real {s.name} = {pr.doprint(e)};" for s, e in plan.obs_temps]
    lines.append(f"    o_re = {pr.doprint(plan.obs_out[0])};")
    lines.append(f"    o_im = {pr.doprint(plan.obs_out[1])};")
    lines.append("}")
    # initial state from expressions (None: |0><0| or zero for ODEs handled by the template)
    lines.append("__device__ __forceinline__ void gq_init_expr(real* __restrict__ y, const real px, "
                 "const real py, const real* __restrict__ C) {")
    lines.append("    (void)px; (void)py; (void)C;")
    if rho0_exprs is None:
        lines.append("    for (int i = 0; i < GQ_D; ++i) y[i] = R(0.0);")
        lines.append("    y[0] = R(1.0);")
    else:
        ipr = HipPrinter(names)
        lines += [f"    y[{i}] = {ipr.doprint(sp.sympify(e))};" for i, e in enumerate(rho0_exprs)]
    lines.append("}")
    return "\n".join(lines)


def emit_model_f64_stage(plan):
    """FP64 statics and stage values for the mixed-precision RK4 form (FP64 parameters, time
    terms and state; FP32 right-hand side): ``gq_statics_d`` / ``gq_stage_d`` on double
    parameters and an FP64 time-atom row, and ``gq_stage_narrow`` rounding each stage value to
    the kernel's FP32 ``gq_stage_t`` once per stage time (structs ``gq_static64_t``, ``gq_stage64_t``)."""
    names = _names_for(plan)
    for i, s in enumerate(plan.const_syms):
        names[s.name] = f"(double)C[{i}]"
    pr = HipPrinter(names, literal="DBL")
    lines = ["struct gq_static64_t {"]
    lines += [f"    double {s.name};" for s, _ in plan.statics] or ["    double unused_;"]
    lines.append("};")
    lines.append("struct gq_stage64_t {")
    lines += [f"    double {s.name};" for s, _ in plan.stage] or ["    double unused_;"]
    lines.append("};")
    lines.append("__device__ __forceinline__ void gq_statics_d(const double px, const double py, "
                 "const real* __restrict__ C, gq_static64_t& S) {")
    lines.append("    (void)px; (void)py; (void)C;")
    lines += [f"    S.{s.name} = {pr.doprint(e)};" for s, e in plan.statics]
    lines.append("}")
    lines.append("__device__ __forceinline__ void gq_stage_d(const double px, const double py, "
                 "const real* __restrict__ C, const gq_static64_t& S, "
                 "const double* __restrict__ TT, gq_stage64_t& G) {")
    lines.append("    (void)px; (void)py; (void)C; (void)S; (void)TT;")
    lines += [f"    G.{s.name} = {pr.doprint(e)};" for s, e in plan.stage]
    lines.append("}")
    lines.append("__device__ __forceinline__ void gq_stage_narrow(const gq_stage64_t& Gd, "
                 "gq_stage_t& G) {")
    lines.append("    (void)Gd; (void)G;")
    lines += [f"    G.{s.name} = (real)Gd.{s.name};" for s, _ in plan.stage]
    lines.append("}")
    lines.append("__device__ __forceinline__ void gq_statics_narrow(const gq_static64_t& Sd, "
                 "gq_static_t& S) {")
    lines.append("    (void)Sd; (void)S;")
    lines += [f"    S.{s.name} = (real)Sd.{s.name};" for s, _ in plan.statics]
    lines.append("}")
    return "\n".join(lines)


def _closure(ordered, seeds):
    """The entries of an ordered [(symbol, expr)] list that ``seeds`` reach (in list order);
    an entry's expression may use earlier entries."""
    need = set(seeds)
    keep = []
    for sym, expr in reversed(ordered):
        if sym in need:
            keep.append((sym, expr))
            need |= expr.free_symbols
    keep.reverse()
    return keep, need


def group_rows(d, groups):
    """Rows of each wave of a point group: contiguous blocks of ceil(d / groups)."""
    rw = -(-d // groups)
    return rw, [(g * rw, min(d, (g + 1) * rw)) for g in range(groups)]


def emit_grouped(plan, groups, rho0_exprs=None):
    """Device code of the grouped kernels: ``groups`` wavefronts share one set of 64
    parameter points, wavefront g owning a contiguous block of state rows.  For every g a
    specialisation ``gq_grp<g>`` holds: its first row and row count, the statics and stage
    values its rows (and, for g = 0, the observable) use, renumbered into small per-wave arrays;
    its RHS rows, reading the stage state from LDS (point-minor, ``Y[j * 64 + lane]``); its rows
    of the initial state.  ``GQ_GROUP_DISPATCH(fn, args)`` calls ``fn<g> args`` for the
    wave-uniform g, so each wave runs code specialised to its rows."""
    d = plan.state_size
    rw, blocks = group_rows(d, groups)
    if groups < 2 or rw < 1 or blocks[-1][0] >= d:
        raise ValueError(f"{groups} groups do not split {d} state rows")
    state_names = {s.name: j for j, s in enumerate(plan.states)}
    parts = ["template <int GI> struct gq_grp;"]
    for g, (lo, hi) in enumerate(blocks):
        exprs = [sp.sympify(e) for e in plan.rhs_out[lo:hi]]
        seeds = set().union(*[e.free_symbols for e in exprs]) if exprs else set()
        obs, obs_temps = [], []
        if g == 0:
            obs = [sp.sympify(e) for e in plan.obs_out]
            obs_temps, reach = _closure(plan.obs_temps, set().union(*[e.free_symbols for e in obs]))
            seeds |= reach
        temps, seeds = _closure(plan.rhs_temps, seeds)
        stage, seeds = _closure(plan.stage, seeds)
        statics, seeds = _closure(plan.statics, seeds)
        names = {"ParX": "px", "ParY": "py", "t": "t_now"}
        for i, c in enumerate(plan.const_syms):
            names[c.name] = f"C[{i}]"
        for i, (sym, _) in enumerate(plan.time_atoms):
            names[sym.name] = f"TT[{i}]"
        for k, (sym, _) in enumerate(statics):
            names[sym.name] = f"Sg[{k}]"
        for k, (sym, _) in enumerate(stage):
            names[sym.name] = f"Gg[{k}]"
        for sym, _ in list(temps) + list(obs_temps):
            names[sym.name] = sym.name
        for s_name, j in state_names.items():
            names[s_name] = f"y{j}"
        pr = HipPrinter(names)
        dev = "__device__ static __forceinline__ void"
        parts.append(f"template <> struct gq_grp<{g}> {{")
        parts.append(f"    static constexpr int R0 = {lo}, NR = {hi - lo}, "
                     f"NS = {max(1, len(statics))}, NG = {max(1, len(stage))};")
        parts.append(f"    {dev} statics(const real px, const real py, "
                     "const real* __restrict__ C, real* __restrict__ Sg) {")
        parts.append("        (void)px; (void)py; (void)C; (void)Sg;")
        parts += [f"        Sg[{k}] = {pr.doprint(e)};" for k, (_, e) in enumerate(statics)]
        parts.append("    }")
        parts.append(f"    {dev} stage(const real px, const real py, const real* __restrict__ C, "
                     "const real* __restrict__ Sg, const real* __restrict__ TT, "
                     "real* __restrict__ Gg) {")
        parts.append("        (void)px; (void)py; (void)C; (void)Sg; (void)TT; (void)Gg;")
        parts += [f"        Gg[{k}] = {pr.doprint(e)};" for k, (_, e) in enumerate(stage)]
        parts.append("    }")
        cols = sorted({state_names[s.name] for e in exprs + [t for _, t in temps]
                       for s in e.free_symbols if s.name in state_names})
        parts.append(f"    {dev} rhs(const real* __restrict__ Y, const int lane, const real px, "
                     "const real py, const real* __restrict__ C, const real* __restrict__ Sg, "
                     "const real* __restrict__ Gg, real* __restrict__ dy) {")
        parts.append("        (void)Y; (void)lane; (void)px; (void)py; (void)C; (void)Sg; (void)Gg;")
        if temps:
            parts += [f"        const real y{j} = Y[{j} * 64 + lane];" for j in cols]
            parts += [f"        const real {sym.name} = {pr.doprint(e)};" for sym, e in temps]
            parts += [f"        dy[{i - lo}] = {pr.doprint(e)};"
                      for i, e in zip(range(lo, hi), exprs)]
        else:
            # each LDS value is read just before the first row that uses it, so a value is
            # live from its first to its last use rather than through the whole evaluation
            loaded = set()
            for i, e in zip(range(lo, hi), exprs):
                first = sorted({state_names[s.name] for s in e.free_symbols
                                if s.name in state_names} - loaded)
                loaded.update(first)
                parts += [f"        const real y{j} = Y[{j} * 64 + lane];" for j in first]
                parts.append(f"        dy[{i - lo}] = {pr.doprint(e)};")
        parts.append("    }")
        if g == 0:
            obs_cols = sorted({state_names[s.name] for e in obs + [t for _, t in obs_temps]
                               for s in e.free_symbols if s.name in state_names})
            parts.append(f"    {dev} obs(const real* __restrict__ Y, const int lane, "
                         "const real px, const real py, const real* __restrict__ C, "
                         "const real* __restrict__ Sg, const real* __restrict__ Gg, "
                         "real& o_re, real& o_im) {")
            parts.append("        (void)Y; (void)lane; (void)px; (void)py; (void)C; (void)Sg; "
                         "(void)Gg;")
            parts += [f"        const real y{j} = Y[{j} * 64 + lane];" for j in obs_cols]
            parts += [f"        const real {sym.name} = {pr.doprint(e)};" for sym, e in obs_temps]
            parts.append(f"        o_re = {pr.doprint(obs[0])};")
            parts.append(f"        o_im = {pr.doprint(obs[1])};")
            parts.append("    }")
        parts.append(f"    {dev} init(real* __restrict__ y, const real px, const real py, "
                     "const real* __restrict__ C) {")
        parts.append("        (void)px; (void)py; (void)C;")
        if rho0_exprs is None:
            parts += [f"        y[{i - lo}] = R({1.0 if i == 0 else 0.0});" for i in range(lo, hi)]
        else:
            ipr = HipPrinter(names)
            parts += [f"        y[{i - lo}] = {ipr.doprint(sp.sympify(rho0_exprs[i]))};"
                      for i in range(lo, hi)]
        parts.append("    }")
        parts.append("};")
    head = [f"#define GQ_D {d}", f"#define GQ_G {groups}", f"#define GQ_RW {rw}",
            f"#define GQ_NATOMS {len(plan.time_atoms)}",
            f"#define GQ_NATOMS_ALLOC {max(1, len(plan.time_atoms))}"]
    cases = " ".join(f"case {g}: fn<{g}> args; break;" for g in range(len(blocks)))
    head.append(f"#define GQ_GROUP_DISPATCH(g, fn, args) switch (g) {{ {cases} default: break; }}")
    # GQ_NBUF LDS copies of the state (set by the caller): the one after buffer b
    head.append("#ifndef GQ_NBUF\n#define GQ_NBUF 2\n#endif")
    head.append("#define GQ_NEXT_BUF(b) (GQ_NBUF == 2 ? ((b) ^ 1) : 0)")
    # time atoms, shared by every wave
    anames = {"t": "t_now"}
    for i, c in enumerate(plan.const_syms):
        anames[c.name] = f"(double)C[{i}]"
    apr = HipPrinter(anames, literal="DBL")
    head.append("__device__ __forceinline__ void gq_atoms_f64(const double t_now, "
                "const real* __restrict__ C, double* __restrict__ TT) {")
    head.append("    (void)t_now; (void)C; (void)TT;")
    head += [f"    TT[{i}] = {apr.doprint(e)};" for i, (_, e) in enumerate(plan.time_atoms)]
    head.append("}")
    rnames = {"t": "t_now"}
    for i, c in enumerate(plan.const_syms):
        rnames[c.name] = f"C[{i}]"
    rpr = HipPrinter(rnames)
    head.append("__device__ __forceinline__ void gq_atoms_real(const real t_now, "
                "const real* __restrict__ C, real* __restrict__ TT) {")
    head.append("    (void)t_now; (void)C; (void)TT;")
    head += [f"    TT[{i}] = {rpr.doprint(e)};" for i, (_, e) in enumerate(plan.time_atoms)]
    head.append("}")
    head += _atoms_stage_lines(plan, {c.name: f"(double)C[{i}]" for i, c in
                                      enumerate(plan.const_syms)},
                               {c.name: f"C[{i}]" for i, c in enumerate(plan.const_syms)})
    return "\n".join(head + parts)


def kernel_suffix(fp64):
    """Kernel names carry the precision (gq_rk4_f64, gq_adaptive_f32, ...)."""
    return "f64" if fp64 else "f32"


def kernel_name(base, fp64):
    return f"{base}_{kernel_suffix(fp64)}"


def _sincos64_lines():
    """FP64 sin / cos without a library call, for the adaptive kernels' stage-time atoms:
    x = k pi/2 + r with |r| <= pi/4 by Cody and Waite's reduction (pi/2 in two FP64 parts; with
    FMA the first difference x - k * (pi/2)_hi is exact for |k| < 2^31), the Taylor polynomials
    of sin and cos in r to degree 17 / 16 by Horner's rule (the first omitted terms are below
    1e-19 at |r| = pi/4), and the quadrant k mod 4.  About 30 FP64 instructions for both."""
    from fractions import Fraction
    from math import factorial
    pio2 = sp.pi.evalf(40) / 2
    hi = float(pio2)
    lo = float(pio2 - sp.Float(hi, 40))
    inv = float(2 / sp.pi.evalf(40))

    def coef(n):
        return repr(float(Fraction((-1) ** (n // 2), factorial(n))))
    # sin r = r + r^3 (c3 + r^2 (c5 + ...)), cos r = 1 + r^2 (c2 + r^2 (c4 + ...))
    sin_terms = [coef(n) for n in range(17, 2, -2)]
    cos_terms = [coef(n) for n in range(16, 1, -2)]
    lines = ["__device__ __forceinline__ void gq_sincos64(double x, double* s, double* c) {",
             f"    const double k = rint(x * {inv!r});",
             f"    double r = fma(-k, {hi!r}, x);",
             f"    r = fma(-k, {lo!r}, r);",
             "    const double r2 = r * r;",
             f"    double ps = {sin_terms[0]};"]
    lines += [f"    ps = fma(ps, r2, {c});" for c in sin_terms[1:]]
    lines += [f"    double pc = {cos_terms[0]};"]
    lines += [f"    pc = fma(pc, r2, {c});" for c in cos_terms[1:]]
    lines += ["    const double sr = fma(r * r2, ps, r);",
              "    const double cr = fma(r2, pc, 1.0);",
              "    const int q = (int)k & 3;",
              "    const double sa = (q & 1) ? cr : sr;",
              "    const double ca = (q & 1) ? sr : cr;",
              "    *s = (q & 2) ? -sa : sa;",
              "    *c = ((q + 1) & 2) ? -ca : ca;",
              "}",
              "__device__ __forceinline__ double gq_sin64(double x) {",
              "    double s, c;", "    gq_sincos64(x, &s, &c);", "    return s;", "}",
              "__device__ __forceinline__ double gq_cos64(double x) {",
              "    double s, c;", "    gq_sincos64(x, &s, &c);", "    return c;", "}"]
    return lines


def _sincos32_lines():
    """FP32 sin / cos of an FP64 angle without a library call, for the FP32 adaptive kernels'
    stage-time atoms: the same Cody and Waite reduction as ``gq_sincos64`` in FP64 to |r| <=
    pi/4, then r rounded to FP32 and the Taylor polynomials of sin and cos to degree 11 / 12 in
    FP32 (the first omitted terms are below 2e-10 at |r| = pi/4), and the quadrant.  The FP32
    library function would carry its large-argument reduction, which the reduced angle never
    needs, and with it about 90 registers."""
    from fractions import Fraction
    from math import factorial
    pio2 = sp.pi.evalf(40) / 2
    hi = float(pio2)
    lo = float(pio2 - sp.Float(hi, 40))
    inv = float(2 / sp.pi.evalf(40))

    def coef(n):
        return repr(float(Fraction((-1) ** (n // 2), factorial(n)))) + "f"
    sin_terms = [coef(n) for n in range(11, 2, -2)]
    cos_terms = [coef(n) for n in range(12, 1, -2)]
    lines = ["__device__ __forceinline__ void gq_sincos32(double x, float* s, float* c) {",
             f"    const double k = rint(x * {inv!r});",
             f"    double rd = fma(-k, {hi!r}, x);",
             f"    rd = fma(-k, {lo!r}, rd);",
             "    const float r = (float)rd;",
             "    const float r2 = r * r;",
             f"    float ps = {sin_terms[0]};"]
    lines += [f"    ps = fmaf(ps, r2, {c});" for c in sin_terms[1:]]
    lines += [f"    float pc = {cos_terms[0]};"]
    lines += [f"    pc = fmaf(pc, r2, {c});" for c in cos_terms[1:]]
    lines += ["    const float sr = fmaf(r * r2, ps, r);",
              "    const float cr = fmaf(r2, pc, 1.0f);",
              "    const int q = (int)k & 3;",
              "    const float sa = (q & 1) ? cr : sr;",
              "    const float ca = (q & 1) ? sr : cr;",
              "    *s = (q & 2) ? -sa : sa;",
              "    *c = ((q + 1) & 2) ? -ca : ca;",
              "}",
              "__device__ __forceinline__ float gq_sin32(double x) {",
              "    float s, c;", "    gq_sincos32(x, &s, &c);", "    return s;", "}",
              "__device__ __forceinline__ float gq_cos32(double x) {",
              "    float s, c;", "    gq_sincos32(x, &s, &c);", "    return c;", "}"]
    return lines


def prelude(fp64, acc64, block_size, defines):
    real = "double" if fp64 else "float"
    acc = "double" if (fp64 or acc64) else "float"
    out = [f"typedef {real} real;",
           f"typedef {acc} acc_t;",
           f"typedef {real}2 real2;",
           "#define R(x) ((real)(x))",
           "#define DBL(x) ((double)(x))",
           f"#define GQ_FP64 {1 if fp64 else 0}",
           "#define GQ_CAT_(a, b) a##_##b",
           "#define GQ_CAT(a, b) GQ_CAT_(a, b)",
           f"#define GQ_KSUFFIX {kernel_suffix(fp64)}",
           "#define GQ_KERNEL(name) GQ_CAT(name, GQ_KSUFFIX)",
           f"#define GQ_BLOCK {int(block_size)}",
           "__device__ __forceinline__ real gq_sign(real x) { return (x > R(0)) - (x < R(0)); }",
           "// x - 2 pi rint(x / 2 pi) in FP64, 2 pi split in two parts (Cody and Waite)",
           "__device__ __forceinline__ double gq_reduce_2pi(double x) {",
           "    const double k = rint(x * 0.15915494309189535);",
           "    return fma(-k, 2.4492935982947064e-16, fma(-k, 6.283185307179586, x));",
           "}"] + _sincos64_lines() + _sincos32_lines() + [
           "// 1 / x for the model code's negative powers: an IEEE division in FP64; in FP32 the",
           "// hardware reciprocal (within 1 ulp of 1 / x, about a fifth of the instructions of an",
           "// IEEE FP32 division; the FP32 right-hand side carries 6e-8 relative rounding anyway)",
           "__device__ __forceinline__ double gq_rcp(double x) { return 1.0 / x; }",
           "__device__ __forceinline__ float gq_rcp(float x) {",
           "#if defined(__HIP_DEVICE_COMPILE__)",
           "    return __builtin_amdgcn_rcpf(x);",
           "#else",
           "    return 1.0f / x;   // host builds of the generated code (CPU checks)",
           "#endif",
           "}"]
    for k, v in defines.items():
        out.append(f"#define {k} {v}")
    out += ["#ifndef GQ_HILBERT", "#define GQ_HILBERT 0", "#endif"]
    return "\n".join(out)


def assemble(plan, templates, *, fp64, acc64=True, block_size=256, defines=None,
             rho0_exprs=None, extra="", groups=1):
    """Full hiprtc source: ABI header + prelude + model + solver constants + templates
    (``groups`` > 1: the grouped model functions, ``emit_grouped``)."""
    with open(_header_path()) as fh:
        header = fh.read()
    model = (emit_model(plan, rho0_exprs) if groups == 1
             else emit_grouped(plan, groups, rho0_exprs))
    defines = dict(defines or {})
    defines.setdefault("GQ_GROUPED", int(groups > 1))
    parts = [header, prelude(fp64, acc64, block_size, defines), model, extra]
    parts += [read_kernel_template(t) for t in (["common.hip"] + list(templates))]
    return "\n".join(parts)


HIPRTC_BASE_OPTIONS = ("-O3", "-std=c++17", "-ffp-contract=fast",
                       # packed FP32 instructions are off for every kernel of this library
                       "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops")


def strip_comments(src):
    return re.sub(r"//[^\n]*", "", src)
