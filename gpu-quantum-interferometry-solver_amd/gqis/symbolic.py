"""Symbolic front end: reduced Lindblad equations and the hoisting plan of the RHS.

Written from the mathematics and the reference's documented behaviour; no code is shared.

* Reduced layout (reference ``gqis/solver.py:202-253``, ``GQIS_API.md:110-118,544-549``): the
  first ``N-1`` populations, then ``Re``/``Im`` of the upper-triangle coherences in row-major
  order; the last population is ``1 - sum(populations)``.
* Reduced Lindblad right-hand side (reference ``gqis/solver.py:399-493``): the real equations of
  ``-i[H, rho] + sum_L (L rho L^+ - 1/2 {L^+ L, rho})`` in that layout, each simplified per state
  coefficient.
* ``make_plan`` replaces the reference's CSE/hoisting plan (``gqis/solver.py:550-625,749-886``):
  every state-independent sub-expression is evaluated once per parameter point ("static") or once
  per stage time ("stage"); every time-only sub-expression becomes a "time atom" that the
  fixed-step kernels read from a table computed in FP64 once per launch.
"""
from dataclasses import dataclass, field
import numbers

import numpy as np
import sympy as sp

PARX = sp.Symbol("ParX", real=True)
PARY = sp.Symbol("ParY", real=True)
TIME = sp.Symbol("t", real=True)


def state_symbols(size):
    return [sp.Symbol(f"rho[{i}]", real=True) for i in range(size)]


def build_independent_rho(N):
    """Hermitian unit-trace ``N x N`` matrix of ``N*N-1`` real symbols, and layout metadata.

    Returns ``(rho, meta)`` with meta keys ``N, rho_syms, num_diag, num_coherences, vec_len``
    (reference ``gqis/solver.py:202-253``).
    """
    n = int(N)
    num_diag = n - 1
    num_coh = n * (n - 1) // 2
    syms = state_symbols(num_diag + 2 * num_coh)
    pos = {}
    k = num_diag
    for i in range(n):
        for j in range(i + 1, n):
            pos[(i, j)] = k
            k += 2
    last = 1 - sp.Add(*syms[:num_diag])

    def entry(i, j):
        if i == j:
            return syms[i] if i < num_diag else last
        a, b = (i, j) if i < j else (j, i)
        re, im = syms[pos[(a, b)]], syms[pos[(a, b)] + 1]
        return re + sp.I * im if i < j else re - sp.I * im

    rho = sp.Matrix(n, n, entry)
    meta = {"N": n, "rho_syms": syms, "num_diag": num_diag, "num_coherences": num_coh,
            "vec_len": len(syms)}
    return rho, meta


def rho_matrix_to_independent_exprs(rho0):
    """Reduced real expressions of a density matrix, or a checked reduced list
    (reference ``gqis/solver.py:256-290``)."""
    if isinstance(rho0, sp.MatrixBase):
        rows, cols = rho0.shape
        if rows != cols:
            raise ValueError("rho0 matrix must be square.")
        n = rows
        out = [sp.simplify(sp.re(rho0[i, i])) for i in range(n - 1)]
        for i in range(n):
            for j in range(i + 1, n):
                out.append(sp.simplify(sp.re(rho0[i, j])))
                out.append(sp.simplify(sp.im(rho0[i, j])))
        return out
    exprs = [sp.sympify(e) for e in rho0]
    n = int(round(np.sqrt(len(exprs) + 1)))
    if n * n - 1 != len(exprs):
        raise ValueError("rho0 reduced vector length must match N*N-1.")
    return [sp.simplify(e) for e in exprs]


def simplify_affine(expr, states):
    """Collect ``expr`` by the state symbols and simplify each coefficient on its own.

    Simplifying the whole equation would put drive terms over a common denominator; per
    coefficient keeps one multiply-add per state (reference ``gqis/solver.py:325-337``).
    """
    expanded = sp.expand(expr)
    groups = sp.collect(expanded, states, evaluate=False)
    if any(key != 1 and key not in states for key in groups):
        return sp.simplify(expr)
    terms = []
    for key, coeff in groups.items():
        coeff = coeff if coeff.is_number else sp.simplify(coeff)
        terms.append(coeff * key)
    return sp.Add(*terms)


def _lindblad_entry(i, j, h, rho, ops):
    """(i, j) entry of the Lindblad right-hand side, written element by element."""
    n = h.shape[0]
    val = sp.Integer(0)
    for k in range(n):
        if h[i, k] != 0:
            val += -sp.I * h[i, k] * rho[k, j]
        if h[k, j] != 0:
            val += sp.I * rho[i, k] * h[k, j]
    for op, opd, gram in ops:
        for k in range(n):
            if op[i, k] == 0:
                continue
            for m in range(n):
                if opd[m, j] != 0:
                    val += op[i, k] * rho[k, m] * opd[m, j]
        for k in range(n):
            if gram[i, k] != 0:
                val -= sp.Rational(1, 2) * gram[i, k] * rho[k, j]
            if gram[k, j] != 0:
                val -= sp.Rational(1, 2) * rho[i, k] * gram[k, j]
    return val


def build_reduced_lindblad_rhs(N, H, Col_Ops, mean_operator, *, pre_expand=False,
                               collect_rho=False, factor_terms=False, symbolic_workers=1,
                               max_symbolic_workers=None):
    """Real reduced Lindblad equations and the observable (reference ``gqis/solver.py:399-493``).

    Returns ``(drho_eqs, mean_re, mean_im, meta)``: ``N*N-1`` expressions in the reduced layout,
    the real and imaginary parts of ``Tr(mean_operator rho)``, and the layout metadata.
    ``symbolic_workers`` is accepted for signature parity; the element-wise build used here is
    fast enough that it runs in this process.
    """
    n = int(N)
    rho, meta = build_independent_rho(n)
    syms = meta["rho_syms"]
    h = sp.Matrix(H)
    ops = []
    for op in Col_Ops:
        op = sp.Matrix(op)
        opd = op.H
        ops.append((op, opd, opd * op))
    eqs = []
    order = [(i, i) for i in range(n - 1)]
    order += [(i, j) for i in range(n) for j in range(i + 1, n)]
    for i, j in order:
        val = _lindblad_entry(i, j, h, rho, ops)
        if pre_expand:
            val = sp.expand(val)
        parts = [sp.re(val)] if i == j else [sp.re(val), sp.im(val)]
        for part in parts:
            if factor_terms:
                part = sp.factor_terms(part)
            part = simplify_affine(part, syms)
            if collect_rho:
                part = sp.collect(part, syms)
            eqs.append(part)
    mean_val = sp.expand(sum(mean_operator[a, b] * rho[b, a]
                             for a in range(n) for b in range(n) if mean_operator[a, b] != 0))
    mean_re = simplify_affine(sp.re(mean_val), syms)
    mean_im = simplify_affine(sp.im(mean_val), syms)
    return eqs, mean_re, mean_im, meta


# ----------------------------------------------------------------------------------------------
# Hoisting plan
# ----------------------------------------------------------------------------------------------

@dataclass
class RHSPlan:
    """Evaluation plan of an ODE right-hand side, independent of the device language.

    Symbols used by the expressions: the states ``rho[i]``, ``ParX``/``ParY``, the runtime
    constants (``const_syms``, read from ``Const_arr``), the static values ``S_k``, the time
    atoms ``TT_k`` and the stage values ``G_k``.
    """
    state_size: int
    states: tuple
    const_syms: tuple
    statics: list           # [(S_k, expr in ParX, ParY, consts, earlier S)]
    time_atoms: list        # [(TT_k, expr in t and consts)]
    stage: list             # [(G_k, expr in ParX, ParY, consts, S, TT, earlier G)]
    rhs_temps: list         # [(symbol, expr)] state-dependent CSE temporaries
    rhs_out: list           # D expressions
    obs_temps: list
    obs_out: list           # [re, im]
    affine: bool            # every equation is affine in the states
    nnz: int = 0            # state coefficients that are not zero (affine case)
    extra: dict = field(default_factory=dict)


class _Hoister:
    """Replaces state-free sub-expressions by static or stage symbols (deduplicated)."""

    def __init__(self, dynamic, force=()):
        self.dynamic = set(dynamic)     # symbols that make an expression non-hoistable
        self.force = set(force)         # bare symbols that still need a hoisted value (time)
        self.values = {}                # expr -> symbol
        self.order = []                 # [(symbol, expr)]
        self.count = 0

    def _new(self, expr, prefix):
        sym = self.values.get(expr)
        if sym is None:
            sym = sp.Symbol(f"{prefix}{self.count}", real=True)
            self.count += 1
            self.values[expr] = sym
            self.order.append((sym, expr))
        return sym

    def __call__(self, expr, prefix):
        if not (expr.free_symbols & self.dynamic):
            if expr.is_Number or not expr.free_symbols:
                return expr
            if expr.is_Symbol and expr not in self.force:
                return expr
            return self._new(expr, prefix)
        if expr.is_Add or expr.is_Mul:
            free = [a for a in expr.args if not (a.free_symbols & self.dynamic)]
            dyn = [a for a in expr.args if a.free_symbols & self.dynamic]
            new_dyn = [self(a, prefix) for a in dyn]
            if free:
                lump = expr.func(*free)
                if not (lump.is_Number or not lump.free_symbols
                        or (lump.is_Symbol and lump not in self.force)):
                    lump = self._new(lump, prefix)
                return expr.func(lump, *new_dyn)
            return expr.func(*new_dyn)
        if expr.args:
            return expr.func(*[self(a, prefix) for a in expr.args])
        return expr


def _affine_coefficients(expr, states):
    """{state or 1: coefficient} if expr is affine in the states, else None."""
    try:
        poly = sp.Poly(expr, *states)
    except sp.PolynomialError:
        return None
    if poly.total_degree() > 1:
        return None
    out = {}
    for monom, coeff in poly.terms():
        if any(e not in (0, 1) for e in monom):
            return None
        key = sp.Integer(1)
        for e, s in zip(monom, states):
            if e:
                key = s
        if coeff.free_symbols & set(states):
            return None
        out[key] = coeff
    return out


def _split_time_atoms(expr, time_syms, const_syms, atoms):
    """Replace maximal sub-expressions of time and constants (containing time) by atoms."""
    allowed = set(time_syms) | set(const_syms)

    def walk(e):
        fs = e.free_symbols
        if fs and fs <= allowed and fs & set(time_syms):
            if e.is_Number:
                return e
            sym = atoms.get(e)
            if sym is None:
                sym = sp.Symbol(f"TT{len(atoms)}", real=True)
                atoms[e] = sym
            return sym
        if e.is_Add or e.is_Mul:
            # keep time-only factors/summands of an Add/Mul together as one atom
            timed = [a for a in e.args if a.free_symbols and a.free_symbols <= allowed
                     and a.free_symbols & set(time_syms)]
            rest = [a for a in e.args if a not in timed]
            if len(timed) > 1:
                return e.func(walk(e.func(*timed)), *[walk(a) for a in rest])
        if e.args:
            return e.func(*[walk(a) for a in e.args])
        return e
    return walk(expr)


def make_plan(equations, obs_re, obs_im, states, drive_map, const_syms, *,
              cse_simplify=False, time_symbol=TIME):
    """Build the static / time-atom / stage / RHS plan of an ODE right-hand side.

    ``equations``/``obs_*`` may contain drive placeholder symbols (keys of ``drive_map``), whose
    expressions depend on ``time_symbol``, ``ParX``, ``ParY`` and the constants.
    """
    states = list(states)
    stset = set(states)
    drive_map = dict(drive_map or {})
    eqs = [sp.sympify(e).xreplace(drive_map) for e in equations]
    obs = [sp.sympify(obs_re).xreplace(drive_map), sp.sympify(obs_im).xreplace(drive_map)]

    # 1. Affine decomposition where possible: one coefficient per (equation, state).
    affine = True
    nnz = 0
    decomposed = []
    for e in eqs:
        coeffs = _affine_coefficients(e, states)
        if coeffs is None:
            affine = False
            decomposed.append(e)
            continue
        nnz += sum(1 for k in coeffs if k != 1)
        terms = [c * (k if k != 1 else 1) for k, c in coeffs.items()]
        decomposed.append(sp.Add(*terms) if terms else sp.Integer(0))
    obs_dec = []
    for e in obs:
        coeffs = _affine_coefficients(e, states)
        obs_dec.append(e if coeffs is None else sp.Add(*[c * (k if k != 1 else 1)
                                                          for k, c in coeffs.items()]))

    # 2. Hoist state-free parts: they become "H" values (static or stage, decided below).
    hoist = _Hoister(stset, force={time_symbol})
    rhs = [hoist(e, "H") for e in decomposed]
    obs_h = [hoist(e, "H") for e in obs_dec]

    # 3. Classify hoisted values; tabulate time atoms of the time-dependent ones.
    tsyms = {time_symbol}
    atoms = {}
    static_vals, stage_vals = [], []
    for sym, expr in hoist.order:
        if expr.free_symbols & tsyms:
            stage_vals.append((sym, _split_time_atoms(expr, tsyms, const_syms, atoms)))
        else:
            static_vals.append((sym, expr))
    atom_syms = set(atoms.values())

    # 4. Inside stage values, hoist the atom-free parts to statics (one FMA chain per stage).
    inner = _Hoister(atom_syms)
    stage_final = [(sym, inner(sp.expand(expr) if _cheap_expand(expr) else expr, "P"))
                   for sym, expr in stage_vals]
    static_vals = static_vals + inner.order

    # 5. CSE every level.
    def cse(pairs, prefix, ignore=()):
        if not pairs:
            return [], []
        syms, exprs = zip(*pairs)
        temps, reduced = sp.cse(list(exprs), symbols=sp.numbered_symbols(prefix),
                                ignore=tuple(ignore))
        if cse_simplify:
            reduced = [sp.simplify(r) for r in reduced]
        return temps, list(zip(syms, reduced))

    s_temps, statics = cse(static_vals, "sx")
    g_temps, stage = cse(stage_final, "gx")
    out_temps, rhs_red = sp.cse(rhs + obs_h, symbols=sp.numbered_symbols("tx"),
                                ignore=tuple(states))
    if cse_simplify:
        rhs_red = [sp.simplify(r) for r in rhs_red]
    # Split the CSE temporaries: obs-only temporaries are recomputed in the observable.
    plan = RHSPlan(
        state_size=len(states), states=tuple(states), const_syms=tuple(const_syms),
        statics=list(s_temps) + statics,
        time_atoms=[(sym, expr) for expr, sym in atoms.items()],
        stage=list(g_temps) + stage,
        rhs_temps=list(out_temps), rhs_out=rhs_red[:len(rhs)],
        obs_temps=list(out_temps), obs_out=rhs_red[len(rhs):],
        affine=affine, nnz=nnz)
    _prune_temps(plan)
    return plan


def _cheap_expand(expr):
    return sp.count_ops(expr) < 200


def _prune_temps(plan):
    """Keep only the CSE temporaries each block needs (RHS and observable separately)."""
    def needed(temps, outs):
        need = set().union(*(sp.sympify(o).free_symbols for o in outs)) if outs else set()
        keep = []
        for sym, expr in reversed(temps):
            if sym in need:
                keep.append((sym, expr))
                need |= expr.free_symbols
        return list(reversed(keep))
    plan.rhs_temps = needed(plan.rhs_temps, plan.rhs_out)
    plan.obs_temps = needed(plan.obs_temps, plan.obs_out)


def count_flops(plan):
    """Operation count of one RHS evaluation (adds + multiplies; functions counted once)."""
    total = 0
    for _, e in plan.rhs_temps:
        total += sp.count_ops(e)
    for e in plan.rhs_out:
        total += sp.count_ops(e)
    return int(total)
