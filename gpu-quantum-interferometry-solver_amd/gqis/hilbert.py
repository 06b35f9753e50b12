"""Hilbert-space (MFMA) path for larger Lindblad models: RK4 on the full density matrix with
the commutator as one complex GEMM per stage on the matrix cores (``kernels/hilbert_rk4.hip``),
or, for the split-precision forms at N = 16 and 32, as both of its products
(``kernels/hilbert_split.hip``).

Applies when H = H0 + diag(v) with H0 constant and v(ParX, ParY, t) a linear combination of
a few basis functions with numeric row coefficients, every collapse operator is constant and
diagonal or single-entry, and the observable is a constant diagonal.  The operators are
numeric here, so no reduced symbolic equations are built: an N = 32 model (1023 reduced
equations) is prepared in milliseconds.
"""
from dataclasses import dataclass
import time

import numpy as np
import sympy as sp

from . import codegen, runtime
from .codegen import HipPrinter
from .symbolic import PARX, PARY, TIME, _split_time_atoms

SUPPORTED_N = (8, 16, 32)


@dataclass
class HilbertModel:
    n: int
    h_eff: np.ndarray        # complex N x N: H0 - (i/2) sum L^+ L
    dep: np.ndarray          # real N x N: sum_k d_k(i) d_k(j) over diagonal collapse operators
    jump: np.ndarray         # real N x N: |l|^2 at [a][b] for single-entry l|a><b|, a != b
    obs: np.ndarray          # real N: diagonal observable
    vcoef: np.ndarray        # real K x N: v_i = sum_k vcoef[k, i] g_k
    basis: list              # K SymPy expressions g_k(ParX, ParY, t, consts)
    runtime_syms: tuple
    key: tuple
    # collapse operators with at most one nonzero per row that are neither diagonal nor a single
    # entry (a qubit's lowering operator in a multi-qubit space): L[i][mcol[k, i]] = ml[k, i],
    # mcol -1 for a zero row; their jump term (L rho L^+)_ij = ml_i conj(ml_j) rho[mcol_i][mcol_j]
    mcol: np.ndarray = None  # int M x N
    ml: np.ndarray = None    # complex M x N


def _numeric(e):
    e = sp.sympify(e)
    if e.free_symbols:
        return None
    return complex(sp.N(e))


def extract(model, drive_map):
    """HilbertModel from a prepared model (``solver._prepare_model`` output), or None."""
    h = model["H"].xreplace(drive_map) if drive_map else model["H"]
    n = h.shape[0]
    if n not in SUPPORTED_N or model["rho0_exprs"] is not None:
        return None
    h0 = np.zeros((n, n), complex)
    dyn = []
    for i in range(n):
        for k in range(n):
            e = sp.expand(h[i, k])
            if i != k:
                v = _numeric(e)
                if v is None:
                    return None
                h0[i, k] = v
            else:
                const, rest = e.as_coeff_Add()
                v = _numeric(const)
                if v is None or abs(v.imag) > 0:
                    return None
                h0[i, i] = v.real
                dyn.append(rest)
    if not np.allclose(h0, h0.conj().T, atol=1e-12):
        return None
    basis, rows = [], []
    for i, e in enumerate(dyn):
        terms = {} if e == 0 else sp.Add.make_args(e)
        coefs = {}
        for t in terms:
            c, m = t.as_coeff_Mul()
            if m.free_symbols - {PARX, PARY, TIME} - set(model["runtime_const_syms"]):
                return None
            if not c.is_number or sp.im(c) != 0:
                return None
            coefs[m] = coefs.get(m, 0.0) + float(c)
        rows.append(coefs)
        for m in coefs:
            if m not in basis:
                basis.append(m)
    if not basis:
        basis = [sp.Integer(0)]
    vcoef = np.array([[r.get(m, 0.0) for r in rows] for m in basis], np.float64)
    dep = np.zeros((n, n))
    jump = np.zeros((n, n))
    mono = []
    h_eff = h0.astype(complex)
    for op in model["Col_Ops"]:
        mat = np.zeros((n, n), complex)
        for i in range(n):
            for k in range(n):
                v = _numeric(op[i, k])
                if v is None:
                    return None
                mat[i, k] = v
        nz = np.argwhere(np.abs(mat) > 0)
        h_eff = h_eff - 0.5j * (mat.conj().T @ mat)
        if len(nz) == 0:
            continue
        if np.all(nz[:, 0] == nz[:, 1]):
            d = np.diag(mat)
            if np.any(np.abs(d.imag) > 0):
                return None
            dep += np.outer(d.real, d.real)
        elif len(nz) == 1:
            a, b = nz[0]
            jump[a, b] += abs(mat[a, b]) ** 2
        elif np.all(np.count_nonzero(np.abs(mat) > 0, axis=1) <= 1):
            col = np.full(n, -1)
            coef = np.zeros(n, complex)
            for i, k in nz:
                col[i], coef[i] = k, mat[i, k]
            mono.append((col, coef))
        else:
            return None
    o = np.zeros(n)
    for i in range(n):
        for k in range(n):
            v = _numeric(model["mean_operator"][i, k])
            if v is None or (i != k and v != 0) or (i == k and v.imag != 0):
                return None
            if i == k:
                o[i] = v.real
    mcol = np.array([c for c, _ in mono], int).reshape(len(mono), n)
    ml = np.array([l for _, l in mono], complex).reshape(len(mono), n)
    key = ("hilbert", n, h_eff.tobytes(), dep.tobytes(), jump.tobytes(), o.tobytes(),
           vcoef.tobytes(), tuple(sp.srepr(b) for b in basis), mcol.tobytes(), ml.tobytes())
    return HilbertModel(n, h_eff, dep, jump, o, vcoef, basis,
                        tuple(model["runtime_const_syms"]), key, mcol, ml)


def pack_a(m):
    """MFMA A operand in the kernel's K order: [tile][step][lane] (see hilbert_rk4.hip)."""
    n = m.n
    big = np.zeros((2 * n, 2 * n))
    big[:n, :n] = m.h_eff.real
    big[:n, n:] = -m.h_eff.imag
    big[n:, :n] = m.h_eff.imag
    big[n:, n:] = m.h_eff.real
    tiles = (2 * n + 31) // 32
    out = np.zeros((tiles, n, 64))
    for t in range(tiles):
        for s in range(n):
            for lane in range(64):
                c, h = lane & 31, lane >> 5
                kappa = 8 * (s // 4) + 4 * h + (s % 4)
                row = 32 * t + c
                if row < 2 * n:
                    out[t, s, lane] = big[row, kappa]
    return out


def dense_enough(m):
    """Whether the dense matrix-core product beats the per-point kernels (below 32 levels):
    their work grows with the nonzeros of H_eff (a reduced element couples to the elements its
    row and column of H_eff reach), the product's does not.  Measured (2000 RK4 steps): dense
    random H_eff at N = 8 in FP64, 52 ms on the matrix cores against 76 ms with 8 lanes per
    point (256^2); the tridiagonal 8-level chain about 3x faster on the lanes; the 4-qubit chain
    (5 nonzeros per row of 16) 17.9 ms on 8 lanes per point against 103 ms with FP16 hi/lo
    products (128^2).  Break-even is about 5.5 nonzeros per row at N = 8."""
    per_row = np.count_nonzero(np.abs(m.h_eff) > 0) / m.n
    return per_row >= 0.75 * m.n


def dmfma_order(n):
    """Logical Rbig row (Re rho_i: i, Im rho_i: n + i) held at each row 16 T + 4 r + g of the
    FP64 matrix-core kernel's tiles: slot m = 4 T + r of lane group g is Re rho_{g n/4 + m} for
    m < n/4, else Im rho_{g n/4 + m - n/4} (kernels/hilbert_rk4_dmfma.hip)."""
    q = n // 4
    perm = np.zeros(2 * n, int)
    for t in range(n // 8):
        for r in range(4):
            for g in range(4):
                m = 4 * t + r
                perm[16 * t + 4 * r + g] = g * q + m if m < q else n + g * q + (m - q)
    return perm


def pack_a_dmfma(m):
    """A operand of the FP64 matrix-core kernel: [To][T][r][lane], lane (i, g) = (lane % 16,
    lane // 16) holding Hbig[16 To + i][16 T + 4 r + g] with rows and columns in dmfma_order."""
    n = m.n
    big = np.zeros((2 * n, 2 * n))
    big[:n, :n], big[:n, n:] = m.h_eff.real, -m.h_eff.imag
    big[n:, :n], big[n:, n:] = m.h_eff.imag, m.h_eff.real
    perm = dmfma_order(n)
    bp = big[np.ix_(perm, perm)]
    tr = n // 8
    lane = np.arange(64)
    out = np.zeros((tr, tr, 4, 64))
    for to in range(tr):
        for t in range(tr):
            for r in range(4):
                out[to, t, r] = bp[16 * to + lane % 16, 16 * t + 4 * r + lane // 16]
    return out


SPLIT_KINDS = {None: 0, "f16": 1, "bf16": 2}
# the sparse form uses compile-time XOR masks when the model has the qubit-register structure
# (``xor_form``); False forces its table form (kept for tests of that form)
SPARSE_XOR = True


def _round_half(x, kind):
    """Round FP64 values to FP16 or BF16 (round to nearest even), returned as FP64."""
    if kind == "f16":
        return x.astype(np.float16).astype(np.float64)
    f = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    rounded = (f + 0x7FFF + ((f >> 16) & 1)) & 0xFFFF0000
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def kernel_base(split=None, fp64=False, valu=False, sparse=False):
    """Kernel name (before the precision suffix): FP32 MFMA / split-precision forms, the FP64
    matrix-core form (gq_hilbert_rk4_dmfma_f64), with ``valu`` the FP64 VALU form, with
    ``sparse`` the wavefront-per-point form for a sparse H_eff."""
    if sparse:
        return "gq_hilbert_sparse"
    if fp64 and not valu:
        return "gq_hilbert_rk4_dmfma"
    return {None: "gq_hilbert_rk4", "f16": "gq_hilbert_rk4_f16x3",
            "bf16": "gq_hilbert_rk4_bf16x3"}[split]


def pack_a_split(m, kind):
    """A operand of the split-precision kernel, hi and lo parts: [tile][K block][lane][8],
    element j of lane (r, h) in block b holding Hbig[32 t + r][kappa(8 b + j, h)] with
    kappa(s, h) = 8 (s // 4) + 4 h + s % 4, the Rbig row of the state value x[s] that lanes of
    half h hold (the B operand is the state itself, no reordering)."""
    n = m.n
    big = np.zeros((2 * n, 2 * n))
    big[:n, :n], big[:n, n:] = m.h_eff.real, -m.h_eff.imag
    big[n:, :n], big[n:, n:] = m.h_eff.imag, m.h_eff.real
    tiles, blocks = (2 * n + 31) // 32, n // 8
    full = np.zeros((tiles, blocks, 64, 8))
    for t in range(tiles):
        for b in range(blocks):
            for lane in range(64):
                r, h = lane & 31, lane >> 5
                if 32 * t + r >= 2 * n:
                    continue
                for j in range(8):
                    s_idx = 8 * b + j
                    kappa = 8 * (s_idx // 4) + 4 * h + (s_idx % 4)
                    full[t, b, lane, j] = big[32 * t + r, kappa]
    hi = _round_half(full, kind)
    lo = _round_half(full - hi, kind)
    return hi, lo


# The two-sided split-precision kernel (kernels/hilbert_split.hip, N = 16 and 32) runs both
# products of the commutator on the matrix cores, so it needs no transposed copy of the state.
# Its derivative part from H_eff is, per element (i, j) and real part P (0 = Re, 1 = Im),
#     -i (H_eff rho)_ij + i (rho H_eff^+)_ij
#   = sum_{Q,k} left[P][i][Q][k] rho_Q[k][j] + sum_{Q,k} rho_Q[k][i] right[P][Q][k][j],
# the second sum using rho_ik = conj(rho_ki) (rho is Hermitian): the state values a lane holds
# for column i of rho are, unchanged, the A operand of row i of the second product.  Since
# right[P][Q][k][j] = (+1, -1)[P] left[P][j][Q][k], the right coefficients in the B layout are
# the left ones in the A layout (the same registers), negated for P = 1: the kernel loads the
# left operands only and subtracts the P = 1 right products.
def two_sided_coefficients(m):
    """Real coefficient arrays ``left`` [P][i][Q][k] and ``right`` [P][Q][k][j] of the
    two-sided form (see above)."""
    h = m.h_eff
    n = m.n
    left = np.zeros((2, n, 2, n))
    left[0, :, 0, :], left[0, :, 1, :] = h.imag, h.real      # Re(-i H rho) = Im(H rho)
    left[1, :, 0, :], left[1, :, 1, :] = -h.real, h.imag     # Im(-i H rho) = -Re(H rho)
    right = np.zeros((2, 2, n, n))
    ht = h.T                                                 # [k][j] = H_jk
    right[0, 0], right[0, 1] = ht.imag, ht.real
    right[1, 0], right[1, 1] = ht.real, -ht.imag
    return left, right


def pack_two_sided(m, kind):
    """Operands of ``kernels/hilbert_split.hip`` (v_mfma_f32_16x16x32: A lane l holds
    A[l % 16][8 (l // 16) + e], B lane l holds B[8 (l // 16) + e][l % 16], e < 8), as hi and lo
    parts.  With NB = N / 16 blocks of 16 rows, K block kb covers the Rbig rows 16 kb .. 16 kb
    + 15 of both real parts: K slot 8 g + e is part Q = e // 4 of row k = 16 kb + 4 g + e % 4.
    Returns (left_hi, left_lo, right_hi, right_lo):
      left  [P][rb][kb][lane][8]: A operand, row i = 16 rb + lane % 16 of part P;
      right [P][kb][cb][lane][8]: B operand, column j = 16 cb + lane % 16 of part P."""
    n = m.n
    nb = n // 16
    left, right = two_sided_coefficients(m)
    lp = np.zeros((2, nb, nb, 64, 8))
    rp = np.zeros((2, nb, nb, 64, 8))
    for p in range(2):
        for b1 in range(nb):
            for kb in range(nb):
                for lane in range(64):
                    c, g = lane % 16, lane // 16
                    for e in range(8):
                        q, k = e // 4, 16 * kb + 4 * g + e % 4
                        lp[p, b1, kb, lane, e] = left[p, 16 * b1 + c, q, k]
                        rp[p, kb, b1, lane, e] = right[p, q, k, 16 * b1 + c]
    lhi = _round_half(lp, kind)
    rhi = _round_half(rp, kind)
    return lhi, _round_half(lp - lhi, kind), rhi, _round_half(rp - rhi, kind)


def two_sided_applies(n, split):
    """The split-precision forms at N = 16 and 32 use the two-sided kernel."""
    return bool(split) and n in (16, 32)


def ell_table(m):
    """Off-diagonal nonzeros of H_eff as an ELL table: K = the largest count in a row (at
    least 1); row i's entries (column, value), padded with (i, 0)."""
    n = m.n
    off = [[k for k in range(n) if k != i and m.h_eff[i, k] != 0] for i in range(n)]
    width = max(1, max(len(r) for r in off))
    cols = np.tile(np.arange(n)[:, None], (1, width))
    vals = np.zeros((n, width), complex)
    for i, r in enumerate(off):
        cols[i, :len(r)] = r
        vals[i, :len(r)] = m.h_eff[i, r]
    return width, cols, vals


def xor_form(m):
    """The qubit-register structure of H_eff and the jump operators, or None: every row i has
    its off-diagonal nonzeros at columns i ^ mask_k with values v_k that do not depend on i
    (sums of sx / sy on single qubits), and every general jump operator maps row i to i ^ b for
    the rows with i & b == cond, all with one value l (a qubit's lowering or raising operator).
    Returns (masks, values, [(b, cond, |l|^2)])."""
    n = m.n
    _, cols, vals = ell_table(m)
    ref = None
    for i in range(n):
        row = {int(c) ^ i: complex(v) for c, v in zip(cols[i], vals[i]) if v != 0}
        if ref is None:
            ref = row
        elif row.keys() != ref.keys() or not all(
                np.isclose(row[k], ref[k], rtol=1e-14, atol=0.0) for k in row):
            return None
    masks = sorted(ref) or [0]
    values = [ref.get(k, 0.0) for k in masks]
    mono = []
    for kk in range(n_mono(m)):
        col, lv = m.mcol[kk], m.ml[kk]
        rows = [i for i in range(n) if col[i] >= 0]
        flips = {int(col[i]) ^ i for i in rows}
        if len(flips) != 1:
            return None
        b = flips.pop()
        conds = {i & b for i in rows}
        if len(conds) != 1:
            return None
        cond = conds.pop()
        if any(((i & b) == cond) != (col[i] >= 0) for i in range(n)):
            return None
        if not np.allclose(lv[rows], lv[rows[0]], rtol=1e-14, atol=0.0):
            return None
        mono.append((b, cond, abs(complex(lv[rows[0]])) ** 2))
    return masks, values, mono


def n_mono(m):
    """Number of collapse operators handled by the general gather (one nonzero per row)."""
    return 0 if m.mcol is None else int(m.mcol.shape[0])


def jump_width(m):
    """Largest number of populations feeding one population (at least 1)."""
    return max(1, int(np.max(np.count_nonzero(m.jump, axis=1))))


def constants(m, runtime_values, split=None):
    """Flat constant buffer and the offsets the kernel reads (FP16 / BF16 operand parts are
    stored as the FP32 values they represent exactly)."""
    rc = np.asarray(runtime_values if runtime_values else [0.0], np.float64)
    jmax = jump_width(m)
    jw = np.zeros((m.n, jmax))
    jb = np.zeros((m.n, jmax))
    for a in range(m.n):
        for e, b in enumerate(np.nonzero(m.jump[a])[0]):
            jw[a, e], jb[a, e] = m.jump[a, b], b
    n = m.n
    big = np.zeros((2 * n, 2 * n))
    big[:n, :n], big[:n, n:] = m.h_eff.real, -m.h_eff.imag
    big[n:, :n], big[n:, n:] = m.h_eff.imag, m.h_eff.real
    mono = n_mono(m)
    mcol = m.mcol if mono else np.full((1, n), -1)
    ml = m.ml if mono else np.zeros((1, n), complex)
    _, ecol, evals = ell_table(m)
    parts = {"RC": rc, "APACK": pack_a(m).ravel(), "HBIG": big.ravel(),
             "ECOL": ecol.ravel().astype(np.float64), "ERE": evals.real.ravel(),
             "EIM": evals.imag.ravel(), "HDR": np.diag(m.h_eff).real.copy(),
             "HDI": np.diag(m.h_eff).imag.copy(),
             "DPACK": pack_a_dmfma(m).ravel(), "DEP": m.dep.ravel(),
             "JW": jw.ravel(), "JB": jb.ravel(), "VCOEF": m.vcoef.ravel(), "OBS": m.obs,
             "MCOL": mcol.ravel().astype(np.float64), "MRE": ml.real.ravel(),
             "MIM": ml.imag.ravel()}
    if split:
        hi, lo = pack_a_split(m, split)
        parts.update(AHI=hi.ravel(), ALO=lo.ravel())
    if two_sided_applies(n, split):
        lhi, llo, _, _ = pack_two_sided(m, split)
        parts.update(LHI=lhi.ravel(), LLO=llo.ravel())
    offs, buf, pos = {}, [], 0
    for k, v in parts.items():
        offs[k] = pos
        buf.append(v)
        pos += v.size
    return np.concatenate(buf).astype(np.float64), offs


LDS_BYTES_PER_CU = 160 * 1024      # gfx950


def lds_bytes(n, fp64, block, jmax, split=None, valu=False, mono=0, sparse_k=0, nbasis=0):
    """Static LDS of the kernel (the ``__shared__`` arrays of ``kernels/hilbert_rk4.hip`` and,
    for FP64 without ``valu``, ``kernels/hilbert_rk4_dmfma.hip``); ``mono`` general jump
    operators add a per-wavefront copy of the state and their row tables; ``nbasis`` drive
    basis terms add their row coefficients (``hilbert_rk4.hip``).  ``sparse_k`` > 0:
    ``kernels/hilbert_sparse.hip`` with that ELL width (a state copy per wavefront)."""
    if sparse_k:
        # state copies (rows padded to a power of two for the swizzled layout), the static
        # diagonal coefficients, drive rows (up to 8 basis terms), jump lists and, for the
        # table form, the ELL and jump-operator tables
        r = 8 if fp64 else 4
        npad = max(8, 1 << (n - 1).bit_length())
        return ((block // 64) * n * npad * 2 * r + n * n * 2 * r + 8 * n * r
                + n * jmax * (r + 4) + n * sparse_k * (4 + 2 * r) + mono * n * (4 + 2 * r))
    if fp64 and not valu:
        waves, pts = block // 64, (2 if n == 8 else 1)
        total = (n // 8) ** 2 * 4 * 64 * 8 + waves * (2 * pts * n * (n + 1) + pts * n + 64) * 8 \
            + n * n * 8 + n * jmax * 12
        if mono:
            total += waves * 2 * pts * n * (n + 1) * 8 + mono * n * (4 + 16)
        return total
    r = 8 if fp64 else 4
    if two_sided_applies(n, split) and not fp64:
        # kernels/hilbert_split.hip: coefficient operands (N = 32 only), the per-wavefront
        # populations and state copy (for its Hermitian part) and, with general jump
        # operators, a second state copy and their row tables
        waves, nb = block // 64, n // 16
        total = (2 * 2 * nb * nb * 64 * 16 if nb > 1 else 0) + waves * n * 4 \
            + waves * 2 * n * (n + 1) * 4 + nbasis * n * 4
        if mono:
            total += waves * 2 * n * (n + 1) * 4 + mono * n * 12
        return total
    if mono:
        waves, pts = block // 64, 32 // n
        extra = waves * 2 * pts * n * (n + 1) * r + mono * n * (4 + 2 * r)
        return lds_bytes(n, fp64, block, jmax, split, valu, nbasis=nbasis) + extra
    waves, pts, tiles = block // 64, 32 // n, (2 * n + 31) // 32
    total = waves * 2 * pts * n * (n + 1) * r + waves * pts * n * r + n * n * r
    total += n * jmax * (r + 4) + waves * 64 * 8 + nbasis * n * r
    if fp64:
        total += 4 * n * n * 8 + waves * pts * 2 * n * n * 8
    elif split:
        total += 2 * tiles * (n // 8) * 64 * 8 * 2
    else:
        total += tiles * n * 64 * 4
    return total


def block_for(m, fp64, block_size=256, valu=False, split=None, sparse=False):
    """Work-group size: the requested one, halved until the kernel's LDS fits a compute unit
    (the FP64 VALU kernel at N = 32, and the FP64 matrix-core kernel at N = 32 with general
    jump operators, run two waves per work-group)."""
    if block_size == "auto":
        return block_size
    b = int(block_size)
    kw = ell_table(m)[0] if sparse else 0
    while b > 64 and lds_bytes(m.n, fp64, b, jump_width(m), split, valu,
                               n_mono(m), kw, len(m.basis)) > LDS_BYTES_PER_CU:
        b //= 2
    return b


def emit(m, offs, mode, block_size=256, fp64=False, split=None, valu=False, trace=False,
         sparse=False, solver="rk4", count_steps=False):
    """hiprtc source of the Hilbert-space RK4 kernel for model m (``split``: the commutator
    GEMM as three FP16 or BF16 products of hi/lo parts with FP32 accumulation; FP64 runs on the
    FP64 matrix cores, or with ``valu`` as VALU dot products; ``sparse``: the wavefront-per-
    point kernel over an ELL table of H_eff, any precision, which also integrates with the
    other fixed-step methods and the adaptive ones, ``solver``; ``count_steps``: the adaptive
    kernel writes accepted / rejected step counts)."""
    if solver != "rk4" and not sparse:
        raise ValueError("the matrix-core forms integrate with RK4 only")
    split_kind = SPLIT_KINDS[split]
    split_name = split
    atoms = {}
    names = {"ParX": "px", "ParY": "py"}
    split = [_split_time_atoms(sp.sympify(b), {TIME}, m.runtime_syms, atoms) for b in m.basis]
    atom_list = [(sym, expr) for expr, sym in atoms.items()]
    for i, (sym, _) in enumerate(atom_list):
        names[sym.name] = f"TT[{i}]"
    for i, s in enumerate(m.runtime_syms):
        names[s.name] = f"C[{i}]"
    pr = HipPrinter(names)
    ftype = "double" if fp64 else "float"
    lines = [f"#define GQ_HN {m.n}", f"#define GQ_NBASIS {len(m.basis)}",
             f"#define GQ_JMAX {jump_width(m)}", f"#define GQ_NMONO {n_mono(m)}",
             f"#define GQ_NATOMS {len(atom_list)}", f"#define GQ_NATOMS_ALLOC {max(1, len(atom_list))}",
             f"#define GQ_HK {ell_table(m)[0]}"]
    xf = xor_form(m) if sparse and SPARSE_XOR else None
    lines.append(f"#define GQ_XOR {int(xf is not None)}")
    if xf is not None:
        masks, values, mono = xf
        lines.append(f"static constexpr int gq_xmask[{len(masks)}] = "
                     f"{{{', '.join(str(k) for k in masks)}}};")
        lines.append(f"static constexpr real gq_xre[{len(masks)}] = "
                     f"{{{', '.join(repr(float(v.real)) for v in values)}}};")
        lines.append(f"static constexpr real gq_xim[{len(masks)}] = "
                     f"{{{', '.join(repr(float(v.imag)) for v in values)}}};")
        if mono:
            lines.append(f"static constexpr int gq_mmask[{len(mono)}] = "
                         f"{{{', '.join(str(b) for b, _, _ in mono)}}};")
            lines.append(f"static constexpr int gq_mcond[{len(mono)}] = "
                         f"{{{', '.join(str(c) for _, c, _ in mono)}}};")
            lines.append(f"static constexpr real gq_mw[{len(mono)}] = "
                         f"{{{', '.join(repr(float(w)) for _, _, w in mono)}}};")
    for k, v in offs.items():
        lines.append(f"#define GQ_OFF_{k} {v}")
    dnames = dict(names)
    dnames["t"] = "t_now"
    for i, s in enumerate(m.runtime_syms):
        dnames[s.name] = f"(double)C[{i}]"
    dpr = HipPrinter(dnames, literal="DBL")
    lines.append("__device__ __forceinline__ void gq_atoms_f64(const double t_now, "
                 "const real* __restrict__ C, double* __restrict__ TT) {")
    lines.append("    (void)t_now; (void)C; (void)TT;")
    lines += [f"    TT[{i}] = {dpr.doprint(e)};" for i, (_, e) in enumerate(atom_list)]
    lines.append("}")
    lines.append(f"__device__ __forceinline__ void gq_basis(const {ftype} px, const {ftype} py, "
                 f"const {ftype}* __restrict__ TT, const {ftype}* __restrict__ C, "
                 f"{ftype}* __restrict__ g) {{")
    lines.append("    (void)px; (void)py; (void)TT; (void)C;")
    lines += [f"    g[{k}] = {pr.doprint(e)};" for k, e in enumerate(split)]
    lines.append("}")
    model_src = "\n".join(lines)
    with open(codegen._header_path()) as fh:
        header = fh.read()
    defines = {"GQ_HILBERT": 1, "GQ_OUT_MEAN": int(mode == "mean"),
               "GQ_OUT_FINAL": int(mode == "final"), "GQ_OUT_FINAL_RHO": int(mode == "final_rho"),
               "GQ_OUT_MEAN_STATE": 0, "GQ_TRACE": int(bool(trace)),
               "GQ_RHO0_VALUES": "GQ_RHO0_FROM_ARGS",
               "GQ_MFMA_SPLIT": split_kind}
    template = ("hilbert_sparse.hip" if sparse else
                "hilbert_rk4_dmfma.hip" if fp64 and not valu else
                "hilbert_split.hip" if two_sided_applies(m.n, split_kind and split_name) else
                "hilbert_rk4.hip")
    defines["GQ_HAS_DEP"] = int(bool(np.any(m.dep != 0)))
    table = "rk4_table.hip"
    if solver in ADAPTIVE:
        # stage times come from the step controller: atoms per stage in FP64, no table
        from . import tableaux
        model_src += "\n" + tableaux.emit_arrays(tableaux.DOP853 if solver == "dop853a"
                                                  else tableaux.TSIT5)
        defines.update({"GQ_ADAPTIVE": 1, "GQ_METHOD_DOP853": int(solver == "dop853a"),
                        "GQ_COUNT_STEPS": int(bool(count_steps))})
        table = "step_control.hip"
    elif solver != "rk4":
        from . import tableaux
        model_src += "\n" + tableaux.emit_fixed(solver)
        defines["GQ_TIME_TABLE"] = 1
        table = "tableau_table.hip"
    parts = [header, codegen.prelude(fp64, True, block_size, defines), model_src,
             codegen.read_kernel_template("common.hip"), codegen.read_kernel_template(table),
             codegen.read_kernel_template(template)]
    return "\n".join(parts)


def points_per_wave(n, fp64, valu=False, sparse=False, split=None):
    """Points per wavefront: 32 / N columns of one 32-column MFMA tile (FP32 forms, FP64 VALU),
    16 / N for the FP64 matrix-core form (one point in two column tiles at N = 32), one for the
    sparse form and the two-sided split-precision form."""
    if sparse or (two_sided_applies(n, split) and not fp64):
        return 1
    if fp64 and not valu:
        return 2 if n == 8 else 1
    return 32 // n


_SOURCES = {}
SPARSE_UNROLL_OPTIONS = ("-mllvm", "-pragma-unroll-threshold=200000")
ADAPTIVE = ("tsit5a", "dop853a")


def run(m, *, mode, var_arrays, rho0_values, runtime_values, dt, t_in, num_steps, warmup_steps,
        block_size=256, collect=False, options=(), fp64=False, split=None, valu=False,
        prepare=False, num_trace=0, trace_stride=1, trace_offset=0, sparse=False, spec=None):
    """Launch the Hilbert-space kernel; returns the same dict as ``device.run_sweep``
    (``valu``: the FP64 VALU form instead of the FP64 matrix-core form; ``spec``: the fixed-step
    solver, RK4 when None, other methods on the sparse form only).  ``prepare``: build
    buffers and kernels but launch nothing; returns the launch plan of ``device.run_sweep``
    (for ``prepare_sweep``), whose ``set_consts`` rewrites the runtime constants in place."""
    import torch
    from .device import _axis, _initial_states
    setup_start = time.time()
    ctx = runtime.Context.get()
    dev = torch.device("cuda", ctx.device)
    if split and fp64:
        raise ValueError("split-precision matrix-core products are an FP32-mode option")
    consts, offs = constants(m, runtime_values, split)
    uses_rho0 = rho0_values is not None
    if valu and n_mono(m):
        raise ValueError("the FP64 VALU form has no general jump operators")
    if sparse and (split or valu):
        raise ValueError("the sparse form has no split-precision or VALU variant")
    solver = spec.name if spec is not None else "rk4"
    adaptive = solver in ADAPTIVE
    count_steps = adaptive and bool(spec.options.get("count_steps"))
    if adaptive:
        if not sparse:
            raise ValueError("the matrix-core forms integrate with fixed-step methods only")
        if prepare:
            raise ValueError("adaptive sweeps on the sparse form are not prepared (graph) sweeps")
        # one wavefront per work-group: a work-group ends with its point, so the hardware's
        # dispatch of new work-groups balances points of different step counts
        block_size = 64
    block_size = block_for(m, fp64, block_size, valu, split, sparse)
    tune = block_size == "auto"
    if tune:
        block_size = block_for(m, fp64, 256, valu, split, sparse)
    trace = num_trace > 0
    skey = (m.key, mode, block_size, uses_rho0, tuple(options), bool(fp64), split, bool(valu),
            trace, bool(sparse), bool(sparse and SPARSE_XOR), solver, count_steps)
    src = _SOURCES.get(skey)
    if src is None:
        src = emit(m, offs, mode, block_size, fp64, split, valu, trace, sparse, solver,
                   count_steps).replace(
            "GQ_RHO0_FROM_ARGS", "1" if uses_rho0 else "0")
        _SOURCES[skey] = src
    options = codegen.HIPRTC_BASE_OPTIONS + tuple(options)
    if solver != "rk4":
        # unroll every stage of the step, whatever its size, so the stage vectors stay in
        # registers (DOP853 at 32 levels: 8.3 against 20 ns per RHS and point in FP64)
        options += SPARSE_UNROLL_OPTIONS
    module, cached, build_s = ctx.compile(src, options)
    rdt = torch.float64 if fp64 else torch.float32
    axes = list(var_arrays.values())
    px = _axis(axes[0], rdt, dev)
    py = _axis(axes[1], rdt, dev) if len(axes) > 1 else torch.zeros(1, dtype=rdt, device=dev)
    nx, ny = int(px.numel()), int(py.numel())
    npts = nx * ny
    d = m.n * m.n - 1
    rho0 = _initial_states(rho0_values, nx, ny, d, rdt, dev) if uses_rho0 else None
    cbuf = torch.as_tensor(consts, dtype=rdt, device=dev)
    if mode == "final_rho":
        result = torch.empty((d, max(npts, 1)), dtype=rdt, device=dev)
    else:
        result = torch.empty(max(npts, 1), dtype=torch.complex128 if fp64 else torch.complex64,
                             device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    trace_buf = (torch.zeros((num_trace, max(npts, 1)),
                             dtype=torch.complex128 if fp64 else torch.complex64, device=dev)
                 if trace else None)
    from . import tableaux
    rows = (tableaux.rows_per_step(solver) * num_steps + 1) if not adaptive else 0
    table_kernel = codegen.kernel_name("gq_rk4_time_table" if solver == "rk4"
                                       else "gq_fixed_time_table", fp64)
    n_atoms = int(src.split("#define GQ_NATOMS ")[1].split()[0]) if not adaptive else 0
    steps_out = (torch.zeros((max(npts, 1), 2), dtype=torch.int32, device=dev)
                 if count_steps else None)
    table = torch.empty(max(1, rows * max(1, n_atoms)), dtype=rdt, device=dev)
    a = runtime.SweepArgs()
    a.dt, a.t_in = float(dt), float(t_in)
    a.num_x, a.num_y = nx, ny
    a.point_begin, a.point_count = 0, npts
    a.num_steps, a.warmup_steps = int(num_steps), int(warmup_steps)
    a.trace_stride, a.trace_offset, a.num_trace = int(trace_stride), int(trace_offset), int(num_trace)
    a.trace_point_stride = npts
    a.trace = trace_buf.data_ptr() if trace_buf is not None else None
    a.par_x, a.par_y, a.consts = px.data_ptr(), py.data_ptr(), cbuf.data_ptr()
    a.rho0 = rho0.data_ptr() if rho0 is not None else None
    a.time_table = table.data_ptr()
    a.result, a.status = result.data_ptr(), status.data_ptr()
    for i, v in enumerate(spec.parameters if spec is not None else ()):
        a.solver_param[i] = float(v)
    a.adaptive_steps = steps_out.data_ptr() if steps_out is not None else None
    stream = torch.cuda.current_stream(dev).cuda_stream
    kname = codegen.kernel_name(kernel_base(split, fp64, valu, sparse), fp64)
    tune_s, tune_times = 0.0, {}
    if tune and npts > 0 and num_steps > 0:
        from . import autotune
        from .device import _timed_launches
        t0 = time.time()
        if n_atoms > 0:
            ctx.launch(module, table_kernel, (rows + 255) // 256,
                       256, a, stream)
        trial_steps = min(int(num_steps), 64)
        jmax = jump_width(m)
        kw = ell_table(m)[0] if sparse else 0
        fits = tuple(b for b in autotune.CANDIDATES
                     if lds_bytes(m.n, fp64, b, jmax, split, valu, n_mono(m), kw,
                                  len(m.basis)) <= LDS_BYTES_PER_CU)
        spare = torch.zeros(1, dtype=torch.int32, device=dev)

        def trial(mod, block):
            t = runtime.SweepArgs.from_buffer_copy(a)
            t.num_steps, t.warmup_steps, t.num_trace, t.trace = trial_steps, 0, 0, None
            per = (block // 64) * points_per_wave(m.n, fp64, valu, sparse, split)
            return _timed_launches(ctx, mod, kname, (npts + per - 1) // per, block, t, stream,
                                   spare)

        if npts * trial_steps >= autotune.MIN_TRIAL_POINT_STEPS:
            block_size, tune_times = autotune.choose(ctx, src, options, trial, fits)
        else:
            block_size = autotune.lookup(ctx, autotune.source_key(src, options, ctx.arch)) \
                or block_size
        module, _, _ = ctx.compile(autotune.with_block(src, block_size), options)
        ctx.read_status(status, stream, clear=True)
        tune_s = time.time() - t0
    per_block = (block_size // 64) * points_per_wave(m.n, fp64, valu, sparse, split)
    if prepare:
        launches = []
        if npts > 0:
            if n_atoms > 0:
                launches.append((table_kernel,
                                 (rows + 255) // 256, 256))
            launches.append((kname, (npts + per_block - 1) // per_block, block_size))
        n_rc = len(m.runtime_syms)

        def set_consts(values):
            # the runtime constants lead the constant buffer ("RC"); the packed operators
            # do not depend on them
            cbuf[:n_rc].copy_(torch.as_tensor(np.asarray(values, np.float64), dtype=rdt,
                                              device=dev))
        return {"ctx": ctx, "module": module, "args": a, "launches": launches, "fp64": fp64,
                "mode": mode, "nx": nx, "ny": ny, "size": d, "num_trace": int(num_trace),
                "state_trace": False, "adaptive": False, "set_consts": set_consts,
                "runtime_syms": tuple(m.runtime_syms),
                "buffers": {"px": px, "py": py, "consts": cbuf, "rho0": rho0, "result": result,
                            "trace": trace_buf, "status": status, "cursor": None,
                            "table": table, "steps": None},
                "info": {"build_s": build_s, "cached_kernel": cached,
                         "block_size": int(block_size), "lanes_per_point": 1,
                         "autotune_s": tune_s, "autotune": tune_times, "kernel": kname}}
    timer = runtime.EventTimer(ctx, stream) if collect else None
    setup_s = time.time() - setup_start
    if npts > 0:
        if timer:
            timer.record_start()
        if n_atoms > 0:
            ctx.launch(module, table_kernel, (rows + 255) // 256,
                       256, a, stream)
        ctx.launch(module, kname, (npts + per_block - 1) // per_block, block_size, a, stream)
        if timer:
            timer.record_stop()
    flags = ctx.read_status(status, stream)
    gpu_s = timer.elapsed_ms() / 1000.0 if timer else 0.0
    if timer:
        timer.close()
    if mode == "final_rho":
        out = result[:, :npts].t().reshape(nx, ny, d).contiguous()
    else:
        out = result[:npts].reshape(nx, ny)
    tr = (trace_buf[:, :npts].t().reshape(nx, ny, num_trace).contiguous()
          if trace_buf is not None else None)
    return {"result": out, "trace": tr, "status": flags, "gpu_s": gpu_s, "build_s": build_s,
            "cached_kernel": cached, "setup_s": setup_s,
            "adaptive_steps": (steps_out[:npts].reshape(nx, ny, 2).cpu().numpy()
                               if steps_out is not None else None),
            "block_size": int(block_size), "autotune_s": tune_s, "autotune": tune_times,
            "kernel": kname}
