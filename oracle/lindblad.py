"""Numeric reduced-coordinate Lindblad generator (FP64 oracle; test infrastructure only).

The reference derives the reduced real equations symbolically (``gqis/solver.py:399-493``).
This module builds the same affine map ``dy/dt = M y + b`` numerically from the operator
matrices, so it is an independent check of the symbolic front end.

Reduced layout (``gqis/solver.py:202-253``): the first ``N-1`` populations ``rho[i,i]``,
then ``Re rho[i,j]``, ``Im rho[i,j]`` for ``i < j`` in row-major order; the last population is
``1 - sum(populations)``.
"""
import numpy as np


def reduced_size(n):
    return n * n - 1


def upper_pairs(n):
    """Coherence index pairs (i, j), i < j, in row-major order (``gqis/solver.py:225-227``)."""
    return [(i, j) for i in range(n) for j in range(i + 1, n)]


def coords(x):
    """Reduced real coordinates of an N x N matrix (trace part dropped)."""
    n = x.shape[0]
    out = [x[k, k].real for k in range(n - 1)]
    for i, j in upper_pairs(n):
        out.append(x[i, j].real)
        out.append(x[i, j].imag)
    return np.array(out, dtype=np.float64)


def reduced_basis(n):
    """Matrices B_k with rho(y) = E_{n-1,n-1} + sum_k y_k B_k (unit trace built in)."""
    basis = []
    for k in range(n - 1):
        b = np.zeros((n, n), complex)
        b[k, k] = 1.0
        b[n - 1, n - 1] = -1.0
        basis.append(b)
    for i, j in upper_pairs(n):
        b = np.zeros((n, n), complex)
        b[i, j] = b[j, i] = 1.0
        basis.append(b)
        b = np.zeros((n, n), complex)
        b[i, j], b[j, i] = 1j, -1j
        basis.append(b)
    return basis


def base_state(n):
    b = np.zeros((n, n), complex)
    b[n - 1, n - 1] = 1.0
    return b


def reduced_to_rho(y):
    """Full density matrix from a reduced vector (``gqis/solver.py:229-248``)."""
    d = len(y)
    n = int(round(np.sqrt(d + 1)))
    rho = base_state(n).astype(complex)
    for yk, bk in zip(y, reduced_basis(n)):
        rho = rho + yk * bk
    return rho


def rho_to_reduced(rho):
    """Reduced vector of a density matrix (``gqis/solver.py:256-281`` matrix branch)."""
    return coords(np.asarray(rho, complex))


def lindblad(h, collapse, rho):
    """-i[H, rho] + sum_L (L rho L^+ - 1/2 {L^+ L, rho}) (textbook form, ``README.md:196-206``)."""
    out = -1j * (h @ rho - rho @ h)
    for op in collapse:
        opd = op.conj().T
        opdop = opd @ op
        out = out + op @ rho @ opd - 0.5 * (opdop @ rho + rho @ opdop)
    return out


def affine_generator(h, collapse=()):
    """(M, b) with d y/dt = M y + b for the reduced coordinates of the Lindblad equation."""
    h = np.asarray(h, complex)
    n = h.shape[0]
    collapse = [np.asarray(c, complex) for c in collapse]
    b = coords(lindblad(h, collapse, base_state(n)))
    # The Lindblad map is linear in rho, so each column is the image of one basis matrix.
    m = np.stack([coords(lindblad(h, collapse, bk)) for bk in reduced_basis(n)], axis=1)
    return m, b


def observable_form(op):
    """(o0, o) complex with Tr(O rho(y)) = o0 + o . y (``gqis/solver.py:431-433``)."""
    op = np.asarray(op, complex)
    n = op.shape[0]
    o0 = np.trace(op @ base_state(n))
    o = np.array([np.trace(op @ bk) for bk in reduced_basis(n)])
    return o0, o
