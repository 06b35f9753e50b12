"""Numeric benchmark models for the oracle (test infrastructure only).

A Lindblad sweep model is affine in the reduced state:
``dy/dt = sum_k g_k(t, px, py) (M_k y + b_k)`` with scalar coefficient functions ``g_k``;
the observable is ``o0 + o . y`` (complex).  ``AffineSweep`` turns one into the vectorised
``f(t, y)`` / ``obs(t, y)`` callables that ``oracle.integrate`` expects, for a flattened
``(nx, ny)`` grid in the reference's row-major point order (``gqis/N_Level_Kernel.cu:92-99``).
"""
import numpy as np

from . import lindblad as lb

SZ = np.array([[1, 0], [0, -1]], complex)
SX = np.array([[0, 1], [1, 0]], complex)
SM = np.array([[0, 1], [0, 0]], complex)   # the reference's sm (Benchmarks/Benchmark_01_two_level.py:187)


class AffineSweep:
    def __init__(self, terms, obs_op, px, py):
        """terms: list of (g(t, px, py) -> scalar or (P,) array, H_k, collapse_k).

        Each term contributes the Lindblad generator of (H_k, collapse_k) scaled by g_k; the
        collapse operators of all terms must sit in a term whose g is the constant 1 (the
        dissipator is not linear in a scaled operator).
        """
        self.px = np.repeat(np.asarray(px, np.float64), len(py))
        self.py = np.tile(np.asarray(py, np.float64), len(px))
        self.parts = []
        for g, h, col in terms:
            m, b = lb.affine_generator(h, col)
            self.parts.append((g, m, b))
        self.o0, self.o = lb.observable_form(obs_op)
        self.npts = self.px.size
        self.dim = self.parts[0][1].shape[0]

    def f(self, t, y, idx=None):
        """dy/dt at time t for the points idx (all points when None)."""
        px = self.px if idx is None else self.px[idx]
        py = self.py if idx is None else self.py[idx]
        out = np.zeros_like(y)
        for g, m, b in self.parts:
            coef = g(t, px, py) if callable(g) else g
            out += np.asarray(coef) * (m @ y + b[:, None])
        return out

    def obs(self, t, y, idx=None):
        return self.o0 + self.o @ y

    def ground_state(self):
        n = int(round(np.sqrt(self.dim + 1)))
        y = np.zeros((self.dim, self.npts))
        if n > 1:
            y[0] = 1.0   # |0><0|: rho[0] = 1, rest zero (gqis/solver.py:1661-1662)
        return y


def one(t, px, py):
    return 1.0


def lzs_two_level(px, py, omega=1.0, delta=1.0):
    """Two-level Landau-Zener-Stuckelberg qubit: H = (eps/2) sz + (Delta/2) sx,
    drive A cos(omega t) sz/2, collapse sqrt(0.01) s-, sqrt(0.005) sz, observable sz."""
    terms = [
        (one, delta / 2 * SX, [np.sqrt(0.01) * SM, np.sqrt(0.005) * SZ]),
        (lambda t, x, y: x, SZ / 2, []),
        (lambda t, x, y: y * np.cos(omega * t), SZ / 2, []),
    ]
    return AffineSweep(terms, SZ, px, py)


def four_level_chain(px, py, tunnel=0.5):
    """Four-level interferometer: diag(eps, -eps, eps+2, -eps-2) + nearest-neighbour tunnelling,
    drive A cos(t) diag(1,-1,1,-1)/2, relaxation sqrt(0.02)|i><i+1|, dephasing
    sqrt(0.01) diag(1,-1,1,-1), observable |0><0|."""
    h0 = np.diag([0.0, 0.0, 2.0, -2.0]).astype(complex)
    for i in range(3):
        h0[i, i + 1] = h0[i + 1, i] = tunnel
    zz = np.diag([1.0, -1.0, 1.0, -1.0]).astype(complex)
    col = []
    for i in range(3):
        op = np.zeros((4, 4), complex)
        op[i, i + 1] = np.sqrt(0.02)
        col.append(op)
    col.append(np.sqrt(0.01) * zz)
    terms = [(one, h0, col),
             (lambda t, x, y: x, zz, []),
             (lambda t, x, y: y * np.cos(t), zz / 2, [])]
    proj = np.zeros((4, 4), complex)
    proj[0, 0] = 1.0
    return AffineSweep(terms, proj, px, py)


def random_n_level_operators(n, seed):
    """Seeded operators of the N-level scaling model: H0 GUE with E|G_ij|^2 = 1,
    H1 and H2 and the observable random diagonal in [-1, 1]."""
    rng = np.random.default_rng(seed)
    g = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)
    h0 = (g + g.conj().T) / 2.0
    h1 = np.diag(rng.uniform(-1.0, 1.0, n)).astype(complex)
    h2 = np.diag(rng.uniform(-1.0, 1.0, n)).astype(complex)
    obs = np.diag(rng.uniform(-1.0, 1.0, n)).astype(complex)
    return h0, h1, h2, obs


def random_n_level(n, seed, px, py, rate=0.01):
    h0, h1, h2, obs = random_n_level_operators(n, seed)
    col = []
    for i in range(n - 1):
        op = np.zeros((n, n), complex)
        op[i, i + 1] = np.sqrt(rate)
        col.append(op)
    terms = [(one, h0, col),
             (lambda t, x, y: x, h1, []),
             (lambda t, x, y: y * np.cos(t), h2, [])]
    return AffineSweep(terms, obs, px, py)


def multi_qubit(n_qubits, px, py, delta=1.0, w_over_delta=1.14, gamma1_per_t=0.05,
                gamma_phi_per_t=0.04, gap_step=0.05, j_over_delta=0.1):
    """The reference's multi-qubit chain (``Benchmarks/Benchmark_04_hilbert_space_scaling.py:
    128-173``): gaps delta (1 + gap_step i) sx_i / 2, J sz_i sz_{i+1}, drive (eps + A sin(w t))
    sz_0 / 2, collapse sqrt(gamma1) s-_i with s- = [[0, 1], [0, 0]] and sqrt(gamma2) sz_i,
    observable qubit 0 in |1><1|."""
    sx = np.array([[0, 1], [1, 0]], complex)
    sm = np.array([[0, 1], [0, 0]], complex)

    def on(op, index):
        out = np.ones((1, 1), complex)
        for k in range(n_qubits):
            out = np.kron(out, op if k == index else np.eye(2))
        return out
    w = w_over_delta * delta
    period = 2.0 * np.pi / w
    gamma1 = gamma1_per_t / period
    gamma2 = gamma1 / 2.0 + gamma_phi_per_t / period
    h0 = sum(0.5 * delta * (1.0 + gap_step * i) * on(sx, i) for i in range(n_qubits))
    h0 = h0 + sum(j_over_delta * delta * on(SZ, i) @ on(SZ, i + 1) for i in range(n_qubits - 1))
    col = []
    for i in range(n_qubits):
        col += [np.sqrt(gamma1) * on(sm, i), np.sqrt(gamma2) * on(SZ, i)]
    hz = on(SZ, 0) / 2
    terms = [(one, h0, col),
             (lambda t, x, y: x, hz, []),
             (lambda t, x, y: y * np.sin(w * t), hz, [])]
    return AffineSweep(terms, on(np.diag([0.0, 1.0]).astype(complex), 0), px, py)


def qubit_resonator_operators(n_photon):
    """Qubit (x) cavity operators of the reference's qubit-resonator model
    (``Benchmarks/Benchmark_04_hilbert_space_scaling.py:63-94``): s- = [[0, 0], [1, 0]] on the
    qubit, sz on the qubit, a with a[k-1, k] = sqrt(k) on the cavity."""
    lower = np.array([[0, 0], [1, 0]], complex)
    ann = np.diag(np.sqrt(np.arange(1, n_photon)), 1).astype(complex)
    eye = np.eye(n_photon)
    return np.kron(lower, eye), np.kron(SZ, eye), np.kron(np.eye(2), ann)


def qubit_resonator(n_photon, px, py, p):
    """The qubit-resonator sweep (``Benchmarks/Benchmark_02_four_level_Interferometry.py:
    140-168``) as three affine terms: Ap (a + a^+) with the collapse operators sqrt(gamma1) s-,
    sqrt(gamma2) sz, sqrt(kappa) a; g1 delta (s+ a + s- a^+) scaled by 1 / E; sz scaled by
    (E - wr2) / 2; E = sqrt(delta^2 + (eps + A cos(w t))^2).  ``p`` holds delta, w, wr2,
    gamma1, gamma2, kappa, ap, g1.  Observable a."""
    sm, sz, a = qubit_resonator_operators(n_photon)

    def energy(t, x, y):
        return np.sqrt(p["delta"] ** 2 + (x + y * np.cos(p["w"] * t)) ** 2)
    col = [np.sqrt(p["gamma1"]) * sm, np.sqrt(p["gamma2"]) * sz, np.sqrt(p["kappa"]) * a]
    jc = p["g1"] * p["delta"] * (sm.conj().T @ a + sm @ a.conj().T)
    terms = [(one, p["ap"] * (a + a.conj().T), col),
             (lambda t, x, y: 1.0 / energy(t, x, y), jc, []),
             (lambda t, x, y: 0.5 * (energy(t, x, y) - p["wr2"]), sz, [])]
    return AffineSweep(terms, a, px, py)


class Duffing:
    """x'' + delta x' + alpha x + beta x^3 = gamma cos(w t) as (x, v); sweep (gamma, x0)."""

    def __init__(self, gamma, x0, delta=0.2, alpha=-1.0, beta=1.0, w=1.2):
        self.gamma = np.repeat(np.asarray(gamma, np.float64), len(x0))
        self.x0 = np.tile(np.asarray(x0, np.float64), len(gamma))
        self.delta, self.alpha, self.beta, self.w = delta, alpha, beta, w
        self.npts = self.gamma.size

    def f(self, t, y, idx=None):
        gam = self.gamma if idx is None else self.gamma[idx]
        x, v = y
        return np.stack([v, -self.delta * v - self.alpha * x - self.beta * x ** 3
                         + gam * np.cos(self.w * t)])

    def y0(self):
        return np.stack([self.x0, np.zeros_like(self.x0)])
