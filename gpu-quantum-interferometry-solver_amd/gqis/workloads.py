"""Symbolic definitions of the benchmark sweep workloads (inputs for ``mesolve_2D`` /
``odesolve_2D``), with the sizes they are quoted at.

* ``lzs_two_level``: two-level Landau-Zener-Stuckelberg qubit, 64 x 64 grid, 2000 RK4 steps.
* ``four_level``: four-level interferometer, 2048 x 2048 grid, t in [0, 100], dt = 0.05.
* ``random_n_level``: N-level scaling model (seeded GUE H0, random diagonal H1/H2/observable).
* ``duffing``: forced Duffing oscillator, 1024 x 1024 grid over (gamma, x0), t in [0, 500].
"""
import numpy as np
import sympy as sp

EPS, AMP, T, DRIVE = sp.symbols("eps A t Drive", real=True)


def lzs_two_level(n_eps=64, n_amp=64, n_t=2001, t_end=200.0):
    """H = (eps/2) sz + (Delta/2) sx, Delta = 1; drive A cos(t) sz/2; collapse sqrt(0.01) s-
    (s- = [[0,1],[0,0]]) and sqrt(0.005) sz; observable sz; rho0 = |0><0|."""
    sz = sp.Matrix([[1, 0], [0, -1]])
    sx = sp.Matrix([[0, 1], [1, 0]])
    sm = sp.Matrix([[0, 1], [0, 0]])
    h = EPS / 2 * sz + sp.Rational(1, 2) * sx + DRIVE * sz / 2
    col = [sp.sqrt(sp.Rational(1, 100)) * sm, sp.sqrt(sp.Rational(5, 1000)) * sz]
    return {
        "H": h, "Drive": {DRIVE: AMP * sp.cos(T)}, "Col_Ops": col, "mean_operator": sz,
        "tlist": np.linspace(0.0, t_end, n_t),
        "var_arrays": {EPS: np.linspace(-10.0, 10.0, n_eps), AMP: np.linspace(0.0, 20.0, n_amp)},
    }


def four_level(n_eps=2048, n_amp=2048, n_t=2001, t_end=100.0, tunnel=sp.Rational(1, 2)):
    """H = diag(eps, -eps, eps+2, -eps-2) + nearest-neighbour tunnelling 0.5
    + A cos(t) diag(1,-1,1,-1)/2; collapse sqrt(0.02)|i><i+1| (i = 0..2) and
    sqrt(0.01) diag(1,-1,1,-1); observable |0><0|."""
    zz = sp.diag(1, -1, 1, -1)
    h = sp.diag(EPS, -EPS, EPS + 2, -EPS - 2) + DRIVE * zz / 2
    for i in range(3):
        h[i, i + 1] = h[i + 1, i] = tunnel
    col = []
    for i in range(3):
        op = sp.zeros(4)
        op[i, i + 1] = sp.sqrt(sp.Rational(2, 100))
        col.append(op)
    col.append(sp.sqrt(sp.Rational(1, 100)) * zz)
    proj = sp.zeros(4)
    proj[0, 0] = 1
    return {
        "H": h, "Drive": {DRIVE: AMP * sp.cos(T)}, "Col_Ops": col, "mean_operator": proj,
        "tlist": np.linspace(0.0, t_end, n_t),
        "var_arrays": {EPS: np.linspace(-15.0, 15.0, n_eps), AMP: np.linspace(0.0, 30.0, n_amp)},
    }


def random_n_level_operators(n, seed):
    """Seeded numeric operators: H0 = (G + G^+)/2 with complex normal E|G_ij|^2 = 1;
    H1, H2 and the observable random diagonal in [-1, 1]."""
    rng = np.random.default_rng(seed)
    g = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)
    h0 = (g + g.conj().T) / 2.0
    h1 = rng.uniform(-1.0, 1.0, n)
    h2 = rng.uniform(-1.0, 1.0, n)
    obs = rng.uniform(-1.0, 1.0, n)
    return h0, h1, h2, obs


def random_n_level(n, seed=2024, n_eps=256, n_amp=256, n_t=2001, t_end=20.0, rate=0.01):
    h0, h1, h2, obs = random_n_level_operators(n, seed)

    def num(z):
        z = complex(z)
        return sp.Float(z.real, 17) + sp.I * sp.Float(z.imag, 17) if z.imag else sp.Float(z.real, 17)
    h = sp.Matrix(n, n, lambda i, j: num(h0[i, j]))
    h += EPS * sp.diag(*[sp.Float(v, 17) for v in h1]) + DRIVE * sp.diag(*[sp.Float(v, 17) for v in h2])
    col = []
    for i in range(n - 1):
        op = sp.zeros(n)
        op[i, i + 1] = sp.sqrt(sp.Float(rate, 17))
        col.append(op)
    return {
        "H": h, "Drive": {DRIVE: AMP * sp.cos(T)}, "Col_Ops": col,
        "mean_operator": sp.diag(*[sp.Float(v, 17) for v in obs]),
        "tlist": np.linspace(0.0, t_end, n_t),
        "var_arrays": {EPS: np.linspace(-5.0, 5.0, n_eps), AMP: np.linspace(0.0, 5.0, n_amp)},
    }


X, V, GAMMA, X0 = sp.symbols("x v gamma x0", real=True)


def duffing(n_gamma=1024, n_x0=1024, t_end=500.0, n_out=5001, delta=0.2, alpha=-1.0, beta=1.0,
            w=1.2):
    """x'' + delta x' + alpha x + beta x^3 = gamma cos(w t) as (x, v); sweep (gamma, x0),
    v0 = 0; tlist is the output grid (dt_out = 0.1 for 5001 points over [0, 500])."""
    rhs = [V, -delta * V - alpha * X - beta * X ** 3 + GAMMA * sp.cos(w * T)]
    return {
        "rhs": rhs, "state_symbols": [X, V], "y0": [X0, 0],
        "tlist": np.linspace(0.0, t_end, n_out),
        "var_arrays": {GAMMA: np.linspace(0.1, 0.6, n_gamma), X0: np.linspace(-2.0, 2.0, n_x0)},
    }


def chain(n, n_eps=64, n_amp=64, eps_range=(-10.0, 10.0), amp_max=20.0, n_t=2001, t_end=100.0):
    """N-level generalisation of the four-level interferometer: detunings
    (-1)^i (eps + 2 floor(i/2)), nearest-neighbour tunnelling 0.5, drive
    A cos(t) diag((-1)^i)/2, relaxation sqrt(0.02)|i><i+1|, dephasing sqrt(0.01) diag((-1)^i),
    observable |0><0| (n = 4 is ``four_level``)."""
    signs = [(-1) ** i for i in range(n)]
    h = sp.diag(*[s * (EPS + 2 * (i // 2)) for i, s in enumerate(signs)])
    h += DRIVE * sp.diag(*signs) / 2
    for i in range(n - 1):
        h[i, i + 1] = h[i + 1, i] = sp.Rational(1, 2)
    col = []
    for i in range(n - 1):
        op = sp.zeros(n)
        op[i, i + 1] = sp.sqrt(sp.Rational(2, 100))
        col.append(op)
    col.append(sp.sqrt(sp.Rational(1, 100)) * sp.diag(*signs))
    proj = sp.zeros(n)
    proj[0, 0] = 1
    return {
        "H": h, "Drive": {DRIVE: AMP * sp.cos(T)}, "Col_Ops": col, "mean_operator": proj,
        "tlist": np.linspace(0.0, t_end, n_t),
        "var_arrays": {EPS: np.linspace(*eps_range, n_eps), AMP: np.linspace(0.0, amp_max, n_amp)},
    }


def multi_qubit(n_qubits, n_eps=64, n_amp=64, n_t=2001, t_end=20.0, delta=1.0, w_over_delta=1.14,
                gamma1_per_t=0.05, gamma_phi_per_t=0.04, gap_step=0.05, j_over_delta=0.1,
                max_over_w=10.0):
    """The reference's multi-qubit scaling model (``Benchmarks/Benchmark_04_hilbert_space_
    scaling.py:128-173``, defaults of ``:1084-1093``): n qubits on a chain (dimension 2^n) with
    gaps delta (1 + gap_step i) sx_i / 2, ZZ coupling J sz_i sz_j on neighbours, the drive
    (eps + A sin(w t)) sz_0 / 2 on qubit 0, collapse sqrt(gamma1) s-_i (s- = [[0,1],[0,0]]) and
    sqrt(gamma2) sz_i with gamma1 = gamma1_per_T / T, gamma2 = gamma1/2 + gamma_phi_per_T / T,
    T = 2 pi / w; observable: qubit 0 in |1><1|.  eps in [-eps_max, eps_max], A in [0, A_max]
    with both maxima max_over_w * w."""
    sx = np.array([[0.0, 1.0], [1.0, 0.0]])
    sz = np.array([[1.0, 0.0], [0.0, -1.0]])
    sm = np.array([[0.0, 1.0], [0.0, 0.0]])
    excited = np.array([[0.0, 0.0], [0.0, 1.0]])

    def on_qubit(op, index):
        out = np.ones((1, 1))
        for k in range(n_qubits):
            out = np.kron(out, op if k == index else np.eye(2))
        return out
    w = w_over_delta * delta
    period = 2.0 * np.pi / w
    gamma1 = gamma1_per_t / period
    gamma2 = gamma1 / 2.0 + gamma_phi_per_t / period
    h = np.zeros((2 ** n_qubits, 2 ** n_qubits))
    col = []
    for i in range(n_qubits):
        h += 0.5 * delta * (1.0 + gap_step * i) * on_qubit(sx, i)
        col += [np.sqrt(gamma1) * on_qubit(sm, i), np.sqrt(gamma2) * on_qubit(sz, i)]
    for i in range(n_qubits - 1):
        h += j_over_delta * delta * on_qubit(sz, i) @ on_qubit(sz, i + 1)
    big = max_over_w * w
    return {
        "H": sp.Matrix(h) + DRIVE * sp.Matrix(on_qubit(sz, 0)) / 2,
        "Drive": {DRIVE: EPS + AMP * sp.sin(w * T)},
        "Col_Ops": [sp.Matrix(c) for c in col], "mean_operator": sp.Matrix(on_qubit(excited, 0)),
        "tlist": np.linspace(0.0, t_end, n_t),
        "var_arrays": {EPS: np.linspace(-big, big, n_eps), AMP: np.linspace(0.0, big, n_amp)},
        "numeric": {"h0": h, "hz": on_qubit(sz, 0) / 2, "col": col,
                    "obs": on_qubit(excited, 0), "w": w},
    }


def qubit_resonator_parameters(regime="wd500"):
    """Physical constants of the reference's qubit-resonator benchmark
    (``Benchmarks/Benchmark_02_four_level_Interferometry.py:851-861`` with the defaults of
    ``:1074-1091``; the same conversion in ``Benchmarks/Benchmark_04_hilbert_space_scaling.py:
    97-109``): wq1 = 4.71 * 1.15 scales the normalised inputs."""
    wq1 = 4.71 * 1.15
    w_over_delta, eps_over_w, amp_over_w = {
        "wd500": (0.0923105326317733, 22.74020028, 27.8315904255),
        "wd1500": (0.27693159789532, 7.58006676, 12.0603558511)}[regime]
    delta = wq1 * 1.0
    w = w_over_delta * delta
    gamma1 = 2.0e-3
    return {"delta": delta, "w": w, "wr2": 7.6767, "gamma1": gamma1, "gamma2": gamma1 / 2.0,
            "kappa": 5.0e-3, "ap": 2.0e-5 * wq1, "g1": 0.04033970276 / 1.23 * wq1,
            "eps_max": eps_over_w * w, "amp_max": amp_over_w * w}


def qubit_resonator(n_photon=2, n_eps=2048, n_amp=2048, periods=40, steps_per_period=146,
                    regime="wd500"):
    """The reference's qubit-resonator model with ``n_photon`` Fock states (dimension
    2 n_photon; ``Benchmarks/Benchmark_02_four_level_Interferometry.py:140-168`` for
    n_photon = 2, ``Benchmarks/Benchmark_04_hilbert_space_scaling.py:73-94`` in general):
    with E = sqrt(delta^2 + Drive^2) and Drive = eps + A cos(w t),
    H = Ap (a + a^+) + g1 delta (s+ a + s- a^+) / E + sz (E - wr2) / 2 on qubit (x) cavity,
    s- = [[0,0],[1,0]] on the qubit, a the cavity's annihilation operator; collapse
    sqrt(gamma1) s-, sqrt(gamma2) sz, sqrt(kappa) a; observable a (complex mean).  The grid is
    eps in [-eps_max, eps_max], A in [0, A_max]; ``periods`` drive periods of
    ``steps_per_period`` RK4 steps (146 is the reference's published 2048 x 2048 row,
    ``Benchmarks/results/Benchmark_02_full_benchmark.csv:26``)."""
    p = qubit_resonator_parameters(regime)
    sz = sp.Matrix([[1, 0], [0, -1]])
    lower = sp.Matrix([[0, 0], [1, 0]])
    ann = sp.zeros(n_photon)
    for k in range(1, n_photon):
        ann[k - 1, k] = float(np.sqrt(k))
    s_minus = sp.kronecker_product(lower, sp.eye(n_photon))
    s_z = sp.kronecker_product(sz, sp.eye(n_photon))
    a = sp.kronecker_product(sp.eye(2), ann)
    energy = sp.sqrt(p["delta"] ** 2 + DRIVE ** 2)
    h = (p["ap"] * (a.H + a) + p["g1"] * p["delta"] * (s_minus.H * a + s_minus * a.H) / energy
         + s_z / 2 * (energy - p["wr2"]))
    col = [np.sqrt(p["gamma1"]) * s_minus, np.sqrt(p["gamma2"]) * s_z, np.sqrt(p["kappa"]) * a]
    period = 2.0 * np.pi / p["w"]
    steps = int(round(periods * steps_per_period))
    return {
        "H": h, "Drive": {DRIVE: EPS + AMP * sp.cos(p["w"] * T)}, "Col_Ops": col,
        "mean_operator": a, "tlist": np.linspace(0.0, periods * period, steps + 1),
        "var_arrays": {EPS: np.linspace(-p["eps_max"], p["eps_max"], n_eps),
                       AMP: np.linspace(0.0, p["amp_max"], n_amp)},
        "params": p,
    }


def serving_stream(count=1000, rate=20.0, seed=7, sides=(32, 128, 512, 1024)):
    """Seeded Poisson stream of sweep-request descriptions: model (two-level LZS or an N-level
    chain, N in {2, 4, 8}), square grid side, amplitude range, solver (rk4 / tsit5a) and
    precision (fp32 / fp64).  Returns a list of dicts with the arrival time in seconds."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for i in range(count):
        style = rng.choice(["lzs", "chain"])
        n = 2 if style == "lzs" else int(rng.choice([2, 4, 8]))
        out.append({"arrival": float(t[i]), "style": str(style), "n": n,
                    "side": int(rng.choice(sides)), "amp_max": float(rng.uniform(5.0, 30.0)),
                    "solver": str(rng.choice(["rk4", "tsit5a"])),
                    "fp64": bool(rng.integers(0, 2))})
    return out
