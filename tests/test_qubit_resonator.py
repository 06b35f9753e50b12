"""The reference's qubit-resonator benchmark model on the CPU: its constants against the values
the reference's published result files record, the oracle's numeric operators against the
symbolic workload, and physical invariants of the coupling."""
import numpy as np
import pytest
import sympy as sp

from gqis import workloads
from oracle import integrate, lindblad, models


def test_constants_match_the_published_headers():
    # Benchmarks/results/Benchmark_02_full_benchmark.csv:33-40 (wd500): Delta 5.4165, w 0.5,
    # gamma1 0.002, gamma2 0.001, eps in +-11.3701, A up to 13.9157953;
    # Benchmarks/results/Benchmark_04_hilbert_space_scaling_calibrated.csv:15: g1 / wq1
    p = workloads.qubit_resonator_parameters()
    assert p["delta"] == pytest.approx(5.4165, abs=1e-12)
    assert p["w"] == pytest.approx(0.5, abs=1e-9)
    assert p["gamma1"] == 0.002 and p["gamma2"] == 0.001 and p["kappa"] == 0.005
    # the headers print the FP32 grid ends (the benchmark passes float32 axes)
    assert float(np.float32(p["eps_max"])) == pytest.approx(11.3701, abs=5e-7)
    assert float(np.float32(p["amp_max"])) == pytest.approx(13.9157953, abs=5e-8)
    assert p["g1"] / 5.4165 == pytest.approx(0.03279650630894309, rel=1e-14)
    assert p["ap"] / 5.4165 == pytest.approx(2e-5, rel=1e-14)
    # the published 2048 x 2048 RK4 row: 40 periods of 146 steps = 5840 steps
    w = workloads.qubit_resonator(2, n_eps=2, n_amp=2)
    assert len(w["tlist"]) == 5841
    assert w["tlist"][-1] == pytest.approx(40 * 2 * np.pi / p["w"])


@pytest.mark.parametrize("n_photon", [2, 3])
def test_oracle_operators_equal_the_symbolic_model(n_photon):
    # the workload's SymPy matrices and the oracle's NumPy terms are written independently; at
    # t = 0 with A = 0 the drive is eps, and the oracle's sum of g_k H_k must equal H(eps)
    p = workloads.qubit_resonator_parameters()
    w = workloads.qubit_resonator(n_photon, n_eps=2, n_amp=2)
    (drive_symbol,) = w["Drive"].keys()
    for eps in (-3.0, 0.7):
        h_sym = np.array(w["H"].subs(drive_symbol, eps).evalf(), dtype=complex)
        m = models.qubit_resonator(n_photon, [eps], [0.0], p)
        h_num = sum(lb_g(g, 0.0, eps) * h for g, h in _terms(n_photon, p))
        assert np.max(np.abs(h_sym - h_num)) < 1e-13
        assert m.dim == (2 * n_photon) ** 2 - 1
    for got, want in zip(w["Col_Ops"], _terms(n_photon, p, collapse=True)):
        assert np.max(np.abs(np.array(sp.Matrix(got).evalf(), dtype=complex) - want)) < 1e-15
    sm, sz, a = models.qubit_resonator_operators(n_photon)
    assert np.array_equal(np.array(w["mean_operator"].evalf(), dtype=complex), a)


def lb_g(g, t, eps):
    return g(t, np.array([eps]), np.array([0.0]))[0] if callable(g) else g


def _terms(n_photon, p, collapse=False):
    # the oracle's terms, read back from its AffineSweep construction inputs
    sm, sz, a = models.qubit_resonator_operators(n_photon)
    if collapse:
        return [np.sqrt(p["gamma1"]) * sm, np.sqrt(p["gamma2"]) * sz, np.sqrt(p["kappa"]) * a]

    def energy(t, x, y):
        return np.sqrt(p["delta"] ** 2 + (x + y * np.cos(p["w"] * t)) ** 2)
    jc = p["g1"] * p["delta"] * (sm.conj().T @ a + sm @ a.conj().T)
    return [(1.0, p["ap"] * (a + a.conj().T)), (lambda t, x, y: 1.0 / energy(t, x, y), jc),
            (lambda t, x, y: 0.5 * (energy(t, x, y) - p["wr2"]), sz)]


@pytest.mark.parametrize("n_photon", [2, 4])
def test_coupling_conserves_excitations(n_photon):
    # s+ a + s- a^+ moves one quantum between qubit and cavity: it commutes with the excitation
    # number s+ s- + a^+ a (a transposed s- or a would not); sz commutes with it too
    sm, sz, a = models.qubit_resonator_operators(n_photon)
    n_exc = sm.conj().T @ sm + a.conj().T @ a
    jc = sm.conj().T @ a + sm @ a.conj().T
    assert np.max(np.abs(jc @ n_exc - n_exc @ jc)) < 1e-14
    assert np.max(np.abs(sz @ n_exc - n_exc @ sz)) < 1e-14
    assert np.max(np.abs(a.conj().T @ a - np.kron(np.eye(2), np.diag(np.arange(n_photon))))) < 1e-14


def test_cavity_relaxation_closed_form():
    # with the probe, the coupling and the qubit terms off, a coherent cavity amplitude decays as
    # <a>(t) = <a>(0) exp(-kappa t / 2): the oracle's generator and RK4 reproduce it
    p = dict(workloads.qubit_resonator_parameters(), ap=0.0, g1=0.0, gamma1=0.0, gamma2=0.0,
             kappa=0.3)
    m = models.qubit_resonator(3, [0.0], [0.0], p)
    # start in |up> (x) (|0> + |1>) / sqrt(2): <a> = 1/2
    psi = np.zeros(6, complex)
    psi[0] = psi[1] = 1 / np.sqrt(2)
    y0 = lindblad.rho_to_reduced(np.outer(psi, psi.conj()))[:, None]
    # sz (E - wr2) / 2 rotates the qubit only: the cavity coherence is untouched by it
    _, trace = integrate.rk4(m.f, y0, 0.01, 400, obs=m.obs, mode="final", n_trace=400)
    t = 0.01 * np.arange(1, 401)
    assert np.max(np.abs(np.abs(trace[0]) - 0.5 * np.exp(-0.15 * t))) < 1e-9
