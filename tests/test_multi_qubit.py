"""The reference's multi-qubit chain benchmark model on the CPU: its constants against the values
the reference's published result file records, the oracle's numeric model against the symbolic
workload, and a closed form of the uncoupled chain."""
import numpy as np
import pytest
import sympy as sp

from gqis import workloads
from oracle import integrate, lindblad, models


def test_constants_match_the_published_header():
    # Benchmarks/results/Benchmark_04_hilbert_space_scaling_calibrated.csv:18 (multi_qubit
    # physics): Delta 1, w/Delta 1.14, J/Delta 0.1, gap_step 0.05, gamma1_per_T 0.05,
    # gamma_phi_per_T 0.04, eps_max/w = A_max/w = 10; ode_variables 63 / 255 (:94-95)
    for nq, states in ((3, 63), (4, 255)):
        w = workloads.multi_qubit(nq, n_eps=3, n_amp=3)
        assert w["H"].shape == (2 ** nq, 2 ** nq) and 4 ** nq - 1 == states
        eps, amp = w["var_arrays"].values()
        assert eps[0] == pytest.approx(-11.4) and eps[-1] == pytest.approx(11.4)
        assert amp[0] == 0.0 and amp[-1] == pytest.approx(11.4)
        assert w["numeric"]["w"] == pytest.approx(1.14)
        h0 = w["numeric"]["h0"]
        # qubit k's gap 1 + 0.05 k: the sx_k coefficient is half of it; ZZ neighbours 0.1
        assert h0[0, 1] == pytest.approx(0.5 * (1 + 0.05 * (nq - 1)))   # last qubit's sx
        assert h0[0, 2 ** (nq - 1)] == pytest.approx(0.5)               # qubit 0's sx
        assert h0[0, 0] == pytest.approx(0.1 * (nq - 1))                # all sz_i sz_i+1 = +1
        period = 2 * np.pi / 1.14
        g1 = 0.05 / period
        g2 = g1 / 2 + 0.04 / period
        col = w["numeric"]["col"]
        assert len(col) == 2 * nq
        assert np.max(np.abs(col[0])) == pytest.approx(np.sqrt(g1))
        assert np.max(np.abs(col[1])) == pytest.approx(np.sqrt(g2))


@pytest.mark.parametrize("nq", [2, 3])
def test_oracle_generator_equals_the_symbolic_model(nq):
    # the workload's SymPy matrices and the oracle's NumPy terms are written independently:
    # the generator of H(Drive = eps + A sin(w t)) and the workload's collapse operators must
    # equal the oracle's sum of scaled generators at every (eps, A, t)
    w = workloads.multi_qubit(nq, n_eps=2, n_amp=2)
    (drive_symbol,) = w["Drive"].keys()
    col = [np.array(sp.Matrix(c).evalf(), dtype=complex) for c in w["Col_Ops"]]
    rng = np.random.default_rng(nq)
    for eps, amp, t in ((-3.0, 0.0, 0.0), (0.7, 2.5, 1.9), (5.0, 11.0, 33.3)):
        drive = eps + amp * np.sin(1.14 * t)
        h = np.array(w["H"].subs(drive_symbol, drive).evalf(), dtype=complex)
        m_sym, b_sym = lindblad.affine_generator(h, col)
        m = models.multi_qubit(nq, [eps], [amp])
        y = rng.standard_normal((m.dim, 1))
        np.testing.assert_allclose(m.f(t, y)[:, 0], m_sym @ y[:, 0] + b_sym, atol=1e-12)
        o0, o = lindblad.observable_form(np.array(w["mean_operator"].evalf(), dtype=complex))
        assert o0 == m.o0 and np.array_equal(o, m.o)


def test_uncoupled_qubit_zero_rabi_oscillation():
    # with the ZZ coupling, the drive and the dissipation off, qubit 0 (gap delta = 1) starts in
    # |0> and its excited population is sin^2(t / 2) whatever the other qubits do: pins the sx
    # coefficient delta / 2, the qubit ordering of the observable and the ground state
    m = models.multi_qubit(3, [0.0], [0.0], gamma1_per_t=0.0, gamma_phi_per_t=0.0,
                           j_over_delta=0.0)
    _, trace = integrate.rk4(m.f, m.ground_state(), 0.01, 600, obs=m.obs, mode="final",
                             n_trace=600)
    t = 0.01 * np.arange(1, 601)
    assert np.max(np.abs(trace[0].real - np.sin(t / 2) ** 2)) < 1e-9
