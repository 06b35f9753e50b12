"""Pins the FP64 oracle: closed forms, the reference's known answers, convergence order,
SciPy and matrix exponentials.  A plausible slip (wrong stage time, dropped term, transposed
generator, wrong mean window) fails at least one of these."""
import numpy as np
import pytest
import scipy.linalg
from scipy.integrate import solve_ivp

from oracle import integrate, lindblad as lb, models

SX = np.array([[0, 1], [1, 0]], complex)
SZ = np.array([[1, 0], [0, -1]], complex)
SM = np.array([[0, 1], [0, 0]], complex)


def test_reduced_generator_is_the_bloch_equations():
    # reference tests/test_package.py:63-85 (closed-form Bloch equations)
    rng = np.random.default_rng(1)
    for _ in range(5):
        delta, drive, g1, g2 = rng.uniform(0.1, 2.0, 4)
        m, b = lb.affine_generator(delta * SX / 2 + drive * SZ / 2,
                                   [np.sqrt(g1) * SM, np.sqrt(g2) * SZ])
        cd = g1 / 2 + 2 * g2
        # y = (rho00, Re rho01, Im rho01)
        m_ref = np.array([[-g1, 0, -delta], [0, -cd, drive], [delta, -drive, -cd]])
        b_ref = np.array([g1, 0, -delta / 2])
        np.testing.assert_allclose(m, m_ref, atol=1e-14)
        np.testing.assert_allclose(b, b_ref, atol=1e-14)
        o0, o = lb.observable_form(np.diag([0.0, 1.0]))
        assert o0 == 1 and np.allclose(o, [-1, 0, 0])


def test_reduced_layout_round_trip():
    rng = np.random.default_rng(2)
    for n in (2, 3, 4):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        rho = a @ a.conj().T
        rho /= np.trace(rho)
        y = lb.rho_to_reduced(rho)
        assert y.size == n * n - 1
        np.testing.assert_allclose(lb.reduced_to_rho(y), rho, atol=1e-14)


def test_generator_matches_full_lindblad_map():
    rng = np.random.default_rng(3)
    n = 3
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = (a + a.conj().T) / 2
    ops = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(2)]
    m, b = lb.affine_generator(h, ops)
    r = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    rho = r @ r.conj().T
    rho /= np.trace(rho)
    y = lb.rho_to_reduced(rho)
    np.testing.assert_allclose(m @ y + b, lb.coords(lb.lindblad(h, ops, rho)), atol=1e-12)


def _const_model(n=3, seed=4):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = (a + a.conj().T) / 2
    op = np.zeros((n, n), complex)
    op[0, 1] = 0.3
    m, b = lb.affine_generator(h, [op])
    return m, b


def test_rk4_fourth_order_against_matrix_exponential():
    m, b = _const_model()
    d = m.shape[0]
    aug = np.zeros((d + 1, d + 1))
    aug[:d, :d], aug[:d, d] = m, b
    y0 = np.zeros(d)
    y0[0] = 1.0
    exact = (scipy.linalg.expm(aug * 2.0) @ np.append(y0, 1.0))[:d]
    errs = []
    for steps in (40, 80, 160):
        res, _ = integrate.rk4(lambda t, y: m @ y + b[:, None], y0[:, None], 2.0 / steps, steps,
                               mode="final_rho")
        errs.append(np.max(np.abs(res[0] - exact)))
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(14.0 < r < 18.0 for r in ratios), ratios


def test_rk4_stage_times_commuting_hamiltonian():
    # reference tests/test_gpu_solver.py:34-54: H = (1 + 0.4 sin t) sx/2, FP64, 80 steps
    final_time = 1.7
    m, b = lb.affine_generator(SX / 2, [])

    def f(t, y):
        return (1 + 0.4 * np.sin(t)) * (m @ y + b[:, None])
    y0 = np.array([[1.0], [0.0], [0.0]])
    res, _ = integrate.rk4(f, y0, final_time / 80, 80, mode="final_rho")
    phase = final_time + 0.4 * (1.0 - np.cos(final_time))
    assert abs(res[0, 0] - np.cos(0.5 * phase) ** 2) < 2e-7


def test_rk4_time_offset_and_right_endpoint_mean():
    # x' = t is integrated exactly by RK4: x(t_k) = ((t_in + t_k)^2 - t_in^2)/2;
    # the mean is the average over the step endpoints after the warmup.
    grid = np.linspace(0.0, 1.0, 65)
    dt = grid[1]

    def f(t, y):
        return np.full_like(y, t)
    res, _ = integrate.rk4(f, np.zeros((1, 1)), dt, 64, t_in=2.0, mode="final_rho")
    assert abs(res[0, 0] - 2.5) < 1e-12          # reference tests/test_ode_api.py:350-355
    obs = lambda t, y: y[0] + 0j  # noqa: E731
    mean, _ = integrate.rk4(f, np.zeros((1, 1)), dt, 64, warmup_steps=16, obs=obs)
    xs = grid[1:] ** 2 / 2
    assert abs(mean[0].real - xs[16:].mean()) < 1e-12
    full, _ = integrate.rk4(f, np.zeros((1, 1)), dt, 64, obs=obs)
    assert abs(full[0].real - xs.mean()) < 1e-12   # reference tests/test_ode_api.py:366-368


def test_rk4_trace_sampling_and_state_mean():
    f = lambda t, y: -y  # noqa: E731
    res, tr = integrate.rk4(f, np.ones((1, 2)), 1 / 128, 128, mode="mean_state",
                            trace_stride=16, n_trace=8, trace_kind="state")
    t_samples = (np.arange(8) * 16 + 1) / 128
    np.testing.assert_allclose(tr[0, :, 0], np.exp(-t_samples), atol=1e-9)
    np.testing.assert_allclose(res[0, 0], np.exp(-np.arange(1, 129) / 128).mean(), atol=1e-9)


def test_installation_known_answers():
    # reference gqis/check_environment.py:302-327
    # Rabi: H = (A/2) sx, P(excited at pi) = sin^2(A pi / 2)
    for amp in (1.0, 2.0):
        m = models.AffineSweep([(models.one, amp / 2 * SX, [])], np.diag([0.0, 1.0]), [0.0], [0.0])
        res, _ = integrate.rk4(m.f, m.ground_state(), np.pi / 256, 256, obs=m.obs, mode="final")
        assert abs(res[0].real - np.sin(amp * np.pi / 2) ** 2) < 1e-4
    # dissipation: pump sqrt(gamma)|1><0|, P = 1 - exp(-gamma t)
    for gamma in (0.5, 2.0):
        pump = np.sqrt(gamma) * np.array([[0, 0], [1, 0]], complex)
        m = models.AffineSweep([(models.one, np.zeros((2, 2)), [pump])], np.diag([0.0, 1.0]),
                               [0.0], [0.0])
        res, _ = integrate.rk4(m.f, m.ground_state(), 1 / 128, 128, obs=m.obs, mode="final")
        assert abs(res[0].real - (1 - np.exp(-gamma))) < 1e-5


def test_tsit5a_closed_forms():
    # reference tests/test_gpu_solver.py:269-310
    m, b = lb.affine_generator(SX / 2, [])

    def f(t, y, idx):
        return (1 + 0.4 * np.sin(t)) * (m @ y + b[:, None])
    res, _, acc, rej = integrate.tsit5a(f, np.array([[1.0], [0.0], [0.0]]), 1.7 / 4, 4,
                                        rtol=1e-12, atol=1e-12, max_substeps=10000)
    phase = 1.7 + 0.4 * (1.0 - np.cos(1.7))
    assert abs(res[0, 0] - np.cos(0.5 * phase) ** 2) < 2e-7

    def g(t, y, idx):
        return m @ y + b[:, None]
    obs = lambda t, y, idx: y[0] + 0j  # noqa: E731
    val, _, acc, _ = integrate.tsit5a(g, np.array([[1.0], [0.0], [0.0]]), 2.0, 5, rtol=1e-10,
                                      atol=1e-10, max_substeps=10000, warmup_steps=1, obs=obs,
                                      mode="mean")
    assert abs(val[0].real - (0.5 + (np.sin(10.0) - np.sin(2.0)) / 16.0)) < 5e-5
    assert acc[0] > 5


@pytest.mark.parametrize("method", ["tsit5a", "dop853a"])
def test_adaptive_oracles_against_scipy(method):
    d = models.Duffing(np.array([0.2, 0.45]), np.array([-1.0, 0.7]))
    fn = integrate.tsit5a if method == "tsit5a" else integrate.dop853a
    res, _, acc, rej = fn(d.f, d.y0(), 0.1, 100, rtol=1e-11, atol=1e-13, max_substeps=5000)
    for p in range(2):
        sol = solve_ivp(lambda t, y: d.f(t, y[:, None], np.array([p]))[:, 0], (0, 10),
                        d.y0()[:, p], method="DOP853", rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(res[p], sol.y[:, -1], atol=5e-8)
    assert np.all(acc > 100)


def test_adaptive_substep_limit_gives_nan():
    d = models.Duffing(np.array([0.3]), np.array([1.0]))
    res, _, acc, rej = integrate.tsit5a(d.f, d.y0(), 0.5, 4, rtol=1e-14, atol=1e-14,
                                        max_substeps=2)
    assert np.all(np.isnan(res))


def test_c_oracle_matches_numpy_oracle():
    from oracle import cbridge
    m = models.four_level_chain(np.linspace(-10, 10, 4), np.linspace(0, 20, 3))
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.05, 400, obs=m.obs, warmup_steps=100)
    got = cbridge.rk4_affine(m, 0.05, 400, warmup=100, threads=2)
    np.testing.assert_allclose(got, ref, atol=1e-13)
    fin, _ = integrate.rk4(m.f, m.ground_state(), 0.05, 400, mode="final_rho")
    np.testing.assert_allclose(cbridge.rk4_affine(m, 0.05, 400, mode="final_rho"), fin, atol=1e-13)


def test_product_dop853_tableau_is_hairers():
    from gqis import tableaux
    c, a, b, e5, e3 = integrate.dop853_tableau()
    t = tableaux.DOP853
    np.testing.assert_array_equal(np.array(t["c"]), c)
    np.testing.assert_array_equal(np.array(t["b"]), b)
    np.testing.assert_array_equal(np.array(t["e5"]), e5)
    np.testing.assert_allclose(np.array(t["e3"]), e3, atol=1e-17)
    for s in range(12):
        np.testing.assert_array_equal(np.array(t["a"][s]), a[s, :len(t["a"][s])])
    np.testing.assert_array_equal(np.array(integrate.TSIT5_B), np.array(tableaux.TSIT5["b"]))


@pytest.mark.parametrize("name,order", [("rk4", 4), ("lserk4", 4), ("dp5", 5), ("tsit5", 5),
                                        ("alshina6", 6), ("ab5", 5), ("anas5", 5)])
def test_fixed_step_oracles_have_their_order(name, order):
    """Convergence order on a time-dependent dissipative model pins each tableau and the
    stage-time semantics (a wrong coefficient or stage time drops the order)."""
    m, b = _const_model()
    y0 = np.zeros((m.shape[0], 1))
    y0[0] = 1.0

    def f(t, y):
        return (1 + 0.3 * np.sin(t)) * (m @ y + b[:, None])
    sol = solve_ivp(lambda t, y: (1 + 0.3 * np.sin(t)) * (m @ y + b), (0, 2), y0[:, 0],
                    method="DOP853", rtol=1e-13, atol=1e-15)
    errs = []
    for steps in (20, 40, 80):
        r, _ = integrate.fixed_step(name, f, y0, 2.0 / steps, steps, mode="final_rho")
        errs.append(np.max(np.abs(r[0] - sol.y[:, -1])))
    observed = np.log2(errs[1] / errs[2])
    assert order - 0.3 < observed < order + 1.0, (name, errs)


def test_dop853_fixed_is_eighth_order():
    m, b = _const_model()
    f = lambda t, y: (m @ y + b[:, None])  # noqa: E731
    d = m.shape[0]
    aug = np.zeros((d + 1, d + 1))
    aug[:d, :d], aug[:d, d] = m, b
    y0 = np.zeros(d)
    y0[0] = 1.0
    exact = (scipy.linalg.expm(aug * 4.0) @ np.append(y0, 1.0))[:d]
    errs = [np.max(np.abs(integrate.fixed_step("dop853", f, y0[:, None], 4.0 / n, n,
                                               mode="final_rho")[0][0] - exact)) for n in (4, 8)]
    assert np.log2(errs[0] / errs[1]) > 7.0


def test_python_cpu_path_matches_c_oracle():
    # the restated per-point Python path (timed CPU comparison) gives the C oracle's map,
    # also through its spawn pool
    from oracle import cbridge, python_cpu
    eps, amp = np.array([-3.0, 7.5]), np.array([0.0, 4.0, 12.0])
    ref = np.asarray(cbridge.rk4_affine(models.four_level_chain(eps, amp), 0.05, 200)).real
    rhs, obs, size = python_cpu.prepare("four_level")
    one = python_cpu.solve_point(rhs, obs, size, eps[1], amp[2], 0.05, 200)
    assert abs(one - ref.reshape(2, 3)[1, 2]) < 1e-13
    out, compute_s, workers = python_cpu.sweep("four_level", eps, amp, 0.05, 200, workers=2)
    assert workers == 2 and compute_s >= 0.0
    np.testing.assert_allclose(out, ref.reshape(2, 3), rtol=0, atol=1e-13)


def _anas5_rotation_phase(omega, dt, steps, monkeypatch=None, unfitted=False):
    """Phase error after ``steps`` Anas5 steps on the oscillator x' = -w y, y' = w x (that is,
    z' = i w z for z = x + i y) started at z = 1."""
    if unfitted:
        monkeypatch.setattr(integrate, "anas5_parameter", lambda w, h: -800.0 / 1071.0)
    rot = np.array([[0.0, -omega], [omega, 0.0]])
    y, _ = integrate.fixed_step("anas5", lambda t, y: rot @ y, np.array([[1.0], [0.0]]), dt,
                                steps, mode="final_rho", frequency=omega)
    z = complex(y[0, 0], y[0, 1])
    return np.angle(z * np.exp(-1j * omega * dt * steps)), abs(z)


@pytest.mark.parametrize("omega,dt", [(1.0, 0.3), (2.0, 0.3), (1.0, 1.0), (5.0, 0.1)])
def test_anas5_fitted_coefficient_has_zero_phase_lag(omega, dt, monkeypatch):
    """The fitted Anas5(w) coefficient (reference gqis/cuda_solvers.py:971-990) is the one that
    removes the method's phase lag at v = w*dt: 50 steps on z' = i w z end at phase 50 v to
    rounding, while the unfitted -800/1071 lags by a v^7-sized amount per step."""
    lag, amp = _anas5_rotation_phase(omega, dt, 50)
    assert abs(lag) < 1e-12 and abs(amp - 1) < 0.05   # dissipation error is not fitted
    lag0, _ = _anas5_rotation_phase(omega, dt, 50, monkeypatch, unfitted=True)
    v = omega * dt
    assert abs(lag0) > 1e-5 * v ** 7 * 50 and abs(lag0) > 1e3 * abs(lag)


def test_product_anas5_coefficient_equals_the_closed_form():
    """The product derives the coefficient from the tableau's stability function; the oracle
    restates the published closed form.  They agree over v = 1e-5 .. 1.4."""
    from gqis.solvers import anas5_coefficient
    for omega, dt in [(1.0, 1e-5), (1.0, 0.05), (0.1, 0.01), (1.0, 0.3), (2.0, 0.5), (3.0, 0.4),
                      (7.0, 0.2)]:
        p = anas5_coefficient(omega, dt)[0]
        assert p == pytest.approx(integrate.anas5_parameter(omega, dt), rel=1e-14, abs=0)
