"""The reference's installation self-test (``gqis/check_environment.py:271-355``) on the GPU: the
stages with a small problem whose answer is exact -- a Rabi rotation, relaxation, the ODE API and
the time-step calibration -- with the reference's own thresholds (the symbolic and code
generation stages run on the CPU, ``tests/test_frontend.py``)."""
import numpy as np
import pytest
import sympy as sp

from gqis import mesolve_2D, odesolve_2D

pytestmark = pytest.mark.gpu

SX = sp.Matrix([[0, 1], [1, 0]])
EXCITED = sp.diag(0, 1)
A, DRIVE, GAMMA, DUMMY = sp.symbols("A Drive gamma dummy", real=True)
ONE_POINT = {DUMMY: np.zeros(1, dtype=np.float32)}
RABI = (DRIVE / 2 * SX, A, [], EXCITED)      # Rabi rotation from level 0


def test_rabi_rotation(gpu):
    # check_environment.py:303-310: P(excited) at t = pi is sin^2(A pi / 2), within 1e-4 (FP32)
    amplitudes = np.array([1.0, 2.0], dtype=np.float32)
    tlist = np.linspace(0.0, np.pi, 257, dtype=np.float32)
    result, info = mesolve_2D(*RABI, tlist, var_arrays={A: amplitudes, **ONE_POINT},
                              output_mode="final", return_timing_info=True)
    exact = np.sin(amplitudes * np.pi / 2) ** 2
    assert np.max(np.abs(np.real(result[:, 0]) - exact)) < 1e-4
    assert info["gpu_kernel_s"] > 0.0


def test_relaxation(gpu):
    # check_environment.py:312-319: pumping sqrt(gamma)|1><0| gives P = 1 - exp(-gamma t)
    rates = np.array([0.5, 2.0], dtype=np.float32)
    tlist = np.linspace(0.0, 1.0, 129, dtype=np.float32)
    pump = sp.sqrt(GAMMA) * sp.Matrix([[0, 0], [1, 0]])
    result = mesolve_2D(sp.zeros(2), 0, [pump], EXCITED, tlist,
                        var_arrays={GAMMA: rates, **ONE_POINT}, output_mode="final")
    assert np.max(np.abs(np.real(result[:, 0]) - (1 - np.exp(-rates)))) < 1e-5


def test_harmonic_oscillator_ode(gpu):
    # check_environment.py:321-327: x'' = -w^2 x from x = 1 gives x(2) = cos(2 w)
    x, v, w = sp.symbols("x v w", real=True)
    frequencies = np.array([1.0, 2.0], dtype=np.float32)
    result = odesolve_2D([v, -w ** 2 * x], [x, v], [1, 0], np.linspace(0, 2, 257),
                         var_arrays={w: frequencies})
    assert np.max(np.abs(result[:, 0, 0] - np.cos(2 * frequencies))) < 1e-5


def test_calibration_refines_a_coarse_grid(gpu):
    # check_environment.py:329-337: 4 steps over [0, pi] are too few; the calibration takes
    # more and the Rabi flip reaches P = 1 within 1e-2
    tlist = np.linspace(0.0, np.pi, 5, dtype=np.float32)
    result, info = mesolve_2D(*RABI, tlist, var_arrays={A: np.array([1.0], np.float32),
                                                        **ONE_POINT},
                              output_mode="final", calibrate_timesteps=True,
                              return_timing_info=True)
    steps = info["timestep_calibration"]["steps"]
    assert steps > 4
    assert abs(float(np.real(result[0, 0])) - 1.0) < 1e-2
