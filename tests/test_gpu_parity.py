"""GPU parity of the sweep kernels against the FP64 oracle (same seeded inputs)."""
import numpy as np
import pytest

from gqis import mesolve_2D, workloads
from oracle import integrate, models

pytestmark = pytest.mark.gpu


def _rel(got, ref):
    """max |a-b| / max|b| (normalised by the map's scale, not per point)."""
    return float(np.max(np.abs(got - ref)) / max(1e-300, float(np.max(np.abs(ref)))))


@pytest.mark.parametrize("fp64", [True, False])
def test_four_level_mean_matches_oracle(gpu, fp64):
    w = workloads.four_level(n_eps=16, n_amp=12, n_t=2001, t_end=100.0)
    got = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], fp64=fp64)
    eps_v, amp_v = w["var_arrays"].values()
    m = models.four_level_chain(eps_v, amp_v)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.05, 2000, obs=m.obs)
    ref = ref.reshape(16, 12)
    assert got.shape == (16, 12)
    assert got.dtype == (np.complex128 if fp64 else np.complex64)
    if fp64:
        assert _rel(got, ref) <= 1e-10
    else:
        # RK4 at dt = 0.05 is unstable where |2 eps + 2 + A cos t| dt exceeds the stability
        # limit (the eps = 15, A = 30 corner): there the state grows beyond a density matrix and
        # any rounding difference is amplified.  Elsewhere: FP32 RHS and state with the FP64
        # time table and mean accumulator stay within 5e-6 of the FP64 oracle.
        final, _ = integrate.rk4(m.f, m.ground_state(), 0.05, 2000, obs=m.obs, mode="final_rho")
        stable = (np.max(np.abs(final), axis=1) <= 1.0).reshape(16, 12)
        assert stable.sum() >= 180
        assert np.max(np.abs(got - ref)[stable]) <= 5e-6


def test_two_level_final_rho_and_trace(gpu):
    w = workloads.lzs_two_level(n_eps=8, n_amp=8, n_t=501, t_end=50.0)
    res, trace, times = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                                   w["tlist"], var_arrays=w["var_arrays"], fp64=True,
                                   output_mode="final_rho", return_time_trace=True,
                                   time_trace_every=7)
    eps_v, amp_v = w["var_arrays"].values()
    m = models.lzs_two_level(eps_v, amp_v)
    n_trace = int(np.ceil(500 / 7))
    ref, ref_tr = integrate.rk4(m.f, m.ground_state(), 0.1, 500, obs=m.obs, mode="final_rho",
                                trace_stride=7, n_trace=n_trace)
    assert res.shape == (8, 8, 3) and trace.shape == (8, 8, n_trace)
    assert _rel(res.reshape(-1, 3), ref) <= 1e-10
    assert _rel(trace.reshape(64, -1), ref_tr) <= 1e-10
    np.testing.assert_allclose(times, (np.arange(n_trace) * 7 + 1) * 0.1)
