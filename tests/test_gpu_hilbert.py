"""Matrix-core (Hilbert-space) RK4 path against the FP64 oracle and the per-lane kernels."""
import numpy as np
import pytest

from gqis import mesolve_2D, workloads
from oracle import integrate, models

pytestmark = pytest.mark.gpu
SEED = 2024


@pytest.mark.parametrize("n,steps", [(8, 400), (16, 200), (32, 100)])
def test_random_n_level_mean_and_state_match_oracle(gpu, n, steps):
    w = workloads.random_n_level(n, seed=SEED, n_eps=3, n_amp=2, n_t=steps + 1,
                                 t_end=0.01 * steps)
    eps_v, amp_v = w["var_arrays"].values()
    got, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                           var_arrays=w["var_arrays"], liouville="mfma", return_timing_info=True)
    rho = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], liouville="mfma", output_mode="final_rho")
    m = models.random_n_level(n, SEED, eps_v, amp_v)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, obs=m.obs)
    fin, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, mode="final_rho")
    assert got.shape == (3, 2) and rho.shape == (3, 2, n * n - 1)
    # FP32 MFMA products and FP32 state; tolerance relative to the map's scale
    assert np.max(np.abs(got.ravel() - ref)) <= 2e-5 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(rho.reshape(6, -1) - fin)) <= 2e-5
    assert info["solver"] == "rk4"


def test_chain_mfma_matches_per_lane_kernel(gpu):
    w = workloads.chain(8, n_eps=8, n_amp=6, n_t=801, t_end=40.0)
    kw = dict(var_arrays=w["var_arrays"])
    mfma = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                      liouville="mfma", **kw)
    lanes64 = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                         liouville="lanes", fp64=True, **kw)
    ok = np.abs(lanes64) <= 1.0
    assert ok.sum() >= 40
    assert np.max(np.abs(mfma - lanes64)[ok]) <= 5e-5


def test_mfma_continuation_from_final_state(gpu):
    w = workloads.random_n_level(16, seed=SEED, n_eps=2, n_amp=2, n_t=101, t_end=1.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"])
    first = mesolve_2D(*args, w["tlist"], var_arrays=w["var_arrays"], liouville="mfma",
                       output_mode="final_rho")
    second = mesolve_2D(*args, w["tlist"], var_arrays=w["var_arrays"], liouville="mfma",
                        output_mode="final_rho", rho0_values=first, t_in=1.0)
    whole = mesolve_2D(*args, np.linspace(0.0, 2.0, 201), var_arrays=w["var_arrays"],
                       liouville="mfma", output_mode="final_rho")
    np.testing.assert_allclose(second, whole, atol=2e-5)


@pytest.mark.parametrize("n,steps", [(8, 200), (16, 100), (32, 40)])
def test_fp64_hilbert_path_parity(gpu, n, steps):
    w = workloads.random_n_level(n, seed=SEED, n_eps=2, n_amp=2, n_t=steps + 1,
                                 t_end=0.01 * steps)
    eps_v, amp_v = w["var_arrays"].values()
    got = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], liouville="mfma", fp64=True)
    m = models.random_n_level(n, SEED, eps_v, amp_v)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, obs=m.obs)
    assert got.dtype == np.complex128
    assert np.max(np.abs(got.ravel() - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref)))
