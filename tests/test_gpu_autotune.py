"""block_size="auto" on the GPU: the measured choice gives the same results as a fixed size,
and the choice is made once per kernel."""
import numpy as np
import pytest

from gqis import autotune, mesolve_2D, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["rk4", "tsit5a"])
def test_auto_block_matches_fixed_block(gpu, solver):
    autotune.clear_memory()
    w = workloads.chain(8, n_eps=256, n_amp=64, n_t=401, t_end=20.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    fixed = mesolve_2D(*args, var_arrays=w["var_arrays"], fp64=True, solver=solver,
                       liouville="lanes", block_size=256, lanes_per_point=1)
    got, info = mesolve_2D(*args, var_arrays=w["var_arrays"], fp64=True, solver=solver,
                           liouville="lanes", block_size="auto", return_timing_info=True)
    print(solver, "chose", info["lanes_per_point"], info["block_size"], info.get("autotune"),
          flush=True)
    # candidates: (lanes per point, work-group size); 63 states also try the grouped kernels
    assert (info["lanes_per_point"], info["block_size"]) in info["autotune"]
    assert {c[0] for c in info["autotune"]} == {1, 2, 4, 8, 16}
    scale = float(np.max(np.abs(fixed)))
    if info["lanes_per_point"] == 1:
        np.testing.assert_array_equal(got, fixed)     # one lane per point: same arithmetic
    else:
        assert float(np.max(np.abs(got - fixed))) <= 1e-12 * scale
    _, again = mesolve_2D(*args, var_arrays=w["var_arrays"], fp64=True, solver=solver,
                          liouville="lanes", block_size="auto", return_timing_info=True)
    assert (again["lanes_per_point"], again["block_size"]) == (info["lanes_per_point"],
                                                               info["block_size"])
    assert "autotune" not in again


def test_auto_block_on_the_matrix_core_path(gpu):
    autotune.clear_memory()
    w = workloads.random_n_level(16, n_eps=128, n_amp=128, n_t=201, t_end=2.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    fixed = mesolve_2D(*args, var_arrays=w["var_arrays"], liouville="mfma", block_size=256)
    got, info = mesolve_2D(*args, var_arrays=w["var_arrays"], liouville="mfma",
                           block_size="auto", return_timing_info=True)
    print("mfma chose", info["block_size"], info.get("autotune"), flush=True)
    assert info["block_size"] in autotune.CANDIDATES
    np.testing.assert_array_equal(got, fixed)
