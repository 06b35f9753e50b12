"""gq_launch_sweep_noalloc (the entry point that allocates nothing) against the ordinary launch
path (gq_launch_sweep, through mesolve_2D): bitwise equal results on two calls in a row, the
second with a smaller input, for a fixed-step sweep whose time table lives in the library's
scratch array and for a persistent adaptive sweep whose work cursor does.

The caller-side table and cursor of the prepared plan are filled with values that would give
wrong results (NaN table rows, a cursor past the last point) and the status word with set bits:
the results come out right only if the entry point uses its own scratch and clears it."""
import numpy as np
import pytest

from gqis import mesolve_2D, prepare_sweep, runtime, workloads
from gqis.device import collect_outputs

pytestmark = pytest.mark.gpu


def _poison(plan):
    b = plan["buffers"]
    if b["table"] is not None:
        b["table"].fill_(float("nan"))
    b["cursor"].fill_(1 << 30)
    b["status"].fill_(7)


def _run_noalloc(prep):
    import torch
    plan = prep.plan
    b = plan["buffers"]
    _poison(plan)
    table = b["table"]
    table_bytes = 0 if table is None else table.numel() * table.element_size()
    stream = torch.cuda.current_stream().cuda_stream
    print(f"noalloc: launches {plan['launches']}, table {table_bytes} bytes", flush=True)
    plan["ctx"].launch_noalloc(plan["module"], plan["launches"], plan["args"], table_bytes,
                               stream)
    flags = plan["ctx"].read_status(b["status"], stream)
    assert flags == 0
    return collect_outputs(plan, flags)["result"].cpu().numpy()


def _four_level(side, steps, **kw):
    w = workloads.four_level(n_eps=side, n_amp=side, n_t=steps + 1, t_end=0.05 * steps)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    kw = dict(var_arrays=w["var_arrays"], fp64=True, **kw)
    return prepare_sweep(*args, **kw), mesolve_2D(*args, **kw)


@pytest.mark.parametrize("solver", ["rk4", "tsit5a"])
def test_noalloc_matches_the_allocating_launch(gpu, solver):
    kw = {"solver": solver}
    if solver == "tsit5a":
        kw["adaptive_options"] = {"step_rtol": 1e-8, "step_atol": 1e-10, "max_substeps": 1000}
    # two calls in a row, the second with fewer points and steps (a shorter table)
    for side, steps in ((16, 400), (8, 120)):
        prep, ref = _four_level(side, steps, **kw)
        if solver == "rk4":
            assert prep.plan["buffers"]["table"] is not None
        got = _run_noalloc(prep)
        np.testing.assert_array_equal(got, ref)
        prep.close()


def test_noalloc_rejects_a_table_beyond_the_scratch(gpu):
    prep, _ = _four_level(4, 10)
    plan = prep.plan
    with pytest.raises(ValueError, match="exceeds"):
        plan["ctx"].launch_noalloc(plan["module"], plan["launches"], plan["args"],
                                   runtime.SCRATCH_BYTES)
    prep.close()
