"""Serving: continuous batching gives every request the same result alone and in the mix."""
import numpy as np
import pytest

from gqis import compile_lindblad_rhs, mesolve_2D, workloads
from gqis.serving import SweepRequest, SweepServer

pytestmark = pytest.mark.gpu


def _models():
    out = {}
    for key, w in (("lzs", workloads.lzs_two_level(n_t=401, t_end=40.0)),
                   ("c2", workloads.chain(2, n_t=401, t_end=20.0)),
                   ("c4", workloads.chain(4, n_t=401, t_end=20.0))):
        out[key] = (compile_lindblad_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                                         list(w["var_arrays"])), w)
    return out


def _requests(models, seed=3, count=14):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        key = ["lzs", "c2", "c4"][i % 3]
        c, w = models[key]
        side = int(rng.choice([4, 16, 33]))
        axes = (np.linspace(-5, 5, side), np.linspace(0, float(rng.uniform(2, 10)), side + 1))
        reqs.append(SweepRequest(c, w["tlist"], axes, solver=["rk4", "tsit5a"][i % 2],
                                 fp64=bool(i % 4 < 2), arrival=0.001 * i))
    return reqs


def test_each_result_is_the_same_alone_and_in_the_mix(gpu):
    models = _models()
    alone = []
    for r in _requests(models):
        srv = SweepServer(streams=1, max_batch_points=1)
        alone.append(srv.serve([r], realtime=False)[0].result)
    mixed_reqs = _requests(models)
    srv = SweepServer(streams=4)
    srv.serve(mixed_reqs, realtime=False)
    assert srv.launches < len(mixed_reqs)          # requests were merged into shared launches
    assert max(r.batch_size for r in mixed_reqs) > 1
    for a, r in zip(alone, mixed_reqs):
        assert r.status == 0
        np.testing.assert_array_equal(a, r.result)


def test_server_matches_the_api(gpu):
    models = _models()
    c, w = models["c4"]
    axes = (np.linspace(-3, 3, 7), np.linspace(0, 6, 5))
    for solver in ("rk4", "tsit5a"):
        for fp64 in (False, True):
            r = SweepRequest(c, w["tlist"], axes, solver=solver, fp64=fp64)
            SweepServer(streams=2).serve([r], realtime=False)
            api = mesolve_2D(c, tlist=w["tlist"], var_arrays=dict(zip(c.sweep_symbols, axes)),
                             solver=solver, fp64=fp64)
            np.testing.assert_array_equal(r.result, api)


def test_non_finite_request_does_not_contaminate_the_batch(gpu):
    models = _models()
    c, w = models["c4"]
    good = SweepRequest(c, w["tlist"], (np.linspace(-1, 1, 8), np.linspace(0, 2, 8)))
    # eps = 1e30 overflows the RHS: that request's points are NaN/Inf, flagged
    bad = SweepRequest(c, w["tlist"], (np.array([1e30]), np.array([1.0])))
    srv = SweepServer(streams=1)
    srv.serve([good, bad], realtime=False)
    assert good.batch_size == 2
    assert bad.status == 1 and good.status == 0
    assert not np.all(np.isfinite(bad.result))
    assert np.all(np.isfinite(good.result))
    ref = mesolve_2D(c, tlist=w["tlist"], var_arrays=dict(zip(c.sweep_symbols, good.axes)))
    np.testing.assert_array_equal(good.result, ref)


def test_large_models_alone_and_in_the_mix(gpu):
    # 63 states: the server's measured choice may split each point over several lanes; the
    # choice is made once per kernel, so alone and mixed launches run the same code
    w = workloads.chain(8, n_t=201, t_end=10.0)
    c = compile_lindblad_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                             list(w["var_arrays"]))
    rng = np.random.default_rng(11)

    def reqs():
        out = []
        for i in range(6):
            side = [64, 128, 20][i % 3]
            axes = (np.linspace(-4, 4, side), np.linspace(0, 3.0 + i, side))
            out.append(SweepRequest(c, w["tlist"], axes, solver=["rk4", "tsit5a"][i % 2],
                                    fp64=i < 4, arrival=0.001 * i))
        return out
    alone = [SweepServer(streams=1, max_batch_points=1).serve([r], realtime=False)[0].result
             for r in reqs()]
    mixed = reqs()
    srv = SweepServer(streams=3)
    srv.serve(mixed, realtime=False)
    for a, r in zip(alone, mixed):
        assert r.status == 0
        np.testing.assert_array_equal(a, r.result)
    for r in mixed:
        api = mesolve_2D(c, tlist=w["tlist"], var_arrays=dict(zip(c.sweep_symbols, r.axes)),
                         solver=r.solver, fp64=r.fp64)
        scale = float(np.max(np.abs(api)))
        tol = 1e-10 if r.fp64 else 1e-5
        assert float(np.max(np.abs(r.result - api))) <= tol * scale
    print("chosen", {k[1:3]: (m["groups"], m["block"]) for k, m in srv.modules.items()},
          flush=True)
