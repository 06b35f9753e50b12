"""Matrix-core (Hilbert-space) RK4 path against the FP64 oracle and the per-lane kernels."""
import numpy as np
import pytest

from gqis import mesolve_2D, workloads
from oracle import integrate, models

pytestmark = pytest.mark.gpu
SEED = 2024


@pytest.mark.parametrize("liouville", ["mfma", "mfma_f16"])
@pytest.mark.parametrize("n,steps", [(8, 400), (16, 200), (32, 100)])
def test_random_n_level_mean_and_state_match_oracle(gpu, n, steps, liouville):
    # mfma_f16: the commutator product as FP16 hi/lo parts (three products, FP32 accumulation)
    w = workloads.random_n_level(n, seed=SEED, n_eps=3, n_amp=2, n_t=steps + 1,
                                 t_end=0.01 * steps)
    eps_v, amp_v = w["var_arrays"].values()
    got, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                           var_arrays=w["var_arrays"], liouville=liouville,
                           return_timing_info=True)
    rho = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], liouville=liouville, output_mode="final_rho")
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


@pytest.mark.parametrize("n,liouville", [(16, "mfma"), (16, "mfma_f16"), (32, "mfma_f16")])
def test_mfma_continuation_from_final_state(gpu, n, liouville):
    # a sweep continued from its final state (rho0 values, t_in) equals the whole sweep; the
    # split forms read the initial state in their own lane layout
    w = workloads.random_n_level(n, seed=SEED, n_eps=2, n_amp=2, n_t=101, t_end=1.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"])
    first = mesolve_2D(*args, w["tlist"], var_arrays=w["var_arrays"], liouville=liouville,
                       output_mode="final_rho")
    second = mesolve_2D(*args, w["tlist"], var_arrays=w["var_arrays"], liouville=liouville,
                        output_mode="final_rho", rho0_values=first, t_in=1.0)
    whole = mesolve_2D(*args, np.linspace(0.0, 2.0, 201), var_arrays=w["var_arrays"],
                       liouville=liouville, output_mode="final_rho")
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


@pytest.mark.parametrize("n", [8, 16, 32])
def test_split_precision_products_over_2000_steps(gpu, n):
    # the random n-level scaling model over its full 2000 steps: FP16 hi/lo products stay at
    # the FP32 path's error level (<= 5e-7 of the map's scale measured), BF16 hi/lo (8 + 8
    # significant bits) at a few 1e-6; both against the FP64 path on the same 8 x 8 points
    w = workloads.random_n_level(n, seed=SEED, n_eps=8, n_amp=8)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    ref = mesolve_2D(*args, var_arrays=w["var_arrays"], fp64=True,
                     liouville="lanes" if n == 8 else "mfma")
    scale = float(np.max(np.abs(ref)))
    for lv, tol in (("mfma", 2e-6), ("mfma_f16", 2e-6), ("mfma_bf16", 5e-5)):
        got = mesolve_2D(*args, var_arrays=w["var_arrays"], liouville=lv)
        err = float(np.max(np.abs(got - ref)))
        print(n, lv, err / scale, flush=True)
        assert err <= tol * scale


def test_auto_uses_fp16_split_products_from_16_levels(gpu):
    w = workloads.random_n_level(16, seed=SEED, n_eps=4, n_amp=4, n_t=101, t_end=1.0)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    auto = mesolve_2D(*args, var_arrays=w["var_arrays"])
    f16 = mesolve_2D(*args, var_arrays=w["var_arrays"], liouville="mfma_f16")
    np.testing.assert_array_equal(auto, f16)
    with pytest.raises(ValueError, match="FP32-mode option"):
        mesolve_2D(*args, var_arrays=w["var_arrays"], liouville="mfma_f16", fp64=True)


def _hilbert_model(w):
    from gqis import hilbert
    from gqis.solver import _drive_map, _prepare_model
    model = _prepare_model(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                           sweep_symbols=tuple(w["var_arrays"]), const_values=None,
                           runtime_consts=None, keep_symbolic_consts=None,
                           auto_runtime_consts=False, rho0=None, ode_size=None)
    return hilbert.extract(model, _drive_map(model["Drive"]))


@pytest.mark.parametrize("n,steps", [(8, 300), (16, 150), (32, 60)])
def test_fp64_matrix_core_form_matches_fp64_valu_form(gpu, n, steps):
    # the FP64 matrix-core kernel (default FP64 form) against the FP64 VALU dot-product kernel:
    # every output mode, a ragged point count (3 x 3: N = 8 puts two points in a wavefront) and
    # a continuation from explicit initial states
    from gqis import hilbert
    w = workloads.random_n_level(n, seed=SEED + n, n_eps=3, n_amp=3, n_t=steps + 1,
                                 t_end=0.01 * steps)
    hm = _hilbert_model(w)
    kw = dict(var_arrays=w["var_arrays"], runtime_values=[], dt=0.01, t_in=0.0,
              num_steps=steps, warmup_steps=steps // 4, fp64=True)
    finals = {}
    for mode in ("mean", "final", "final_rho"):
        got = hilbert.run(hm, mode=mode, rho0_values=None, **kw)
        ref = hilbert.run(hm, mode=mode, rho0_values=None, valu=True, **kw)
        g, r = got["result"].cpu().numpy(), ref["result"].cpu().numpy()
        print(n, mode, float(np.max(np.abs(g - r))), flush=True)
        assert g.shape == r.shape and np.max(np.abs(g - r)) <= 1e-12 * max(1.0, np.max(np.abs(r)))
        finals[mode] = r
    again = hilbert.run(hm, mode="final_rho", rho0_values=finals["final_rho"], **kw)
    ref = hilbert.run(hm, mode="final_rho", rho0_values=finals["final_rho"], valu=True, **kw)
    assert np.max(np.abs(again["result"].cpu().numpy() - ref["result"].cpu().numpy())) <= 1e-12


@pytest.mark.parametrize("n,steps", [(8, 200), (16, 100)])
def test_fp64_matrix_core_final_state_matches_oracle(gpu, n, steps):
    w = workloads.random_n_level(n, seed=SEED, n_eps=2, n_amp=3, n_t=steps + 1,
                                 t_end=0.01 * steps)
    eps_v, amp_v = w["var_arrays"].values()
    rho = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                     var_arrays=w["var_arrays"], liouville="mfma", fp64=True,
                     output_mode="final_rho")
    m = models.random_n_level(n, SEED, eps_v, amp_v)
    fin, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, mode="final_rho")
    assert rho.shape == (2, 3, n * n - 1)
    assert np.max(np.abs(rho.reshape(6, -1) - fin)) <= 1e-12


def test_fp64_matrix_core_chain_matches_per_lane_kernel(gpu):
    # relaxation (single-entry jumps) and dephasing (diagonal collapse) of the 8-level chain
    w = workloads.chain(8, n_eps=5, n_amp=3, n_t=401, t_end=20.0)
    kw = dict(var_arrays=w["var_arrays"], fp64=True)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    for mode in ("mean", "final_rho"):
        mfma = mesolve_2D(*args, liouville="mfma", output_mode=mode, **kw)
        lanes = mesolve_2D(*args, liouville="lanes", output_mode=mode, **kw)
        assert np.max(np.abs(mfma - lanes)) <= 1e-12 * max(1.0, np.max(np.abs(lanes)))


@pytest.mark.parametrize("n,lv,fp64,tol", [(8, "mfma", True, 1e-12), (8, "mfma", False, 2e-5),
                                           (16, "mfma_f16", False, 2e-5)])
def test_matrix_core_time_traces(gpu, n, lv, fp64, tol):
    # observable traces after steps offset + k * stride, with the mean and with the final
    # state, against the per-point FP64 kernel (whose traces are pinned to the oracle's) at
    # N = 8, and at N = 16 against the FP64 matrix-core trace (the 255-state per-point FP64
    # kernel takes minutes to compile)
    w = workloads.random_n_level(n, seed=SEED, n_eps=3, n_amp=2, n_t=121, t_end=1.2)
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    ref_path = "lanes" if n == 8 else "mfma"
    for mode, every in (("mean", 7), ("final_rho", 20)):
        kw = dict(var_arrays=w["var_arrays"], output_mode=mode, return_time_trace=True,
                  time_trace_every=every, return_timing_info=True)
        res, tr, times, info = mesolve_2D(*args, liouville=lv, fp64=fp64, **kw)
        assert info["kernel"].startswith("gq_hilbert_rk4")
        rres, rtr, rtimes, rinfo = mesolve_2D(*args, liouville=ref_path, fp64=True, **kw)
        assert tr.shape == rtr.shape == (3, 2, len(times))
        np.testing.assert_allclose(times, rtimes, rtol=1e-6)
        scale = max(1.0, float(np.max(np.abs(rtr))))
        assert float(np.max(np.abs(tr - rtr))) <= tol * scale
        assert float(np.max(np.abs(res - rres))) <= tol * max(1.0, float(np.max(np.abs(rres))))


@pytest.mark.parametrize("nq,steps", [(3, 200), (4, 100), (5, 30)])
def test_multi_qubit_model_on_the_matrix_cores(gpu, nq, steps):
    # the reference's multi-qubit chain (Benchmark 04): each qubit's lowering operator has one
    # nonzero per row, so its jump term is the kernels' general gather; FP64 matrix-core form
    # against the oracle to 1e-10 of the map's scale, the FP32 and FP16 hi/lo forms to 2e-5
    n = 2 ** nq
    w = workloads.multi_qubit(nq, n_eps=2, n_amp=2, n_t=steps + 1, t_end=0.01 * steps)
    eps_v, amp_v = w["var_arrays"].values()
    args = (w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"])
    m = models.multi_qubit(nq, eps_v, amp_v)
    ref, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, obs=m.obs)
    fin, _ = integrate.rk4(m.f, m.ground_state(), 0.01, steps, mode="final_rho")
    scale = max(1.0, float(np.max(np.abs(ref))))
    forms = [("mfma", True, 1e-10)] + ([("mfma", False, 2e-5), ("mfma_f16", False, 2e-5)]
                                       if nq < 5 else [])
    for lv, fp64, tol in forms:
        got, info = mesolve_2D(*args, var_arrays=w["var_arrays"], liouville=lv, fp64=fp64,
                               return_timing_info=True)
        assert info["kernel"].startswith("gq_hilbert_rk4")
        print(nq, lv, fp64, float(np.max(np.abs(got.ravel() - ref))), flush=True)
        assert np.max(np.abs(got.ravel() - ref)) <= tol * scale
        rho = mesolve_2D(*args, var_arrays=w["var_arrays"], liouville=lv, fp64=fp64,
                         output_mode="final_rho")
        assert rho.shape == (2, 2, n * n - 1)
        assert np.max(np.abs(rho.reshape(4, -1) - fin)) <= tol


def test_automatic_choice_follows_the_density_of_h(gpu):
    # 16 levels: the dense random model takes the FP16 hi/lo matrix-core form, the sparse
    # 4-qubit chain (5 nonzeros per row) the per-point kernels, which are faster for it
    dense = workloads.random_n_level(16, seed=SEED, n_eps=2, n_amp=2, n_t=21, t_end=0.2)
    sparse = workloads.multi_qubit(4, n_eps=2, n_amp=2, n_t=21, t_end=0.2)
    kernels = []
    for w in (dense, sparse):
        _, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                             var_arrays=w["var_arrays"], return_timing_info=True)
        kernels.append(info["kernel"])
    assert kernels[0] == "gq_hilbert_rk4_f16x3_f32"
    assert not kernels[1].startswith("gq_hilbert")
    # 32 levels: the sparse 5-qubit chain takes the wavefront-per-point sparse form
    five = workloads.multi_qubit(5, n_eps=2, n_amp=2, n_t=11, t_end=0.1)
    for fp64 in (False, True):
        _, info = mesolve_2D(five["H"], five["Drive"], five["Col_Ops"], five["mean_operator"],
                             five["tlist"], var_arrays=five["var_arrays"], fp64=fp64,
                             return_timing_info=True)
        assert info["kernel"] == "gq_hilbert_sparse_" + ("f64" if fp64 else "f32")
