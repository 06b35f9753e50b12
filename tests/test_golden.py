"""Golden fixtures (tests/golden/): the oracle still reproduces them, and the benchmark workloads
are the ones recorded in configs.json.  The GPU side is tests/test_gpu_golden.py."""
import json
import os

import numpy as np
import pytest

from gqis import workloads
from oracle import integrate, models

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(HERE, name))


def _close(a, b, rel=1e-12):
    assert np.max(np.abs(a - b)) <= rel * max(1.0, float(np.max(np.abs(b))))


def test_oracle_reproduces_lzs_fixture():
    g = _load("c1_lzs_64x64_rk4_mean_fp64.npz")
    m = models.lzs_two_level(g["eps"], g["amp"])
    res, _ = integrate.rk4(m.f, m.ground_state(), float(g["t_end"]) / (int(g["n_t"]) - 1),
                           int(g["n_t"]) - 1, obs=m.obs)
    _close(res.reshape(64, 64), g["result"])


RANDOM_FIXTURES = {4: "c3_random_4_level_8x8_subset_rk4_mean_fp64.npz",
                   8: "c3_random_8_level_8x8_subset_rk4_mean_fp64.npz",
                   16: "c3_random_16_level_8x8_subset_rk4_mean_fp64.npz",
                   32: "c3_random_32_level_4x4_subset_rk4_mean_fp64.npz"}


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_oracle_reproduces_random_n_level_fixture(n):
    g = _load(RANDOM_FIXTURES[n])
    m = models.random_n_level(n, int(g["seed"]), g["eps"], g["amp"])
    res, _ = integrate.rk4(m.f, m.ground_state(), 0.01, 2000, obs=m.obs)
    assert g["result"].shape == (256 // int(g["stride"]),) * 2
    _close(res.reshape(g["result"].shape), g["result"])
    # the fixture's axes are exact subsets of the 256 x 256 benchmark grid
    np.testing.assert_array_equal(np.linspace(-5.0, 5.0, 256)[::int(g["stride"])], g["eps"])
    np.testing.assert_array_equal(np.linspace(0.0, 5.0, 256)[::int(g["stride"])], g["amp"])


def test_oracle_reproduces_duffing_fixture():
    g = _load("c4_duffing_8x8_subset_t20_tsit5a_fp64.npz")
    d = models.Duffing(g["gamma"], g["x0"])
    res, _, acc, rej = integrate.tsit5a(d.f, d.y0(), 0.1, 200, rtol=float(g["rtol"]),
                                        atol=float(g["atol"]), max_substeps=1000)
    _close(res.reshape(8, 8, 2), g["result"])
    np.testing.assert_array_equal(acc.reshape(8, 8), g["accepted"])
    # layout: entry [i, j] is the point (gamma[i], x0[j]) integrated on its own
    one = models.Duffing(g["gamma"][2:3], g["x0"][5:6])
    r1, _, _, _ = integrate.tsit5a(one.f, one.y0(), 0.1, 200, rtol=float(g["rtol"]),
                                   atol=float(g["atol"]), max_substeps=1000)
    _close(r1.reshape(2), g["result"][2, 5])


def test_fixture_grids_are_subsets_of_the_benchmark_grids():
    cfg = json.load(open(os.path.join(HERE, "configs.json")))
    g = _load("c2_four_level_16x16_subset_rk4_mean_fp64.npz")
    w = workloads.four_level(n_eps=2048, n_amp=2048)
    eps, amp = w["var_arrays"].values()
    np.testing.assert_array_equal(eps[::128], g["eps"])
    np.testing.assert_array_equal(amp[::128], g["amp"])
    c = cfg["four_level_2048"]
    assert [eps[0], eps[-1], eps.size] == c["eps"] and [amp[0], amp[-1], amp.size] == c["amp"]
    assert [w["tlist"][0], w["tlist"][-1], w["tlist"].size] == c["t"]
    c = cfg["random_n_level_256"]
    w = workloads.random_n_level(8, seed=c["seed"])
    eps, amp = w["var_arrays"].values()
    assert [eps[0], eps[-1], eps.size] == c["eps"] and [amp[0], amp[-1], amp.size] == c["amp"]
    assert [w["tlist"][0], w["tlist"][-1], w["tlist"].size] == c["t"]
    # the product's seeded operators are the oracle's
    for a, b in zip(workloads.random_n_level_operators(8, c["seed"]),
                    models.random_n_level_operators(8, c["seed"])):
        np.testing.assert_array_equal(np.diag(b) if b.ndim == 2 and a.ndim == 1 else b, a)
    c = cfg["lzs_two_level_64"]
    w = workloads.lzs_two_level()
    eps, amp = w["var_arrays"].values()
    assert [eps[0], eps[-1], eps.size] == c["eps"] and [amp[0], amp[-1], amp.size] == c["amp"]
    c = cfg["duffing_1024"]
    w = workloads.duffing()
    gam, x0 = w["var_arrays"].values()
    assert [gam[0], gam[-1], gam.size] == c["gamma"] and [x0[0], x0[-1], x0.size] == c["x0"]
    s = workloads.serving_stream(count=50)
    assert {d["side"] for d in s} <= set(cfg["serving_stream"]["sides"])
    assert all(5.0 <= d["amp_max"] <= 30.0 for d in s)
