"""Writes the golden fixtures of this directory from the FP64 oracle (``oracle/``).

The reference ships no result maps (its benchmark CSVs hold only timings), so these fixtures are
the oracle's FP64 answers on exact subsets of the benchmark grids, written once and committed:
the GPU tests compare the HIP kernels with them without recomputing the oracle, and a CPU test
checks that the oracle still reproduces them.  ``configs.json`` records the workload decisions
the fixtures and benchmarks use.  Run from the repository root:
``python tests/golden/make_golden.py [fixture names]``.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import integrate, models  # noqa: E402


def c1_lzs():
    eps, amp = np.linspace(-10.0, 10.0, 64), np.linspace(0.0, 20.0, 64)
    m = models.lzs_two_level(eps, amp)
    res, _ = integrate.rk4(m.f, m.ground_state(), 0.1, 2000, obs=m.obs)
    return dict(eps=eps, amp=amp, t_end=200.0, n_t=2001, result=res.reshape(64, 64))


def c2_four_level_subset():
    eps = np.linspace(-15.0, 15.0, 2048)[::128]
    amp = np.linspace(0.0, 30.0, 2048)[::128]
    m = models.four_level_chain(eps, amp)
    res, _ = integrate.rk4(m.f, m.ground_state(), 0.05, 2000, obs=m.obs)
    return dict(eps=eps, amp=amp, t_end=100.0, n_t=2001, stride=128, result=res.reshape(16, 16))


def c3_random_n_level(n, stride=32):
    eps = np.linspace(-5.0, 5.0, 256)[::stride]
    amp = np.linspace(0.0, 5.0, 256)[::stride]
    m = models.random_n_level(n, 2024, eps, amp)
    res, _ = integrate.rk4(m.f, m.ground_state(), 0.01, 2000, obs=m.obs)
    side = 256 // stride
    return dict(eps=eps, amp=amp, t_end=20.0, n_t=2001, seed=2024, stride=stride,
                result=res.reshape(side, side))


def c4_duffing_short():
    gam = np.linspace(0.1, 0.6, 1024)[::128]
    x0 = np.linspace(-2.0, 2.0, 1024)[::128]
    d = models.Duffing(gam, x0)
    res, _, acc, rej = integrate.tsit5a(d.f, d.y0(), 0.1, 200, rtol=1e-8, atol=1e-10,
                                        max_substeps=1000)
    return dict(gamma=gam, x0=x0, t_end=20.0, n_t=201, rtol=1e-8, atol=1e-10,
                result=res.reshape(8, 8, 2), accepted=acc.reshape(8, 8),
                rejected=rej.reshape(8, 8))


FIXTURES = {
    "c1_lzs_64x64_rk4_mean_fp64.npz": c1_lzs,
    "c2_four_level_16x16_subset_rk4_mean_fp64.npz": c2_four_level_subset,
    "c3_random_4_level_8x8_subset_rk4_mean_fp64.npz": lambda: c3_random_n_level(4),
    "c3_random_8_level_8x8_subset_rk4_mean_fp64.npz": lambda: c3_random_n_level(8),
    "c3_random_16_level_8x8_subset_rk4_mean_fp64.npz": lambda: c3_random_n_level(16),
    "c3_random_32_level_4x4_subset_rk4_mean_fp64.npz": lambda: c3_random_n_level(32, stride=64),
    "c4_duffing_8x8_subset_t20_tsit5a_fp64.npz": c4_duffing_short,
}


def main(names=()):
    """Write every fixture, or only those named on the command line."""
    for name, make in FIXTURES.items():
        if names and name not in names:
            continue
        data = make()
        np.savez_compressed(os.path.join(HERE, name), **data)
        print(name, {k: np.shape(v) for k, v in data.items()})


if __name__ == "__main__":
    main(sys.argv[1:])
