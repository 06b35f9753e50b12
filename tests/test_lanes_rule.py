"""The default wavefronts per point and the LDS copies of the grouped kernels (host logic): the
choices the measurements on gfx950 support, and the LDS limit of one compute unit."""
import pytest

from gqis.solver import LDS_BYTES_PER_CU, auto_lanes_per_point, grouped_buffers
from gqis.solvers import get_solver_spec

RK4 = get_solver_spec("rk4", step_size=0.1)
TSIT5A = get_solver_spec("tsit5a", step_size=0.1)
DP5 = get_solver_spec("dp5", step_size=0.1)
BIG, SMALL = 1 << 20, 64 * 64


@pytest.mark.parametrize("d,fp64,nnz,points,want", [
    (15, True, 66, BIG, 1),            # below 48 states: one lane
    (63, False, 409, BIG, 1),          # sparse FP32 resonator: one lane (4.84 ms vs 6.97 at 8)
    (63, True, 409, BIG, 1),           # sparse FP64: 126 words, one lane (9.27 vs 14.29)
    (63, True, 330, SMALL, 4),         # small grid: 16 FP64 rows per wavefront
    (63, True, 1754, BIG, 8),          # dense rows: 8 lanes
    (99, False, 640, BIG, 1),
    (143, False, 1033, BIG, 4),        # 203 ms vs 267 at 8
    (143, True, 786, 1 << 16, 8),      # FP64 chain: 20.3 ms vs 24.5 at 4
    (195, False, 1400, BIG, 8),        # 4 lanes would fit one work-group per CU
    (255, False, 2000, BIG, 8),
    (399, False, 3000, BIG, 8),        # single LDS copy, 1.40 s vs 1.86 at 16
    (12, False, None, BIG, 1),
    (63, False, None, BIG, 8),         # non-affine RHS: the dense rule
])
def test_rk4_defaults(d, fp64, nnz, points, want):
    assert auto_lanes_per_point(RK4, d, fp64, nnz, points, 2) == want


def test_adaptive_and_tableau_defaults():
    assert auto_lanes_per_point(TSIT5A, 63, False, 409, BIG, 2) == 1     # 0.55 s vs 0.85 at 8
    assert auto_lanes_per_point(TSIT5A, 63, True, 330, BIG, 2) == 8      # 126 words: grouped
    assert auto_lanes_per_point(TSIT5A, 399, False, 3000, BIG, 2) == 16  # 4.79 s vs 7.60 at 8
    assert auto_lanes_per_point(DP5, 63, False, 409, BIG, 2) == 8
    assert auto_lanes_per_point(DP5, 40, False, 200, BIG, 2) == 1
    # fixed-step tableaux keep their stage vectors per lane: 16 lanes once 8 would leave more
    # than 160 words (4-qubit chain, 255 FP32 states: DOP853 6.7 ns per point-step against
    # 35.4 with 8 lanes, Tsit5 1.17 against 1.59); the 2N-storage LSRK4 keeps three vectors
    assert auto_lanes_per_point(get_solver_spec("dop853", step_size=0.1), 255, False, 2000, BIG,
                                2) == 16
    assert auto_lanes_per_point(get_solver_spec("tsit5", step_size=0.1), 255, False, 2000, BIG,
                                2) == 16
    assert auto_lanes_per_point(get_solver_spec("lserk4", step_size=0.1), 255, False, 2000, BIG,
                                2) == 8
    # LSRK4 follows the RK4 rule: one lane at 63 FP32 states (15.8 ms against 30.4 with 8)
    assert auto_lanes_per_point(get_solver_spec("lserk4f", step_size=0.1), 63, False, 409, BIG,
                                2) == 1
    assert auto_lanes_per_point(TSIT5A, 255, False, 2000, BIG, 2) == 8   # adaptive: unchanged


def test_lds_copies():
    # two copies while they fit 160 KB, then one, then no grouped kernel
    assert grouped_buffers(RK4, 255, False, 8) == 2          # 2 x 65 KB
    assert grouped_buffers(RK4, 399, False, 8) == 1          # 2 x 102 KB do not fit
    assert grouped_buffers(RK4, 195, True, 8) == 1
    assert grouped_buffers(RK4, 1023, False, 16) == 0        # one copy is 262 KB
    assert 2 * 320 * 64 * 4 <= LDS_BYTES_PER_CU < 2 * 321 * 64 * 4
    # the adaptive kernels' error-norm partials and stage-time atoms count too
    assert grouped_buffers(TSIT5A, 143, True, 8, natoms=2) == 1
    assert grouped_buffers(RK4, 143, True, 8, natoms=2) == 2
    # a split that leaves a wavefront without rows, or a solver without a grouped kernel
    assert grouped_buffers(RK4, 15, False, 16) == 0
    assert grouped_buffers(get_solver_spec("rk4", step_size=0.1), 63, False, 1) == 0
