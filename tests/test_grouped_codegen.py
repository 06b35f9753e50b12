"""The grouped (lanes-per-point) model code on the CPU: the emitted C++ of every wave group,
compiled with g++ for the host, reproduces the one-lane-per-point model functions row by row
(RHS, observable, initial state) -- so the row partition, the per-group closures of statics and
stage values and their renumbering are checked on the printed code itself.  The hiprtc build of
the grouped kernels for gfx950 is checked too."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import sympy as sp

from gqis import codegen, runtime, workloads
from gqis.gqis_ode_api import ode_model
from gqis.solver import _compile_rhs, _emit_source
from gqis.solvers import get_solver_spec

HARNESS = r"""
#include <cmath>
using std::cos; using std::sin; using std::exp; using std::sqrt; using std::pow;
using std::fabs; using std::fmin; using std::fmax;
#define __device__
#define __forceinline__ inline
#define __restrict__
typedef double real;
#define R(x) ((real)(x))
#define DBL(x) ((double)(x))
#define GQ_FP64 1
inline real gq_sign(real x) { return (x > R(0)) - (x < R(0)); }
inline double gq_reduce_2pi(double x) { return x; }
inline double gq_cos64(double x) { return std::cos(x); }
inline double gq_sin64(double x) { return std::sin(x); }
inline double gq_rcp(double x) { return 1.0 / x; }
namespace lane {
%(lane)s
}
namespace grp {
%(grp)s
template <int GI>
void run(const double* Y, double px, double py, const double* C, double t, double* dy,
         double* y0, double* obs) {
    using P = gq_grp<GI>;
    double Sg[P::NS], Gg[P::NG], TT[GQ_NATOMS_ALLOC], k[P::NR], yi[P::NR];
    P::statics(px, py, C, Sg);
    gq_atoms_real(t, C, TT);
    P::stage(px, py, C, Sg, TT, Gg);
    P::rhs(Y, 0, px, py, C, Sg, Gg, k);
    P::init(yi, px, py, C);
    for (int r = 0; r < P::NR; ++r) { dy[P::R0 + r] = k[r]; y0[P::R0 + r] = yi[r]; }
    if constexpr (GI == 0) P::obs(Y, 0, px, py, C, Sg, Gg, obs[0], obs[1]);
}
}
extern "C" void eval(const double* y, double px, double py, const double* C, double t,
                     double* dy_lane, double* dy_grp, double* y0_lane, double* y0_grp,
                     double* obs_lane, double* obs_grp) {
    lane::gq_static_t S;
    lane::gq_statics(px, py, C, S);
    double TT[GQ_NATOMS_ALLOC];
    lane::gq_atoms_real(t, C, TT);
    lane::gq_stage_t G;
    lane::gq_stage(px, py, C, S, TT, G);
    lane::gq_rhs(y, px, py, C, S, G, dy_lane);
    lane::gq_obs(y, px, py, C, S, G, obs_lane[0], obs_lane[1]);
    lane::gq_init_expr(y0_lane, px, py, C);
    static double Y[GQ_D * 64];
    for (int j = 0; j < GQ_D; ++j) Y[j * 64] = y[j];
    for (int g = 0; g < GQ_G; ++g) {
        GQ_GROUP_DISPATCH(g, grp::run, (Y, px, py, C, t, dy_grp, y0_grp, obs_grp))
    }
}
"""


def _build(tmp_path, compiled, groups):
    lane = codegen.emit_model(compiled.plan, compiled.rho0_exprs)
    grp = codegen.emit_grouped(compiled.plan, groups, compiled.rho0_exprs)
    src = tmp_path / f"grouped_{groups}.cpp"
    src.write_text(HARNESS % {"lane": lane, "grp": grp})
    lib = tmp_path / f"grouped_{groups}.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-w", str(src), "-o", str(lib)],
                   check=True)
    return ctypes.CDLL(str(lib))


def _models():
    w = workloads.chain(8)
    chain8, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                             sweep_symbols=tuple(w["var_arrays"]))
    w = workloads.four_level(n_eps=2, n_amp=2)
    four, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                           sweep_symbols=tuple(w["var_arrays"]))
    w = workloads.duffing(n_gamma=2, n_x0=2)
    m = ode_model(w["rhs"], w["state_symbols"], w["y0"], w["var_arrays"])
    duff, _ = _compile_rhs(m.equations, {}, [], m.observable, sweep_symbols=tuple(m.sweeps),
                           rho0=m.initial, ode_size=m.size, cse_simplify=True)
    return {"chain8": (chain8, (2, 4, 8)), "four_level": (four, (2, 4)), "duffing": (duff, (2,))}


@pytest.mark.parametrize("name", ["chain8", "four_level", "duffing"])
def test_grouped_code_reproduces_per_lane_code(tmp_path, name):
    compiled, group_counts = _models()[name]
    d = compiled.state_size
    nconst = max(1, len(compiled.plan.const_syms))
    rng = np.random.default_rng(3)
    for groups in group_counts:
        lib = _build(tmp_path, compiled, groups)
        for _ in range(3):
            y = rng.standard_normal(d)
            C = rng.standard_normal(nconst)
            px, py, t = rng.uniform(-5, 5), rng.uniform(0, 9), rng.uniform(0, 50)
            out = [np.zeros(d), np.zeros(d), np.zeros(d), np.zeros(d), np.zeros(2), np.zeros(2)]
            ptr = [o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for o in out]
            lib.eval(y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_double(px),
                     ctypes.c_double(py), C.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                     ctypes.c_double(t), *ptr)
            np.testing.assert_array_equal(out[1], out[0])      # RHS rows
            np.testing.assert_array_equal(out[3], out[2])      # initial state
            np.testing.assert_array_equal(out[5], out[4])      # observable


def test_group_partition_and_limits():
    rw, blocks = codegen.group_rows(63, 8)
    assert rw == 8 and blocks[0] == (0, 8) and blocks[-1] == (56, 63)
    assert codegen.group_rows(15, 4) == (4, [(0, 4), (4, 8), (8, 12), (12, 15)])
    w = workloads.lzs_two_level(n_eps=2, n_amp=2)
    c, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                        sweep_symbols=tuple(w["var_arrays"]))
    with pytest.raises(ValueError, match="do not split"):
        codegen.emit_grouped(c.plan, 4, None)          # 3 rows cannot feed 4 waves


def test_hiprtc_compiles_grouped_kernels_for_gfx950():
    compiled, _ = _models()["chain8"]
    spec = get_solver_spec("rk4", step_size=0.05)
    for fp64, groups, mode, trace in ((True, 4, "mean", True), (False, 8, "final_rho", True),
                                      (True, 8, "final", False)):
        src = _emit_source(compiled, spec, mode, False, trace, trace and mode == "final_rho",
                           fp64, True, 256, False, groups)
        assert "gq_rk4_grouped" in src
        assert runtime.compile_check(src, codegen.HIPRTC_BASE_OPTIONS) > 1000


@pytest.mark.parametrize("solver", ["dp5", "lserk4", "ab5", "tsit5a"])
def test_hiprtc_compiles_grouped_tableau_kernels(solver):
    # one method of each fixed-step scheme (tableau, 2N-storage, multistep) and an adaptive one
    compiled, _ = _models()["chain8"]
    spec = get_solver_spec(solver, step_size=0.05)
    for fp64, groups, mode, trace in ((True, 8, "mean", True), (False, 4, "final", False)):
        src = _emit_source(compiled, spec, mode, False, trace, False, fp64, True, 256, False,
                           groups)
        name = "gq_adaptive_grouped" if spec.adaptive else "gq_fixed_grouped"
        assert name in src
        assert runtime.compile_check(src, codegen.HIPRTC_BASE_OPTIONS) > 1000
