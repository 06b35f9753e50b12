"""The adaptive kernels' stage-time atoms (``gq_atoms_stage``), compiled for the CPU from the
emitted model code: in FP32 kernels each sin / cos of an affine function of t is reduced to
|r| <= pi/4 in FP64 and evaluated in FP32 (``gq_sincos32``), so the atoms keep FP32 accuracy at
any t; FP64 kernels evaluate them in FP64 (``gq_sincos64``).  A wrong part of 2 pi, a missing reduction or a sin / cos swap fails
these."""
import ctypes
import subprocess

import numpy as np
import pytest
import sympy as sp

from gqis import codegen, workloads
from gqis.solver import _compile_rhs

HARNESS = r"""
#include <cmath>
using std::cos; using std::sin; using std::exp; using std::sqrt; using std::pow;
using std::fabs; using std::fmin; using std::fmax; using std::rint; using std::fma;
#define __device__
#define __forceinline__ inline
#define __restrict__
typedef %(real)s real;
#define R(x) ((real)(x))
#define DBL(x) ((double)(x))
#define GQ_FP64 %(fp64)d
inline real gq_sign(real x) { return (x > R(0)) - (x < R(0)); }
%(reduce)s
%(model)s
extern "C" void atoms(const double* t, int n, const double* C, double* out) {
    real Cr[8];
    for (int i = 0; i < 8; ++i) Cr[i] = (real)C[i];
    for (int j = 0; j < n; ++j) {
        real TT[GQ_NATOMS_ALLOC];
        gq_atoms_stage(t[j], Cr, TT);
        for (int i = 0; i < GQ_NATOMS; ++i) out[j * GQ_NATOMS + i] = (double)TT[i];
    }
}
"""


def _reduce_source():
    # the device definition, as the kernels' prelude emits it
    pre = codegen.prelude(False, True, 256, {})
    body = pre[pre.index("__device__ __forceinline__ double gq_reduce_2pi"):]
    return body.replace("__device__ __forceinline__", "inline")


def _atoms_lib(tmp_path, plan, fp64):
    src = tmp_path / f"atoms_{int(fp64)}.cpp"
    src.write_text(HARNESS % {"real": "double" if fp64 else "float", "fp64": int(fp64),
                              "reduce": _reduce_source(), "model": codegen.emit_model(plan)})
    lib = tmp_path / f"atoms_{int(fp64)}.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-w",
                    str(src), "-o", str(lib)], check=True)
    fn = ctypes.CDLL(str(lib)).atoms
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return fn


def _run(fn, t, natoms, consts=()):
    c = np.zeros(8)
    c[:len(consts)] = consts
    out = np.zeros((len(t), natoms))
    fn(t.ctypes.data, len(t), c.ctypes.data, out.ctypes.data)
    return out


def test_fp32_atoms_keep_fp32_accuracy_at_large_times(tmp_path):
    # the qubit resonator's atoms cos(w t)^2 and cos(w t), w = 0.5, over 40 drive periods and
    # far beyond (t up to 1e5): FP32 accuracy of the reduced angle, where cos of the FP32
    # time would be off by up to 4e-3
    w = workloads.qubit_resonator(2, n_eps=2, n_amp=2)
    compiled, _ = _compile_rhs(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                               sweep_symbols=tuple(w["var_arrays"]))
    plan = compiled.plan
    assert [str(e) for _, e in plan.time_atoms] == ["cos(0.5*t)**2", "cos(0.5*t)"]
    t = np.concatenate([np.linspace(0.0, 40 * 4 * np.pi, 4001),
                        np.random.default_rng(5).uniform(1e4, 1e5, 2000)])
    want = np.stack([np.cos(0.5 * t) ** 2, np.cos(0.5 * t)], axis=1)
    got32 = _run(_atoms_lib(tmp_path, plan, False), t, 2)
    assert np.max(np.abs(got32 - want)) < 3e-7
    naive = np.cos(t.astype(np.float32).astype(np.float64) * 0.5)
    assert np.max(np.abs(naive - want[:, 1])) > 1e-3          # what the reduction avoids
    got64 = _run(_atoms_lib(tmp_path, plan, True), t, 2)
    assert np.max(np.abs(got64 - want)) < 1e-12


def test_fp32_atoms_sin_with_runtime_constant_and_non_trig_atoms(tmp_path):
    # sin(c t + 1) with c a runtime constant, and exp(-t) (t outside a trig function: FP64,
    # rounded to FP32).  An FP32 kernel holds c in FP32, so the atoms are those of float(c).
    from gqis.symbolic import TIME, make_plan
    t_sym, c, y = TIME, sp.Symbol("c", real=True), sp.Symbol("y0", real=True)
    plan = make_plan([sp.sin(c * t_sym + 1) * y + sp.exp(-t_sym / 7)], sp.Integer(0),
                     sp.Integer(0), [y], {}, (c,))
    atoms = {str(e) for _, e in plan.time_atoms}
    assert any("sin" in a for a in atoms) and any("exp" in a for a in atoms)
    t = np.random.default_rng(6).uniform(0.0, 3e4, 3000)
    got = _run(_atoms_lib(tmp_path, plan, False), t, len(plan.time_atoms), consts=(2.3,))
    for i, (_, e) in enumerate(plan.time_atoms):
        f = sp.lambdify((t_sym, c), e, "numpy")
        want = np.broadcast_to(f(t, float(np.float32(2.3))), t.shape)
        assert np.max(np.abs(got[:, i] - want)) < 3e-7 * max(1.0, float(np.max(np.abs(want))))


SINCOS_HARNESS = r"""
#include <cmath>
using std::rint; using std::fma;
#define __device__
#define __forceinline__ inline
%(reduce)s
extern "C" void sincos64(const double* x, int n, double* s, double* c) {
    for (int j = 0; j < n; ++j) gq_sincos64(x[j], &s[j], &c[j]);
}
"""


def test_fp64_sincos_is_within_two_ulp_of_the_exact_values(tmp_path):
    """gq_sincos64 (Cody-Waite reduction by pi/2 in two FP64 parts, Taylor polynomials to
    degree 17 / 16) against 40-digit mpmath values: within 2.5e-16 absolute (2 ulp of values
    near 1) for arguments up to 1e5, including points next to multiples of pi/4, where the
    quadrant changes; a dropped term, a wrong quadrant sign or an unsplit pi/2 fails this."""
    import mpmath as mp
    src = tmp_path / "sincos.cpp"
    src.write_text(SINCOS_HARNESS % {"reduce": _reduce_source()})
    lib = tmp_path / "sincos.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-w",
                    str(src), "-o", str(lib)], check=True)
    fn = ctypes.CDLL(str(lib)).sincos64
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(11)
    quarter = np.pi / 4 * np.arange(-40, 800)
    x = np.concatenate([rng.uniform(0.0, 600.0, 6000), rng.uniform(-1e5, 1e5, 3000),
                        quarter, np.nextafter(quarter, np.inf), np.nextafter(quarter, -np.inf),
                        [0.0, 1e-300, -0.5, 1.2 * 499.9]])
    s, c = np.zeros_like(x), np.zeros_like(x)
    fn(x.ctypes.data, len(x), s.ctypes.data, c.ctypes.data)
    with mp.workdps(40):
        ws = np.array([float(mp.sin(mp.mpf(float(v)))) for v in x])
        wc = np.array([float(mp.cos(mp.mpf(float(v)))) for v in x])
    assert np.max(np.abs(s - ws)) <= 2.5e-16 and np.max(np.abs(c - wc)) <= 2.5e-16
    # the library functions (glibc) are not closer than that either
    assert np.max(np.abs(np.cos(x) - wc)) <= 2.5e-16


SINCOS32_HARNESS = r"""
#include <cmath>
using std::rint; using std::fma; using std::fmaf;
#define __device__
#define __forceinline__ inline
%(reduce)s
extern "C" void sincos32(const double* x, int n, float* s, float* c) {
    for (int j = 0; j < n; ++j) gq_sincos32(x[j], &s[j], &c[j]);
}
"""


def test_fp32_sincos_is_within_two_fp32_ulp_of_the_exact_values(tmp_path):
    """gq_sincos32 (the FP64 Cody-Waite reduction to |r| <= pi/4, the angle rounded to FP32,
    Taylor polynomials to degree 11 / 12 in FP32) against 40-digit mpmath values: within
    1.2e-7 absolute (two FP32 ulp of values just below 1) for arguments up to 1e5, including
    the points next to multiples of pi/4; a dropped low-order term (the degree-9 one of sin is
    3.1e-7 at pi/4), a wrong quadrant or the FP32 angle before the reduction fails this."""
    import mpmath as mp
    src = tmp_path / "sincos32.cpp"
    src.write_text(SINCOS32_HARNESS % {"reduce": _reduce_source()})
    lib = tmp_path / "sincos32.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-w",
                    str(src), "-o", str(lib)], check=True)
    fn = ctypes.CDLL(str(lib)).sincos32
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(12)
    quarter = np.pi / 4 * np.arange(-40, 800)
    x = np.concatenate([rng.uniform(0.0, 600.0, 6000), rng.uniform(-1e5, 1e5, 3000),
                        quarter, np.nextafter(quarter, np.inf), np.nextafter(quarter, -np.inf),
                        [0.0, -0.5, 1.2 * 499.9]])
    s = np.zeros(len(x), np.float32)
    c = np.zeros(len(x), np.float32)
    fn(x.ctypes.data, len(x), s.ctypes.data, c.ctypes.data)
    with mp.workdps(40):
        ws = np.array([float(mp.sin(mp.mpf(float(v)))) for v in x])
        wc = np.array([float(mp.cos(mp.mpf(float(v)))) for v in x])
    es = np.max(np.abs(s.astype(np.float64) - ws))
    ec = np.max(np.abs(c.astype(np.float64) - wc))
    assert es <= 1.2e-7 and ec <= 1.2e-7, (es, ec)
    # FP32 sin of the FP32-rounded argument, which the reduction in FP64 avoids
    naive = np.sin(x.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(naive - ws)) > 1e-3


def test_negative_powers_print_as_gq_rcp_and_match_the_division_on_the_host(tmp_path):
    """A standalone negative integer power in model code (the form common subexpressions such as
    1 / sqrt(d^2 + D^2) take) goes through gq_rcp, whose FP32 device overload is the hardware
    reciprocal and whose FP64 one an IEEE division; a quotient inside a product stays a
    division.  Compiled for the CPU from the prelude, the FP64 overload equals 1.0 / x bit for
    bit, so FP64 kernels keep their results."""
    x, y = sp.symbols("x y", real=True)
    pr = codegen.HipPrinter({"x": "x", "y": "y"})
    assert pr.doprint(1 / x) == "gq_rcp(x)"
    assert pr.doprint(x**-3) == "gq_rcp(x*x*x)"
    assert pr.doprint(y / x**3) == "y/(x*x*x)"
    src = tmp_path / "rcp.cpp"
    pre = codegen.prelude(True, True, 256, {})
    body = pre[pre.index("__device__ __forceinline__ double gq_rcp"):]
    body = body[:body.index("#ifndef GQ_HILBERT")] if "#ifndef GQ_HILBERT" in body else body
    src.write_text("#define __device__\n#define __forceinline__ inline\n" + body.split("#define ")[0]
                   + '\nextern "C" void rcp(const double* x, int n, double* out) {'
                   " for (int j = 0; j < n; ++j) out[j] = gq_rcp(x[j]); }\n")
    lib = tmp_path / "rcp.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-w",
                    str(src), "-o", str(lib)], check=True)
    fn = ctypes.CDLL(str(lib)).rcp
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    v = np.concatenate([np.random.default_rng(13).uniform(-1e3, 1e3, 4000), [1e-300, -3.0, 7.0]])
    out = np.zeros_like(v)
    fn(v.ctypes.data, len(v), out.ctypes.data)
    assert np.array_equal(out, 1.0 / v)
