"""ctypes access to oracle/liboracle_rk4.so (test infrastructure / CPU baseline only)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_rk4.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} missing: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
        d = ctypes.POINTER(ctypes.c_double)
        _LIB.rk4_affine_sweep.restype = ctypes.c_int
        _LIB.rk4_affine_sweep.argtypes = [ctypes.c_int, d, d, d, d, d, d, ctypes.c_double,
                                          ctypes.c_longlong, d, d, d, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, d, d, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, d, ctypes.c_int]
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def rk4_affine(model, dt, num_steps, *, warmup=0, t_in=0.0, mode="mean", threads=0, omega=1.0,
               y0=None):
    """RK4 sweep of an ``oracle.models.AffineSweep`` built from three terms
    (1, px, py*cos(omega t)); returns complex means (P,) or final states (P, D)."""
    (g0, m0, b0), (g1, m1, b1), (g2, m2, b2) = model.parts
    arrs = [np.ascontiguousarray(x, np.float64) for x in (m0, b0, m1, b1, m2, b2)]
    px = np.ascontiguousarray(model.px, np.float64)
    py = np.ascontiguousarray(model.py, np.float64)
    o_re = np.ascontiguousarray(model.o.real, np.float64)
    o_im = np.ascontiguousarray(model.o.imag, np.float64)
    d = model.dim
    npts = px.size
    out = np.zeros(2 * npts if mode == "mean" else npts * d, np.float64)
    y0a = None if y0 is None else np.ascontiguousarray(y0, np.float64)
    rc = lib().rk4_affine_sweep(d, *[_p(a) for a in arrs], float(omega), npts, _p(px), _p(py),
                                _p(y0a), float(dt), int(num_steps), int(warmup), float(t_in),
                                _p(o_re), _p(o_im), float(model.o0.real), float(model.o0.imag),
                                0 if mode == "mean" else 1, _p(out), int(threads))
    if rc != 0:
        raise ValueError("rk4_affine_sweep: bad arguments")
    if mode == "mean":
        return out[0::2] + 1j * out[1::2]
    return out.reshape(npts, d)
