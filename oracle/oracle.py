"""ctypes wrapper of the CPU oracle (libsse_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of bench.py to check or time the reference algorithm on
the CPU. The product (cloud.jl_amd/) never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsse_oracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.sse_oracle_residual.argtypes = [vp, vp, vp]
        L.sse_oracle_residual.restype = C.c_int
        L.sse_oracle_residual_repeat.argtypes = [vp, vp, vp, C.c_int]
        L.sse_oracle_residual_repeat.restype = C.c_int
        L.sse_oracle_rk_steps.argtypes = [vp, vp, C.c_double, C.c_int]
        L.sse_oracle_rk_steps.restype = C.c_int
        L.sse_oracle_erk_steps.argtypes = [vp, vp, C.c_double, C.c_int, C.c_int, vp, vp]
        L.sse_oracle_erk_steps.restype = C.c_int
        L.sse_oracle_last_error.restype = C.c_char_p
        L.sse_oracle_set_threads.argtypes = [C.c_int]
        L.sse_oracle_max_threads.restype = C.c_int
        L.sse_oracle_ck2n54.argtypes = [dp, dp, dp]
        L.sse_oracle_residual_jvp.argtypes = [vp, vp, vp, C.c_double, C.c_double, vp, vp]
        L.sse_oracle_residual_jvp.restype = C.c_int
        L.sse_oracle_rk_jvp.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_int]
        L.sse_oracle_rk_jvp.restype = C.c_int
        L.sse_oracle_two_point.argtypes = [C.c_int, C.c_int, C.c_double, vp, C.c_int, vp, vp, vp]
        L.sse_oracle_c2e.argtypes = [C.c_int, C.c_double, vp, vp]
        L.sse_oracle_e2c.argtypes = [C.c_int, C.c_double, vp, vp]
        L.sse_oracle_logmean.argtypes = [C.c_double, C.c_double]
        L.sse_oracle_logmean.restype = C.c_double
        L.sse_oracle_inv_logmean.argtypes = [C.c_double, C.c_double]
        L.sse_oracle_inv_logmean.restype = C.c_double
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise RuntimeError("oracle: " + lib().sse_oracle_last_error().decode())


def set_threads(n: int):
    lib().sse_oracle_set_threads(int(n))


def residual(arrays, u: np.ndarray) -> np.ndarray:
    """semi_discrete_residual! on the CPU; ``arrays`` is a ``sse.solvers.DescArrays``."""
    u = np.asfortranarray(u, dtype=np.float64)
    dudt = np.zeros_like(u, order="F")
    _check(lib().sse_oracle_residual(C.addressof(arrays.desc), u.ctypes.data, dudt.ctypes.data))
    return dudt


def residual_repeat(arrays, u: np.ndarray, nrep: int) -> np.ndarray:
    u = np.asfortranarray(u, dtype=np.float64)
    dudt = np.zeros_like(u, order="F")
    _check(lib().sse_oracle_residual_repeat(C.addressof(arrays.desc), u.ctypes.data, dudt.ctypes.data, nrep))
    return dudt


def rk_steps(arrays, u: np.ndarray, dt: float, nsteps: int) -> np.ndarray:
    u = np.array(u, dtype=np.float64, order="F", copy=True)
    _check(lib().sse_oracle_rk_steps(C.addressof(arrays.desc), u.ctypes.data, float(dt), int(nsteps)))
    return u


def erk_steps(arrays, u: np.ndarray, dt: float, nsteps: int, A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """nsteps steps of the explicit RK method with Butcher tableau (A, b)."""
    u = np.array(u, dtype=np.float64, order="F", copy=True)
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    s = len(b)
    assert A.shape == (s, s)
    _check(lib().sse_oracle_erk_steps(C.addressof(arrays.desc), u.ctypes.data, float(dt), int(nsteps), s,
                                      A.ctypes.data, b.ctypes.data))
    return u


def residual_jvp(arrays, u: np.ndarray, du: np.ndarray, dlambda: float = 0.0, db: float = 0.0):
    """(R(u), R'(u) du + dR/dλ dλ + dR/db db) by forward-mode dual numbers through
    the oracle (λ the Lax-Friedrichs coefficient, b the diffusion coefficient)."""
    u = np.asfortranarray(u, dtype=np.float64)
    du = np.asfortranarray(du, dtype=np.float64)
    r, dr = np.zeros_like(u, order="F"), np.zeros_like(u, order="F")
    _check(lib().sse_oracle_residual_jvp(C.addressof(arrays.desc), u.ctypes.data, du.ctypes.data,
                                         float(dlambda), float(db), r.ctypes.data, dr.ctypes.data))
    return r, dr


def rk_jvp(arrays, u: np.ndarray, du: np.ndarray, dt: float, nsteps: int, dlambda: float = 0.0,
           db: float = 0.0):
    """(u_N, du_N): nsteps CK2N54 steps with the tangent carried along."""
    u = np.array(u, dtype=np.float64, order="F", copy=True)
    du = np.array(du, dtype=np.float64, order="F", copy=True)
    _check(lib().sse_oracle_rk_jvp(C.addressof(arrays.desc), u.ctypes.data, du.ctypes.data, float(dlambda),
                                   float(db), float(dt), int(nsteps)))
    return u, du


def ck2n54():
    A, B, Cc = (np.zeros(5) for _ in range(3))
    dp = C.POINTER(C.c_double)
    lib().sse_oracle_ck2n54(A.ctypes.data_as(dp), B.ctypes.data_as(dp), Cc.ctypes.data_as(dp))
    return A, B, Cc


def two_point(law_kind: int, d: int, gamma: float, uL, uR, flux: int = 1, a=None) -> np.ndarray:
    N_c = d + 2 if law_kind == 1 else 1
    uL = np.ascontiguousarray(uL, dtype=np.float64)
    uR = np.ascontiguousarray(uR, dtype=np.float64)
    F = np.zeros(N_c * d)
    av = None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    lib().sse_oracle_two_point(law_kind, d, gamma, None if av is None else av.ctypes.data, flux,
                               uL.ctypes.data, uR.ctypes.data, F.ctypes.data)
    return F.reshape(d, N_c).T  # N_c x d


def c2e(d: int, gamma: float, u) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.zeros(d + 2)
    lib().sse_oracle_c2e(d, gamma, u.ctypes.data, w.ctypes.data)
    return w


def e2c(d: int, gamma: float, w) -> np.ndarray:
    w = np.ascontiguousarray(w, dtype=np.float64)
    u = np.zeros(d + 2)
    lib().sse_oracle_e2c(d, gamma, w.ctypes.data, u.ctypes.data)
    return u


def logmean(x: float, y: float) -> float:
    return lib().sse_oracle_logmean(x, y)


def inv_logmean(x: float, y: float) -> float:
    return lib().sse_oracle_inv_logmean(x, y)
