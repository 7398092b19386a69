"""Conservation and invariant functionals evaluated on the GPU.

Mirrors the reference's analysis types (src/Analysis/conservation.jl:1-189):
``PrimaryConservationAnalysis`` (mass), ``EnergyConservationAnalysis`` (energy
with the solver's mass matrix) and ``EntropyConservationAnalysis`` (mathematical
entropy), each with ``evaluate_conservation(u)`` and
``evaluate_conservation_residual(u, dudt)``, computed by libsse_hip's
``sse_functional`` with a fixed summation order.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .solvers import Solver


class _Analysis:
    kind = None

    def __init__(self, solver: Solver):
        self.solver = solver

    def _call(self, u, dudt):
        from .hip import check
        s = self.solver
        lib, h = s.lib, s.handle
        N_p, N_c, N_e = s.size()
        nout = 1 if self.kind == abi.SSE_ENTROPY else N_c
        out = np.zeros(nout)
        if isinstance(u, np.ndarray):
            u = np.asfortranarray(u, dtype=np.float64)
            du = None if dudt is None else np.asfortranarray(dudt, dtype=np.float64)
            code = lib.sse_functional(h, u.ctypes.data, None if du is None else du.ctypes.data, self.kind,
                                      abi.SSE_HOST, out.ctypes.data_as(C.POINTER(C.c_double)))
        else:
            code = lib.sse_functional(h, u.data_ptr(), None if dudt is None else dudt.data_ptr(), self.kind,
                                      abi.SSE_DEVICE, out.ctypes.data_as(C.POINTER(C.c_double)))
        check(lib, h, code)
        return out

    def evaluate_conservation(self, u) -> np.ndarray:
        return self._call(u, None)

    def evaluate_conservation_residual(self, u, dudt) -> np.ndarray:
        return self._call(u, dudt)


class PrimaryConservationAnalysis(_Analysis):
    """Mass of each conservative variable (conservation.jl:113-117, 145-152)."""
    kind = abi.SSE_MASS


class EnergyConservationAnalysis(_Analysis):
    """1/2 u' M u per variable (conservation.jl:131-143, 154-167)."""
    kind = abi.SSE_ENERGY


class EntropyConservationAnalysis(_Analysis):
    """Mathematical entropy (conservation.jl:119-129, 169-189)."""
    kind = abi.SSE_ENTROPY
