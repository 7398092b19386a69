"""Conservation, invariant and error functionals evaluated on the GPU.

Mirrors the reference's analysis types (src/Analysis/conservation.jl:1-189):
``PrimaryConservationAnalysis`` (mass), ``EnergyConservationAnalysis`` (energy
with the solver's mass matrix) and ``EntropyConservationAnalysis`` (mathematical
entropy), each with ``evaluate_conservation(u)`` and
``evaluate_conservation_residual(u, dudt)``, computed by libsse_hip's
``sse_functional`` with a fixed summation order; ``ErrorAnalysis``
(src/Analysis/error.jl:1-91) through ``sse_l2_error``; and ``analyze`` over the
time steps of a results directory (conservation.jl:191-274), which reads the
solutions that ``io.SaveCallback`` wrote.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Union

import numpy as np

from . import abi
from .solvers import Solver


class _Analysis:
    kind = None

    def __init__(self, solver: Solver, results_path: str = "./"):
        self.solver = solver
        self.results_path = results_path

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
            import torch
            torch.cuda.synchronize()  # inputs may come from another handle's stream; the call blocks anyway
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


class ErrorAnalysis:
    """``ErrorAnalysis(results_path, conservation_law, spatial_discretization)``
    (error.jl:13-57) with the default error quadrature, the volume quadrature:
    V_err = V, w_err = W, and J_q. The sums run on the GPU (``sse_l2_error``)."""

    def __init__(self, solver: Solver, results_path: str = "./", error_quadrature_rule=None):
        if error_quadrature_rule is not None:
            raise NotImplementedError("an error quadrature other than the volume quadrature")
        self.solver = solver
        self.results_path = results_path

    def exact_values(self, exact_solution, t: float) -> np.ndarray:
        """The exact solution at the volume quadrature nodes, (N_q, N_c, N_e)."""
        xyzq = self.solver.spatial_discretization.mesh.xyzq
        ue = np.asarray(exact_solution(tuple(xyzq), t))  # N_q x N_e x N_c
        return np.asfortranarray(np.transpose(ue, (0, 2, 1)), dtype=np.float64)


@dataclass
class ConservationAnalysisResultsWithDerivative:
    """(conservation.jl:41-45): rows are time steps, columns variables."""
    t: np.ndarray
    E: np.ndarray
    dEdt: np.ndarray


def _l2_error(analysis: ErrorAnalysis, sol, exact_solution, t: float, normalize: bool) -> np.ndarray:
    from .hip import check
    s = analysis.solver
    lib, h = s.lib, s.handle
    N_p, N_c, N_e = s.size()
    out = np.zeros(N_c)
    ue = analysis.exact_values(exact_solution, t)
    if isinstance(sol, np.ndarray):
        u = np.asfortranarray(sol, dtype=np.float64)
        code = lib.sse_l2_error(h, u.ctypes.data, ue.ctypes.data, int(normalize), abi.SSE_HOST,
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    else:
        import torch

        from .time_integration import to_device
        ued = to_device(ue)
        torch.cuda.synchronize()  # sol may come from another handle's stream; the call blocks anyway
        code = lib.sse_l2_error(h, sol.data_ptr(), ued.data_ptr(), int(normalize), abi.SSE_DEVICE,
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    check(lib, h, code)
    return out


def analyze(analysis, *args, normalize: bool = False, **kw):
    """The reference's ``analyze`` methods:

    - ``analyze(ErrorAnalysis, sol, exact_solution, t=0.0; normalize)`` (error.jl:58-91):
      the L2 error of each variable;
    - ``analyze(ConservationAnalysis, time_steps::list; normalize)`` (conservation.jl:236-274,
      entropy :191-219): E and dE/dt at the written steps of ``analysis.results_path``;
    - ``analyze(ConservationAnalysis, initial_time_step=0, final_time_step="final")``
      (conservation.jl:221-234): (initial, final, difference).
    """
    if isinstance(analysis, ErrorAnalysis):
        sol, exact_solution = args[0], args[1]
        t = args[2] if len(args) > 2 else kw.get("t", 0.0)
        return _l2_error(analysis, sol, exact_solution, float(t), normalize)
    if not isinstance(analysis, _Analysis):
        raise TypeError(f"no analyze method for {type(analysis).__name__}")
    from .io import load_solution
    path = analysis.results_path
    if args and isinstance(args[0], (list, tuple, np.ndarray)):
        steps: List[int] = [int(x) for x in args[0]]
        factor = 1.0
        if normalize:
            u0, _ = load_solution(path, steps[0])
            factor = np.abs(analysis.evaluate_conservation(u0))
        n = 1 if analysis.kind == abi.SSE_ENTROPY else analysis.solver.size()[1]
        t = np.empty(len(steps))
        E = np.empty((len(steps), n))
        dEdt = np.empty((len(steps), n))
        for i, step in enumerate(steps):
            u, du, t[i] = load_solution(path, step, load_du=True)
            E[i] = analysis.evaluate_conservation(u) / factor
            dEdt[i] = analysis.evaluate_conservation_residual(u, du) / factor
        return ConservationAnalysisResultsWithDerivative(t, E, dEdt)
    initial_step: Union[int, str] = args[0] if len(args) > 0 else kw.get("initial_time_step", 0)
    final_step: Union[int, str] = args[1] if len(args) > 1 else kw.get("final_time_step", "final")
    u0, _ = load_solution(path, initial_step)
    uf, _ = load_solution(path, final_step)
    initial = analysis.evaluate_conservation(u0)
    final = analysis.evaluate_conservation(uf)
    return initial, final, final - initial
