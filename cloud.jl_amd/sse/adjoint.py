"""Derivatives of the residual and of the CK2N54 time integration on the GPU.

The reference has no derivative code; a Julia caller would expose these through
a ``ChainRulesCore.rrule`` of ``semi_discrete_residual!`` (the analogue of a
``torch.autograd.Function``). Here:

* ``residual_vjp(solver, u, ybar)`` -> (ubar, lambda_bar): the pull-back of the
  residual, ``ubar = (dR/du)' ybar`` and ``lambda_bar = (dR/dλ)' ybar`` for the
  Lax-Friedrichs coefficient λ (libsse_hip's ``sse_residual_vjp``);
* ``rrule_semi_discrete_residual(solver, u)`` -> (dudt, pullback);
* ``rk_adjoint(solver, u0, dt, nsteps, checkpoint_every)`` -> (J, dJ/du0, dJ/dλ)
  for the entropy integral J of the final CK2N54 state, with checkpointing and
  one recomputation per segment (``sse_rk_adjoint``).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .solvers import Solver, semi_discrete_residual


def _ptrs(x):
    if isinstance(x, np.ndarray):
        return x.ctypes.data, abi.SSE_HOST
    return x.data_ptr(), abi.SSE_DEVICE


def residual_vjp(solver: Solver, u, ybar, ubar=None):
    """(ubar, lambda_bar). Host arrays are Fortran (N_p, N_c, N_e); device tensors (N_e, N_c, N_p)."""
    from .hip import check
    lib, h = solver.lib, solver.handle
    if ubar is None:
        if isinstance(u, np.ndarray):
            ubar = np.zeros_like(u, order="F")
        else:
            import torch
            ubar = torch.empty_like(u)
    if isinstance(u, np.ndarray):
        u = np.asfortranarray(u, dtype=np.float64)
        ybar = np.asfortranarray(ybar, dtype=np.float64)
    pu, where = _ptrs(u)
    py, _ = _ptrs(ybar)
    po, _ = _ptrs(ubar)
    lam = C.c_double(0.0)
    check(lib, h, lib.sse_residual_vjp(h, pu, py, po, C.byref(lam), where))
    return ubar, lam.value


def rrule_semi_discrete_residual(solver: Solver, u):
    """``ChainRulesCore.rrule``-style: returns dudt and a pull-back ybar -> (ubar, lambda_bar)."""
    if isinstance(u, np.ndarray):
        dudt = np.zeros_like(u, order="F")
    else:
        import torch
        dudt = torch.empty_like(u)
    semi_discrete_residual(dudt, u, solver)

    def pullback(ybar):
        return residual_vjp(solver, u, ybar)

    return dudt, pullback


def rk_adjoint(solver: Solver, u0, dt: float, nsteps: int, checkpoint_every: int = 20):
    """(J, grad_u0, grad_lambda) for J = entropy integral after nsteps CK2N54 steps."""
    from .hip import check
    lib, h = solver.lib, solver.handle
    if isinstance(u0, np.ndarray):
        u0 = np.asfortranarray(u0, dtype=np.float64)
        grad = np.zeros_like(u0, order="F")
    else:
        import torch
        grad = torch.empty_like(u0)
    pu, where = _ptrs(u0)
    pg, _ = _ptrs(grad)
    J, lam = C.c_double(0.0), C.c_double(0.0)
    check(lib, h, lib.sse_rk_adjoint(h, pu, float(dt), int(nsteps), int(checkpoint_every), abi.SSE_ENTROPY,
                                     C.byref(J), pg, C.byref(lam), where))
    return J.value, grad, lam.value
