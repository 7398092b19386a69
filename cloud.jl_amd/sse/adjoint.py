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
  one recomputation per segment (``sse_rk_adjoint``);
* ``energy_adjoint_b(solver, u0, dt, nsteps)`` -> (J, dJ/db, M^{-1} dJ/du0) for
  linear advection-diffusion and J = 1/2 u_N' M u_N, by the discrete adjoint of
  the CK2N54 solve run on the forward kernels (see its docstring).
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


def energy_adjoint_helpers(solver: Solver):
    """The solvers ``energy_adjoint_b`` shares across a sweep of diffusion coefficients on
    this mesh: the law (a, b) and the velocity reversed (-a, b), whose handles it sets to
    each member's b (``sse_set_diffusion``, no new copy of the geometry), the viscous
    operator dR/db (velocity 0 at b = 1; R is affine in b), and the advection operator
    alone (``LinearAdvectionEquation(a)``: R(u) = R_adv(u) + b dR/db(u), and R_adv needs
    no auxiliary gradient)."""
    from .conservation_laws import LinearAdvectionDiffusionEquation, LinearAdvectionEquation
    law = solver.conservation_law
    if not isinstance(law, LinearAdvectionDiffusionEquation):
        raise NotImplementedError("the velocity-reversal adjoint is for linear advection-diffusion")
    adv = Solver(LinearAdvectionEquation(tuple(law.a)), solver.spatial_discretization, solver.form,
                 solver.strategy, solver.mass_solver, solver.parallelism)
    return (solver.with_law(LinearAdvectionDiffusionEquation(tuple(law.a), law.b)),
            solver.with_law(LinearAdvectionDiffusionEquation(tuple(-x for x in law.a), law.b)),
            solver.with_law(LinearAdvectionDiffusionEquation(tuple(0.0 for _ in law.a), 1.0)),
            adv)


def energy_adjoint_b(solver: Solver, u0, dt: float, nsteps: int, helpers=None, keep_bytes: int = 16 << 30):
    """Discrete adjoint of ``nsteps`` CK2N54 steps of the linear advection-diffusion
    residual R(u) = A(a, b) u (StandardForm, skew-symmetric mapping, LF, BR1) for
    J = 1/2 u_N' M u_N, M the solver's mass matrix. Returns (J, dJ/db, g) with
    g = M^{-1} dJ/du0 (the gradient as a state: dJ/du0 . du = <g, du>_M).

    The weak-form operator M A splits into skew-symmetric advection and symmetric
    LF and BR1 terms, so (M A(a, b))' = M A(-a, b): the transposed residual is the
    residual with the velocity reversed, conjugated by M. In the variables
    u~ = M^{-1} ubar, k~ = M^{-1} kbar the reverse of a stage (k = A_s k + dt R(u_s);
    u = u + B_s k) is
        k~ += B_s u~;  u~ += dt R_{-a}(k~);  dJ/db += dt <k~, dR/db(u_s)>_M;  k~ *= A_s
    with R_{-a} the residual with the velocity reversed and the same b, and dR/db(u) =
    R(u; a = 0, b = 1) (R is affine in b), from the solvers of
    ``energy_adjoint_helpers`` (pass them to share them across members; `solver` only
    supplies b, set on the first two with ``sse_set_diffusion``). <x, y>_M is
    ``sse_functional``. u~ starts at u_N. Every operator application is a forward residual
    on the GPU. When the 5 nsteps viscous terms dR/db(u_s) fit in `keep_bytes` of device
    memory, the forward sweep runs stage by stage and keeps them (per stage dR/db(u_s) and
    the advection residual R_adv(u_s), R = R_adv + b dR/db), and the reverse sweep needs no
    recomputation; otherwise the forward sweep runs on the library's RK steps (one
    residual per stage), keeps the step states, and each step's stages are recomputed
    once in the same way. Either way the reverse stage is one residual. `u0` is a device tensor (N_e, 1, N_p).

    Each solver's handle has its own stream, ordered against torch's default stream
    but not against the other handles' streams; the functionals synchronise the
    device before they read."""
    import torch

    from .analysis import EnergyConservationAnalysis
    from .time_integration import CarpenterKennedy2N54 as CK
    from .hip import check
    own = helpers is None
    fwd, rev, visc, adv = energy_adjoint_helpers(solver) if own else helpers
    b = float(solver.conservation_law.b)
    for h_ in (fwd, rev):
        check(h_.lib, h_.handle, h_.lib.sse_set_diffusion(h_.handle, b))
    energy = EnergyConservationAnalysis(visc)  # the mass matrix is the same for every law on the mesh
    r = torch.empty_like(u0)

    def step(u, keep):
        """One step, keeping the viscous terms dR/db(u_s) of the stage states; the stage
        residual is R_adv(u_s) + b dR/db(u_s)."""
        k = torch.zeros_like(u)
        for s in range(5):
            semi_discrete_residual(keep[s], u, visc)
            semi_discrete_residual(r, u, adv)
            r.add_(keep[s], alpha=b)
            k.mul_(CK.A[s]).add_(r, alpha=dt)
            u.add_(k, alpha=CK.B[s])

    try:
        keep_all = 5 * nsteps * u0.numel() * u0.element_size() <= keep_bytes
        u = u0.clone()
        if keep_all:  # stage by stage, keeping dR/db(u_s) of every stage
            kept = torch.empty((nsteps, 5) + tuple(u0.shape), dtype=u0.dtype, device=u0.device)
            for n in range(nsteps):
                step(u, kept[n])
        else:  # the library's RK steps (stage update fused into the residual), step states kept
            from .time_integration import rk_steps
            states = [u0.clone()]
            for _ in range(nsteps):
                rk_steps(fwd, u, dt, 1)
                states.append(u.clone())
            dvis = [torch.empty_like(u) for _ in range(5)]
        J = float(energy.evaluate_conservation(u)[0])
        ut = u.clone()
        kt = torch.empty_like(u)
        dJdb = 0.0
        for n in range(nsteps - 1, -1, -1):
            if keep_all:
                dvis = kept[n]
            else:
                step(states[n].clone(), dvis)
            kt.zero_()
            for s in range(4, -1, -1):
                kt.add_(ut, alpha=CK.B[s])
                semi_discrete_residual(r, kt, rev)
                ut.add_(r, alpha=dt)
                dJdb += dt * float(energy.evaluate_conservation_residual(kt, dvis[s])[0])
                kt.mul_(CK.A[s])
        return J, dJdb, ut
    finally:
        if own:
            for h_ in (fwd, rev, visc, adv):
                h_.close()
