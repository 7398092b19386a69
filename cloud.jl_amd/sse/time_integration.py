"""Fixed-step time integration of the semi-discrete system on the GPU.

The reference hands its ``ODEProblem`` to OrdinaryDiffEq and integrates with
``CarpenterKennedy2N54()`` at fixed dt (``test/test_driver.jl:77-83``), or
``DP8()`` (``test/euler_3d.jl:45-53``). Here ``solve`` advances the state with
libsse_hip's ``sse_rk_steps``, which fuses each stage's low-storage update into
the residual kernel, or with ``sse_erk_steps`` for a general explicit tableau.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import abi
from .solvers import ODEProblem, Solver


class CarpenterKennedy2N54:
    """Five-stage, fourth-order, 2N-storage RK of Carpenter & Kennedy (1994), NASA
    TM-109112 (the coefficients of OrdinaryDiffEq's ``CarpenterKennedy2N54``): stage s
    sets k = A[s] k + dt R(u), then u = u + B[s] k. The kernels hold the same table."""
    A = (0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
         -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0)
    B = (1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
         2277821191437.0 / 14882151754819.0)


class DP8:
    """Dormand-Prince 8(5,3) (Hairer, Norsett & Wanner, "Solving Ordinary Differential
    Equations I", 2nd ed., and Hairer's DOP853 code), the integrator of OrdinaryDiffEq's
    ``DP8()`` that the reference's 3D Euler test runs at fixed dt (test/euler_3d.jl:45-53):
    12 stages, the 8th-order propagating weights. The tableau is the published one as
    SciPy transcribes it (``scipy.integrate._ivp.dop853_coefficients``: A, B)."""

    def __init__(self):
        from scipy.integrate._ivp import dop853_coefficients as dop
        s = dop.N_STAGES
        self.A = np.ascontiguousarray(dop.A[:s, :s], dtype=np.float64)
        self.b = np.ascontiguousarray(dop.B, dtype=np.float64)
        self.c = np.ascontiguousarray(dop.C[:s], dtype=np.float64)

    @property
    def stages(self) -> int:
        return len(self.b)


def to_device(u: np.ndarray):
    """(N_p, N_c, N_e) Fortran-ordered host array -> contiguous CUDA tensor (N_e, N_c, N_p)."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asfortranarray(u).T)).to("cuda")


def to_host(t) -> np.ndarray:
    """Inverse of ``to_device``."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


def rk_steps(solver: Solver, u_dev, dt: float, nsteps: int, t0: float = 0.0):
    """Advance the device state ``u_dev`` in place by ``nsteps`` CK2N54 steps."""
    from .hip import check
    from .solvers import _check_device_tensor
    _check_device_tensor(u_dev, solver)
    lib, h = solver.lib, solver.handle
    check(lib, h, lib.sse_rk_steps(h, u_dev.data_ptr(), float(t0), float(dt), int(nsteps), abi.SSE_DEVICE))
    check(lib, h, lib.sse_synchronize(h))
    return u_dev


def erk_steps(solver: Solver, u_dev, dt: float, nsteps: int, alg, t0: float = 0.0):
    """Advance the device state ``u_dev`` in place by ``nsteps`` steps of the explicit
    RK method ``alg`` (a Butcher tableau: ``alg.A``, ``alg.b``; e.g. ``DP8()``)."""
    import ctypes as C
    from .hip import check
    from .solvers import _check_device_tensor
    _check_device_tensor(u_dev, solver)
    lib, h = solver.lib, solver.handle
    A = np.ascontiguousarray(alg.A, dtype=np.float64)
    b = np.ascontiguousarray(alg.b, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    check(lib, h, lib.sse_erk_steps(h, u_dev.data_ptr(), float(t0), float(dt), int(nsteps), len(b),
                                    A.ctypes.data_as(dp), b.ctypes.data_as(dp), abi.SSE_DEVICE))
    check(lib, h, lib.sse_synchronize(h))
    return u_dev


def _advance(ode: ODEProblem, alg, u, dt, n, t0):
    if n <= 0:
        return
    if alg is None or isinstance(alg, CarpenterKennedy2N54):
        rk_steps(ode.p, u, dt, n, t0)
    else:
        erk_steps(ode.p, u, dt, n, alg, t0)


@dataclass
class ODESolution:
    t: float
    u: np.ndarray


def solve(ode: ODEProblem, alg=None, dt: float = None, nsteps: int = None, callback=None) -> ODESolution:
    """``solve(ode, CarpenterKennedy2N54() or DP8(), dt=dt, adaptive=false)`` from tspan[0]
    to tspan[1] (or for ``nsteps`` steps). As OrdinaryDiffEq does with the end of
    tspan as a stop time, a last shorter step lands exactly on tspan[1].

    `callback` (e.g. ``sse.io.SaveCallback``, the reference's ``save_callback``,
    ``src/File/save.jl:74-86``) is called as callback(solver, u_device, t, step) at
    the start, after every ``callback.interval`` steps, and at the end."""
    if dt is None:
        raise ValueError("fixed-step integration needs dt")
    t0, t1 = ode.tspan
    u = to_device(ode.u0)
    if nsteps is not None:
        n, rest = int(nsteps), 0.0
    else:
        n = int(np.floor((t1 - t0) / dt * (1.0 + 1e-12)))
        rest = t1 - (t0 + n * dt)
        if rest <= 1e-12 * max(abs(t1), dt):
            rest = 0.0
    if callback is None:
        _advance(ode, alg, u, dt, n, t0)
        if rest:
            _advance(ode, alg, u, rest, 1, t0 + n * dt)
    else:
        every = max(1, int(callback.interval))
        callback(ode.p, u, t0, 0)
        step = 0
        while step < n:
            k = min(every - step % every, n - step)
            _advance(ode, alg, u, dt, k, t0 + step * dt)
            step += k
            if step % every == 0:
                callback(ode.p, u, t0 + step * dt, step)
        if rest:
            _advance(ode, alg, u, rest, 1, t0 + n * dt)
            step += 1
        callback(ode.p, u, t1 if nsteps is None else t0 + n * dt, step)  # end of tspan
    t_end = t0 + n * dt + rest if nsteps is None else t0 + n * dt
    return ODESolution(t1 if nsteps is None else t_end, to_host(u))
