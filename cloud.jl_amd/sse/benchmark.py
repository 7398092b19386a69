"""The reference's residual benchmark on the GPU.

``scaling_test_euler_2d(p, M)`` restates ``scaling_test_euler_2d``
(``src/Analysis/benchmark.jl:116-183``): the isentropic vortex (Mach 0.4, angle 0,
strength sqrt(2/(γ-1) (1 - 0.75^(γ-1))), radius 1/10, centred in the unit square)
for 2D Euler on ModalTensor(p) triangles of degree-p mapping, an M x M periodic
mesh of [0,1]^2 warped by ChanWarping(1/16), exact metrics with the projected
Jacobian, FluxDifferencingForm with Lax-Friedrichs interface dissipation; it times
``semi_discrete_residual!`` on the initial state and reports the minimum and the
median over samples.

Samples here are device times of one residual (HIP events on the handle's stream,
``sse_profile_residual``), where the reference's ``@benchmark`` takes host wall
times. With a ``path`` the minima and medians of successive calls are appended to
``minimum.npy`` / ``median.npy`` in that directory (the reference means to keep
these vectors in ``minimum.jld2`` / ``median.jld2``; its append branch saves the
scalar, ``:171-182``).
"""

from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional

import numpy as np

from .conservation_laws import EulerEquations, IsentropicVortex, LaxFriedrichsNumericalFlux
from .geometry import make_spatial_discretization
from .mesh import ChanWarping, uniform_periodic_mesh, warp_mesh
from .reference_approximation import ModalTensor, ReferenceApproximationTri
from .solvers import FluxDifferencingForm, ReferenceOperator, semidiscretize


def scaling_test_euler_2d_problem(p: int, M: int):
    """The ODE problem of ``scaling_test_euler_2d`` (``benchmark.jl:120-159``)."""
    mach, angle, L, gamma = 0.4, 0.0, 1.0, 1.4
    T = L / mach
    strength = math.sqrt(2.0 / (gamma - 1.0) * (1.0 - 0.75 ** (gamma - 1.0)))
    law = EulerEquations(2, gamma)
    exact = IsentropicVortex(gamma=gamma, Ma=mach, theta=angle, R=1.0 / 10.0, beta=strength,
                             x_0=(L / 2, L / 2))
    form = FluxDifferencingForm(inviscid_numerical_flux=LaxFriedrichsNumericalFlux())
    ref = ReferenceApproximationTri(ModalTensor(p), mapping_degree=p)
    mesh = warp_mesh(uniform_periodic_mesh(ref, ((0.0, L), (0.0, L)), (M, M)), ref, ChanWarping(1.0 / 16.0, (L, L)))
    sd = make_spatial_discretization(mesh, ref)  # exact metrics, projected Jacobian
    return semidiscretize(law, sd, exact, form, (0.0, T), ReferenceOperator())


def scaling_test_euler_2d(p: int, M: int, path: Optional[str] = None, samples: int = 200, verbose: bool = True):
    """Returns {"min_ms", "median_ms", "N_e", "DOF", "kernels"} for one residual of the
    benchmark problem; appends to ``path``/minimum.npy and median.npy when given."""
    import torch

    from . import abi
    from .hip import check
    from .time_integration import to_device
    if verbose:
        print("building the problem...")
    ode = scaling_test_euler_2d_problem(p, M)
    solver = ode.p
    u = to_device(ode.u0)
    dudt = torch.empty_like(u)
    lib, h = solver.lib, solver.handle
    ms = (C.c_double * abi.SSE_NUM_KERNELS)()
    check(lib, h, lib.sse_profile_residual(h, u.data_ptr(), dudt.data_ptr(), 3, ms))  # warm-up
    times = []
    for _ in range(samples):
        check(lib, h, lib.sse_profile_residual(h, u.data_ptr(), dudt.data_ptr(), 1, ms))
        times.append(sum(ms))
    min_time, med_time = float(np.min(times)), float(np.median(times))
    if verbose:
        print(f"min = {min_time:.4f} ms  med = {med_time:.4f} ms")
    if path:
        os.makedirs(path, exist_ok=True)
        for name, value in (("minimum.npy", min_time), ("median.npy", med_time)):
            f = os.path.join(path, name)
            old = np.load(f) if os.path.isfile(f) else np.zeros(0)
            np.save(f, np.append(old, value))
    N_p, N_c, N_e = solver.size()
    return {"min_ms": min_time, "median_ms": med_time, "N_e": N_e, "DOF": N_p * N_c * N_e,
            "kernels": solver.describe()}
