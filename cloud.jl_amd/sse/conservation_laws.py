"""Conservation laws, numerical-flux types and initial data.

Mirrors ``src/ConservationLaws/ConservationLaws.jl`` (flux type hierarchy,
``LaxFriedrichsNumericalFlux`` storing halfλ = λ/2),
``src/ConservationLaws/euler_navierstokes.jl`` (``EulerEquations``, entropy
variables, ``IsentropicVortex``, ``EulerPeriodicTest``) and
``src/ConservationLaws/linear_advection_diffusion.jl``. The per-node physics
evaluated inside the residual lives in the HIP kernels; the NumPy functions
here serve initial conditions and post-processing functionals only.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Tuple

import numpy as np


# ----------------------------------------------------------------------------- laws
@dataclass
class EulerEquations:
    """``EulerEquations{d}(γ)`` (``euler_navierstokes.jl:232-244``)."""
    d: int
    gamma: float = 1.4

    @property
    def N_c(self) -> int:
        return self.d + 2

    @property
    def second_order(self) -> bool:
        return False


@dataclass
class LinearAdvectionEquation:
    """``LinearAdvectionEquation(a)`` (``linear_advection_diffusion.jl:1-19``)."""
    a: Tuple[float, ...]

    @property
    def d(self) -> int:
        return len(self.a)

    @property
    def N_c(self) -> int:
        return 1

    @property
    def second_order(self) -> bool:
        return False


@dataclass
class LinearAdvectionDiffusionEquation:
    """``LinearAdvectionDiffusionEquation(a, b)`` (``linear_advection_diffusion.jl:21-51``)."""
    a: Tuple[float, ...]
    b: float

    @property
    def d(self) -> int:
        return len(self.a)

    @property
    def N_c(self) -> int:
        return 1

    @property
    def second_order(self) -> bool:
        return True


# ----------------------------------------------------------------------------- fluxes
@dataclass
class LaxFriedrichsNumericalFlux:
    """Stores halfλ = λ/2; default λ = 1 (``ConservationLaws.jl:54-62``)."""
    lam: float = 1.0

    @property
    def half_lambda(self) -> float:
        return 0.5 * self.lam


class CentralNumericalFlux:
    pass


class EntropyConservativeNumericalFlux:
    pass


class BR1:
    pass


class EntropyConservativeFlux:
    pass


class ConservativeFlux:
    pass


# ----------------------------------------------------------------------------- physics (host, NumPy)
def conservative_to_entropy(law: EulerEquations, u: np.ndarray) -> np.ndarray:
    """(``euler_navierstokes.jl:100-113``); u has shape (..., N_c)."""
    g = law.gamma
    d = law.d
    k = (0.5 / u[..., 0]) * (u[..., 1:d + 1] ** 2).sum(-1)
    p = (g - 1.0) * (u[..., d + 1] - k)
    inv_p = 1.0 / p
    w = np.empty_like(u)
    w[..., 0] = (g - np.log(p / u[..., 0] ** g)) / (g - 1.0) - k * inv_p
    w[..., 1:d + 1] = u[..., 1:d + 1] * inv_p[..., None]
    w[..., d + 1] = -u[..., 0] * inv_p
    return w


def entropy(law, u: np.ndarray) -> np.ndarray:
    """Mathematical entropy S(u) (``euler_navierstokes.jl:93-98``;
    ``linear_advection_diffusion.jl:104``)."""
    if isinstance(law, EulerEquations):
        g = law.gamma
        d = law.d
        p = (g - 1.0) * (u[..., d + 1] - (0.5 / u[..., 0]) * (u[..., 1:d + 1] ** 2).sum(-1))
        return -u[..., 0] * np.log(p / u[..., 0] ** g) / (g - 1.0)
    return 0.5 * u[..., 0] ** 2


# ----------------------------------------------------------------------------- initial data
@dataclass
class IsentropicVortex:
    """``IsentropicVortex`` (``euler_navierstokes.jl:232-274``), p = ρ^γ / γ."""
    gamma: float = 1.4
    Ma: float = 0.4
    theta: float = math.pi / 4
    R: float = 1.0
    beta: float = 1.0
    sigma: float = 1.0
    x_0: Tuple[float, float] = (0.0, 0.0)

    def __call__(self, x, t=0.0):
        g = self.gamma
        xr = ((x[0] - self.x_0[0]) / self.R, (x[1] - self.x_0[1]) / self.R)
        om = self.beta * np.exp(-0.5 / self.sigma ** 2 * (xr[0] ** 2 + xr[1] ** 2))
        dv = (-xr[1] * om, xr[0] * om)
        dT = -0.5 * (g - 1.0) * om ** 2
        rho = (1.0 + dT) ** (1.0 / (g - 1.0))
        v = (self.Ma * math.cos(self.theta) + dv[0], self.Ma * math.sin(self.theta) + dv[1])
        p = rho ** g / g
        E = p / (g - 1.0) + 0.5 * rho * (v[0] ** 2 + v[1] ** 2)
        return np.stack([rho, rho * v[0], rho * v[1], E], axis=-1)


@dataclass
class ShuVortex:
    """Isentropic vortex of Shu (1998) on a free stream (ρ, u, v, p) = (1, u∞, v∞, 1):
    δu = -ε/(2π) e^{(1-r²)/2} (y-y0), δv = ε/(2π) e^{(1-r²)/2} (x-x0),
    T = 1 - (γ-1) ε² e^{1-r²} / (8 γ π²), ρ = T^{1/(γ-1)}, p = ρ^γ.
    Used for the 2D benchmark vortex (strength 5 centred in [0,10]^2), which the
    reference's ``IsentropicVortex`` (p = ρ^γ/γ) cannot express; the reference
    accepts such plain functions through ``evaluate(::Function, ...)``
    (``src/GridFunctions/GridFunctions.jl:180-201``)."""
    gamma: float = 1.4
    strength: float = 5.0
    x_0: Tuple[float, float] = (5.0, 5.0)
    u_inf: Tuple[float, float] = (1.0, 1.0)
    period: Tuple[float, float] = (10.0, 10.0)

    def __call__(self, x, t=0.0):
        g = self.gamma
        eps = self.strength
        # vortex centre advected with the free stream, wrapped to the periodic box
        dx = x[0] - (self.x_0[0] + self.u_inf[0] * t)
        dy = x[1] - (self.x_0[1] + self.u_inf[1] * t)
        dx = dx - self.period[0] * np.round(dx / self.period[0])
        dy = dy - self.period[1] * np.round(dy / self.period[1])
        r2 = dx ** 2 + dy ** 2
        f = eps / (2 * math.pi) * np.exp(0.5 * (1.0 - r2))
        u = self.u_inf[0] - f * dy
        v = self.u_inf[1] + f * dx
        T = 1.0 - (g - 1.0) * eps ** 2 * np.exp(1.0 - r2) / (8.0 * g * math.pi ** 2)
        rho = T ** (1.0 / (g - 1.0))
        p = rho ** g
        E = p / (g - 1.0) + 0.5 * rho * (u ** 2 + v ** 2)
        return np.stack([rho, rho * u, rho * v, E], axis=-1)


@dataclass
class EulerPeriodicTest:
    """ρ = 1 + strength sin(2π Σx / L), unit velocity, p = 1
    (``euler_navierstokes.jl:276-300``)."""
    d: int
    gamma: float = 1.4
    strength: float = 0.2
    L: float = 2.0

    def __call__(self, x, t=0.0):
        s = sum(x[m] for m in range(self.d))
        rho = 1.0 + self.strength * np.sin(2 * math.pi * s / self.L)
        comps = [rho] + [rho] * self.d + [1.0 / (self.gamma - 1.0) + 0.5 * rho * self.d]
        return np.stack(comps, axis=-1)


@dataclass
class InitialDataGaussian:
    """``InitialDataGaussian(A, σ, x₀)`` (``src/GridFunctions/GridFunctions.jl:50-58, 136-142``)."""
    A: float
    sigma: float
    x_0: Tuple[float, ...]

    def __call__(self, x, t=0.0):
        r2 = sum((x[m] - self.x_0[m]) ** 2 for m in range(len(self.x_0)))
        return (self.A * np.exp(-r2 / (2.0 * self.sigma ** 2)))[..., None]


@dataclass
class InitialDataSine:
    """``InitialDataSine(A, k)`` (``src/GridFunctions/GridFunctions.jl:30-38, 124-128``)."""
    A: float
    k: Tuple[float, ...]

    def __call__(self, x, t=0.0):
        return (self.A * np.prod([np.sin(self.k[m] * x[m]) for m in range(len(self.k))], axis=0))[..., None]


@dataclass
class InitialDataCosine:
    """``InitialDataCosine(A, k)`` (``src/GridFunctions/GridFunctions.jl:40-48, 130-134``)."""
    A: float
    k: Tuple[float, ...]

    def __call__(self, x, t=0.0):
        return (self.A * np.prod([np.cos(self.k[m] * x[m]) for m in range(len(self.k))], axis=0))[..., None]


@dataclass
class ExactSolution:
    """``ExactSolution(conservation_law, initial_data)`` for the linear laws
    (``src/ConservationLaws/linear_advection_diffusion.jl:130-182``): the initial
    data translated by a t, times exp(-b |k|^2 t) for a sine under diffusion, or
    the spreading Gaussian. Not periodic-wrapped, like the reference's default."""
    conservation_law: object
    initial_data: object

    def __call__(self, x, t=0.0):
        law, f = self.conservation_law, self.initial_data
        a = law.a
        z = tuple(x[m] - a[m] * t for m in range(len(a)))
        b = getattr(law, "b", 0.0)
        if isinstance(f, InitialDataGaussian) and b > 0.0:
            d = len(a)
            r2 = sum((z[m] - f.x_0[m]) ** 2 for m in range(d))
            t0 = f.sigma ** 2 / (2.0 * b)
            return (f.A * (t0 / (t + t0)) ** (0.5 * d) * np.exp(-r2 / (4.0 * b * (t0 + t))))[..., None]
        out = f(z)
        if isinstance(f, InitialDataSine) and b > 0.0:
            out = out * np.exp(-b * sum(km ** 2 for km in f.k) * t)
        return out
