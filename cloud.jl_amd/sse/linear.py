"""``LinearResidual`` (src/Solvers/linear.jl:2-18): the semi-discrete residual of a
linear conservation law as a linear operator on the flattened state, applied on
the GPU through ``semi_discrete_residual``."""

from __future__ import annotations

import numpy as np

from .conservation_laws import LinearAdvectionDiffusionEquation, LinearAdvectionEquation
from .solvers import SkewSymmetricMapping, Solver, StandardForm, semi_discrete_residual


class LinearResidual:
    """y = R(x) for x, y of length N_p * N_c * N_e in the reference's column-major
    (N_p, N_c, N_e) order; like the reference's ``LinearMap`` it is square and its
    ``@``/``matvec`` runs one residual evaluation."""

    def __init__(self, solver: Solver):
        if not isinstance(solver.conservation_law, (LinearAdvectionEquation, LinearAdvectionDiffusionEquation)):
            raise TypeError("LinearResidual needs a linear conservation law")
        self.solver = solver
        n = int(np.prod(solver.size()))
        self.shape = (n, n)

    def matvec(self, x) -> np.ndarray:
        shape = self.solver.size()
        u = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(shape, order="F"))
        y = np.zeros(shape, order="F")
        semi_discrete_residual(y, u, self.solver, 0.0)
        return y.reshape(-1, order="F")

    def __matmul__(self, x):
        return self.matvec(x)

    def adjoint(self) -> "LinearResidual":
        """The adjoint in the mass-matrix inner product, M^{-1} R' M, as a residual. For
        the skew-symmetric StandardForm of linear advection(-diffusion) it is the
        residual with the velocity reversed: advection is skew in M and the LF and BR1
        terms are symmetric, so <v, R w>_M = <R_{-a} v, w>_M. This is the operator of
        the discrete adjoint in ``adjoint.energy_adjoint_b``."""
        form = self.solver.form
        if not (isinstance(form, StandardForm) and isinstance(form.mapping_form, SkewSymmetricMapping)):
            raise NotImplementedError("the velocity-reversal adjoint needs the skew-symmetric StandardForm")
        law = self.solver.conservation_law
        a = tuple(-x for x in law.a)
        rev = (LinearAdvectionDiffusionEquation(a, law.b) if isinstance(law, LinearAdvectionDiffusionEquation)
               else LinearAdvectionEquation(a))
        return LinearResidual(self.solver.with_law(rev))
