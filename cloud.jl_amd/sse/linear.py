"""``LinearResidual`` (src/Solvers/linear.jl:2-18): the semi-discrete residual of a
linear conservation law as a linear operator on the flattened state, applied on
the GPU through ``semi_discrete_residual``."""

from __future__ import annotations

import numpy as np

from .conservation_laws import LinearAdvectionDiffusionEquation, LinearAdvectionEquation
from .solvers import Solver, semi_discrete_residual


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
