"""One-dimensional orthogonal polynomials and Gauss-type quadrature rules.

The reference obtains these from two third-party Julia packages that are not
vendored in the reference tree:

* Jacobi.jl (compat "0.6.0"): ``zgj/wgj`` (Gauss-Jacobi), ``zglj/wglj``
  (Gauss-Lobatto-Jacobi), ``zgrjm/wgrjm`` (Gauss-Radau-Jacobi with the node at
  x = -1), called from ``src/SpatialDiscretizations/quadrature_rules.jl:76-92``.
* NodesAndModes.jl / StartUpDG.jl (compat "1"): ``jacobiP`` (orthonormal Jacobi
  polynomials) and the 1D Legendre ``vandermonde``/``grad_vandermonde``
  (``src/SpatialDiscretizations/tensor_simplex.jl:97-98, 142-156``).

This module restates the published algorithms: the normalised three-term
recurrence of Hesthaven & Warburton, *Nodal Discontinuous Galerkin Methods*
(2008), Appendix A (JacobiP / GradJacobiP), and the Golub-Welsch eigenvalue
method for Gauss-Jacobi nodes, refined by Newton iteration on the recurrence.
Weights use the Christoffel formula w_i = 1 / sum_n p_n(x_i)^2 for Gauss rules
and exact moment matching for the Lobatto/Radau rules. Weight function is
(1 - x)^a (1 + x)^b, matching Jacobi.jl.
"""

from __future__ import annotations

import math

import numpy as np


def _gamma0(a: float, b: float) -> float:
    return (2.0 ** (a + b + 1.0) / (a + b + 1.0) * math.gamma(a + 1.0)
            * math.gamma(b + 1.0) / math.gamma(a + b + 1.0))


def jacobi_p(x, a: float, b: float, n: int) -> np.ndarray:
    """Orthonormal Jacobi polynomial P_n^{(a,b)}(x), normalised so that
    int_{-1}^{1} P_n P_m (1-x)^a (1+x)^b dx = delta_nm (NodesAndModes ``jacobiP``)."""
    x = np.asarray(x, dtype=np.float64)
    pl = np.empty((n + 1,) + x.shape)
    g0 = _gamma0(a, b)
    pl[0] = 1.0 / math.sqrt(g0)
    if n == 0:
        return pl[0].copy()
    g1 = (a + 1.0) * (b + 1.0) / (a + b + 3.0) * g0
    pl[1] = ((a + b + 2.0) * x / 2.0 + (a - b) / 2.0) / math.sqrt(g1)
    if n == 1:
        return pl[1].copy()
    aold = 2.0 / (2.0 + a + b) * math.sqrt((a + 1.0) * (b + 1.0) / (a + b + 3.0))
    for i in range(1, n):
        h1 = 2.0 * i + a + b
        anew = 2.0 / (h1 + 2.0) * math.sqrt((i + 1.0) * (i + 1.0 + a + b) * (i + 1.0 + a)
                                            * (i + 1.0 + b) / (h1 + 1.0) / (h1 + 3.0))
        bnew = -(a * a - b * b) / h1 / (h1 + 2.0)
        pl[i + 1] = (-aold * pl[i - 1] + (x - bnew) * pl[i]) / anew
        aold = anew
    return pl[n].copy()


def grad_jacobi_p(x, a: float, b: float, n: int) -> np.ndarray:
    """Derivative of the orthonormal Jacobi polynomial."""
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.zeros_like(x)
    return math.sqrt(n * (n + a + b + 1.0)) * jacobi_p(x, a + 1.0, b + 1.0, n - 1)


def _gauss_jacobi_raw(q: int, a: float, b: float) -> np.ndarray:
    """Golub-Welsch nodes of P_q^{(a,b)}, refined by Newton, ascending."""
    if q == 0:
        return np.zeros(0)
    n = np.arange(q, dtype=np.float64)
    ab = a + b
    diag = np.empty(q)
    for i in range(q):
        if i == 0 and abs(ab) < 1e-14:
            diag[i] = (b - a) / (ab + 2.0)
        else:
            diag[i] = (b * b - a * a) / ((2 * i + ab) * (2 * i + ab + 2.0))
    off = np.empty(max(q - 1, 0))
    for i in range(q - 1):
        off[i] = 2.0 / (2 * i + ab + 2.0) * math.sqrt(
            (i + 1.0) * (i + 1.0 + ab) * (i + 1.0 + a) * (i + 1.0 + b)
            / ((2 * i + ab + 1.0) * (2 * i + ab + 3.0)))
    T = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    x = np.sort(np.linalg.eigvalsh(T))
    del n
    for _ in range(3):  # Newton refinement on the orthonormal polynomial
        f = jacobi_p(x, a, b, q)
        df = grad_jacobi_p(x, a, b, q)
        x = x - f / df
    return np.sort(x)


def _interpolatory_weights(x: np.ndarray, a: float, b: float) -> np.ndarray:
    """Weights making the rule exact for all polynomials of degree < len(x)."""
    q = len(x)
    V = np.stack([jacobi_p(x, a, b, n) for n in range(q)], axis=1)
    m = np.zeros(q)
    m[0] = math.sqrt(_gamma0(a, b))
    return np.linalg.solve(V.T, m)


def gauss_jacobi(q: int, a: float = 0.0, b: float = 0.0):
    """q-point Gauss-Jacobi rule (Jacobi.jl ``zgj``/``wgj``)."""
    x = _gauss_jacobi_raw(q, a, b)
    s = np.zeros_like(x)
    for n in range(q):
        s += jacobi_p(x, a, b, n) ** 2
    return x, 1.0 / s


def gauss_lobatto_jacobi(q: int, a: float = 0.0, b: float = 0.0):
    """q-point Gauss-Lobatto-Jacobi rule (Jacobi.jl ``zglj``/``wglj``):
    nodes -1, 1 and the roots of P_{q-2}^{(a+1,b+1)}."""
    if q < 2:
        raise ValueError("Gauss-Lobatto needs at least 2 points")
    x = np.concatenate(([-1.0], _gauss_jacobi_raw(q - 2, a + 1.0, b + 1.0), [1.0]))
    return x, _interpolatory_weights(x, a, b)


def gauss_radau_jacobi_m(q: int, a: float = 0.0, b: float = 0.0):
    """q-point Gauss-Radau-Jacobi rule with a node at x = -1 (Jacobi.jl
    ``zgrjm``/``wgrjm``): nodes -1 and the roots of P_{q-1}^{(a,b+1)}."""
    x = np.concatenate(([-1.0], _gauss_jacobi_raw(q - 1, a, b + 1.0)))
    return x, _interpolatory_weights(x, a, b)


def gauss_lobatto_legendre_nodes(n: int) -> np.ndarray:
    """n+1 Legendre-Gauss-Lobatto nodes (Hesthaven-Warburton JacobiGL(0,0,n))."""
    if n == 1:
        return np.array([-1.0, 1.0])
    return np.concatenate(([-1.0], _gauss_jacobi_raw(n - 1, 1.0, 1.0), [1.0]))


def vandermonde_1d(n: int, r) -> np.ndarray:
    """Legendre (orthonormal) Vandermonde V[i, j] = P_j(r_i), j = 0..n."""
    r = np.asarray(r, dtype=np.float64)
    return np.stack([jacobi_p(r, 0.0, 0.0, j) for j in range(n + 1)], axis=1)


def grad_vandermonde_1d(n: int, r) -> np.ndarray:
    r = np.asarray(r, dtype=np.float64)
    return np.stack([grad_jacobi_p(r, 0.0, 0.0, j) for j in range(n + 1)], axis=1)
