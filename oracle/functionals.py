"""NumPy restatement of the reference's conservation and invariant functionals.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check libsse_hip's
``sse_functional``; the product never imports it.

Follows src/Analysis/conservation.jl:
  evaluate_conservation (Primary)          :113-117  sum_k 1' WJ_k V u_k
  evaluate_conservation (Entropy)          :119-129  sum_k sum_i WJ_k,i S((V u_k)_i)
  evaluate_conservation (Energy)           :131-143  1/2 sum_k u_k' M_k u_k
  evaluate_conservation_residual (Primary) :145-152
  evaluate_conservation_residual (Energy)  :154-167
  evaluate_conservation_residual (Entropy) :169-189  (P w_q)' M_k dudt, P = M_k^{-1} V' WJ_k
with M_k of the weight-adjusted solver, inv(V' diag(W / J_k) V) for M^{-1} = I
(src/Solvers/mass_matrix.jl:150-153), or W J_k for the diagonal solver.
"""

from __future__ import annotations

import numpy as np


def _parts(sd):
    ref = sd.reference_approximation
    V = ref.Vdense() if ref.V is not None else np.eye(ref.N_q)
    return ref, V, sd.geometric_factors.J_q


def mass_matrix(sd, k):
    ref, V, Jq = _parts(sd)
    if ref.V is None:
        return np.diag(ref.W * Jq[:, k])
    return np.linalg.inv(V.T @ ((ref.W / Jq[:, k])[:, None] * V))


def mass(sd, u):
    ref, V, Jq = _parts(sd)
    return np.einsum("qk,qp,pck->c", ref.W[:, None] * Jq, V, u)


def energy(sd, u):
    return np.array([sum(0.5 * u[:, e, k] @ mass_matrix(sd, k) @ u[:, e, k] for k in range(sd.N_e))
                     for e in range(u.shape[1])])


def _entropy_pointwise(law, uq):
    """S(u) (euler_navierstokes.jl:93-98; linear_advection_diffusion.jl:104)."""
    if hasattr(law, "gamma"):
        g, d = law.gamma, law.d
        p = (g - 1.0) * (uq[..., d + 1] - 0.5 / uq[..., 0] * (uq[..., 1:d + 1] ** 2).sum(-1))
        return -uq[..., 0] * np.log(p / uq[..., 0] ** g) / (g - 1.0)
    return 0.5 * uq[..., 0] ** 2


def _entropy_variables(law, uq):
    if hasattr(law, "gamma"):
        g, d = law.gamma, law.d
        k = 0.5 / uq[..., 0] * (uq[..., 1:d + 1] ** 2).sum(-1)
        p = (g - 1.0) * (uq[..., d + 1] - k)
        w = np.empty_like(uq)
        w[..., 0] = (g - np.log(p / uq[..., 0] ** g)) / (g - 1.0) - k / p
        w[..., 1:d + 1] = uq[..., 1:d + 1] / p[..., None]
        w[..., d + 1] = -uq[..., 0] / p
        return w
    return uq.copy()


def entropy(sd, law, u):
    ref, V, Jq = _parts(sd)
    uq = np.einsum("qp,pck->qkc", V, u)
    return np.array([np.sum(ref.W[:, None] * Jq * _entropy_pointwise(law, uq))])


def mass_rate(sd, u, dudt):
    return mass(sd, dudt)


def energy_rate(sd, u, dudt):
    return np.array([sum(u[:, e, k] @ mass_matrix(sd, k) @ dudt[:, e, k] for k in range(sd.N_e))
                     for e in range(u.shape[1])])


def entropy_rate(sd, law, u, dudt):
    ref, V, Jq = _parts(sd)
    out = 0.0
    for k in range(sd.N_e):
        M = mass_matrix(sd, k)
        WJ = ref.W * Jq[:, k]
        wq = _entropy_variables(law, V @ u[:, :, k])
        P = np.linalg.solve(M, V.T * WJ[None, :])  # M_k^{-1} V' WJ_k
        out += sum((P @ wq[:, e]) @ M @ dudt[:, e, k] for e in range(u.shape[1]))
    return np.array([out])
