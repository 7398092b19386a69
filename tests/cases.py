"""Discretisations shared by the tests (the reference's test scenarios and the
benchmark workloads, at sizes the oracle finishes in seconds)."""

import numpy as np

from sse.workloads import (advection_2d_modal_tensor_tri,  # noqa: F401
                           advection_3d_modal_tensor_tet, advection_diffusion_1d_modal_multi,
                           advection_diffusion_3d_config, density_wave_3d_config, euler_1d_gauss,
                           euler_vortex_2d_modal, vortex_2d_config)


def l2_error(sd, u, exact, t=0.0):
    """ErrorAnalysis with the volume quadrature (src/Analysis/error.jl:58-91)."""
    ref = sd.reference_approximation
    uq = np.einsum("qp,pck->qck", ref.Vdense(), u)
    ue = np.transpose(np.asarray(exact(sd.mesh.xyzq, t)), (0, 2, 1))
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    return np.sqrt(np.einsum("ik,ick->c", WJ, (ue - uq) ** 2))


def energy_rate(sd, u, dudt):
    """EnergyConservationAnalysis residual sum_k u' M_k dudt with the weight-adjusted
    mass matrix M_k = (V' diag(W / J_k) V)^{-1} (src/Analysis/conservation.jl:158-167,
    src/Solvers/mass_matrix.jl:150-153); returns (rate, scale)."""
    ref = sd.reference_approximation
    V = ref.Vdense()
    Jq = sd.geometric_factors.J_q
    rate = scale = 0.0
    for k in range(sd.N_e):
        Minv = V.T @ ((ref.W / Jq[:, k])[:, None] * V)
        for e in range(u.shape[1]):
            y = np.linalg.solve(Minv, dudt[:, e, k])
            rate += u[:, e, k] @ y
            scale += np.abs(u[:, e, k]) @ np.abs(y)
    return rate, scale
