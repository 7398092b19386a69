"""Discretisations shared by the tests (the reference's test scenarios and the
benchmark workloads, at sizes the oracle finishes in seconds)."""

import numpy as np

from sse.workloads import (density_wave_3d_config, euler_1d_gauss,  # noqa: F401
                           euler_vortex_2d_modal, vortex_2d_config)


def l2_error(sd, u, exact, t=0.0):
    """ErrorAnalysis with the volume quadrature (src/Analysis/error.jl:144-222)."""
    ref = sd.reference_approximation
    uq = np.einsum("qp,pck->qck", ref.Vdense(), u)
    ue = np.transpose(np.asarray(exact(sd.mesh.xyzq, t)), (0, 2, 1))
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    return np.sqrt(np.einsum("ik,ick->c", WJ, (ue - uq) ** 2))
