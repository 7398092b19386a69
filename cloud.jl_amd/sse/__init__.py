"""MI355X-native semi-discrete residual for StableSpectralElements.jl.

Host-side mirror of the reference's API (ConservationLaws, SpatialDiscretizations,
Solvers): discretisation set-up in NumPy, the residual itself in libsse_hip.so
(HIP kernels for gfx950) behind the C-ABI declared in include/sse_hip.h.
"""

from . import abi, conservation_laws, geometry, mesh, reference_approximation, solvers
from .conservation_laws import *  # noqa: F401,F403
from .geometry import (ChanWilcoxMetrics, ConservativeCurlMetrics, ExactMetrics,  # noqa: F401
                       SpatialDiscretization, check_facet_nodes, check_normals,
                       make_spatial_discretization)
from .mesh import (ChanWarping, DelReyWarping, UniformWarping, uniform_periodic_mesh,  # noqa: F401
                   warp_mesh)
from .reference_approximation import (GaussQuadrature, LGLQuadrature, LGQuadrature,  # noqa: F401
                                      LGRQuadrature, ModalMulti, ModalTensor, NodalMulti,
                                      NodalTensor,
                                      ReferenceApproximationFor, ReferenceApproximationLine,
                                      ReferenceApproximationTet, ReferenceApproximationTri,
                                      check_sbp_property)
from .solvers import (MI355X, DiagonalSolver, FluxDifferencingForm, ODEProblem,  # noqa: F401
                      ReferenceOperator, Solver, StandardForm, WeightAdjustedSolver,
                      semi_discrete_residual, semidiscretize)
