"""ctypes mirror of ``include/sse_hip.h`` (the library's C-ABI)."""

from __future__ import annotations

import ctypes as C

SSE_ABI_VERSION = 3

SSE_OK = 0
SSE_ERR_INVALID_ARGUMENT = 1
SSE_ERR_UNSUPPORTED = 2
SSE_ERR_HIP = 3
SSE_ERR_NONPHYSICAL_STATE = 4
SSE_ERR_OUT_OF_MEMORY = 5

SSE_LINE, SSE_TRI, SSE_TET, SSE_HEX = 1, 2, 3, 5
SSE_EULER, SSE_LINEAR_ADVECTION, SSE_LINEAR_ADVECTION_DIFFUSION = 1, 2, 3
SSE_FLUX_DIFFERENCING, SSE_STANDARD = 1, 2
SSE_LAX_FRIEDRICHS, SSE_CENTRAL, SSE_ENTROPY_CONSERVATIVE_NUMERICAL_FLUX = 1, 2, 3
SSE_SKEW_SYMMETRIC_MAPPING, SSE_STANDARD_MAPPING = 1, 2
SSE_NO_VISCOUS_FLUX, SSE_BR1 = 0, 1
SSE_TWO_POINT_EC, SSE_TWO_POINT_CONSERVATIVE = 1, 2
SSE_MASS_WEIGHT_ADJUSTED, SSE_MASS_DIAGONAL = 1, 2
SSE_C_NONE, SSE_C_DENSE, SSE_C_SPARSE = 0, 1, 2
SSE_DEVICE, SSE_HOST = 0, 1
SSE_NUM_KERNELS = 2
KERNEL_NAMES = ("nodal_values", "time_derivative")

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_int64)


class SseDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("element", C.c_int32), ("d", C.c_int32),
        ("N_p", C.c_int32), ("N_q", C.c_int32), ("N_f", C.c_int32), ("N_fac", C.c_int32),
        ("N_c", C.c_int32), ("N_e", C.c_int32), ("law", C.c_int32), ("form", C.c_int32),
        ("inviscid_flux", C.c_int32), ("two_point_flux", C.c_int32), ("mass_solver", C.c_int32),
        ("S_sparse", C.c_int32), ("C_kind", C.c_int32), ("mapping_form", C.c_int32),
        ("viscous_flux", C.c_int32),
        ("gamma", C.c_double), ("half_lambda", C.c_double), ("a", C.c_double * 3), ("b", C.c_double),
        ("V", _dp), ("R", _dp), ("R_selection", _lp), ("W", _dp), ("B", _dp), ("S", _dp),
        ("C", _dp), ("Minv", _dp), ("D", _dp),
        ("V_A", _dp), ("V_B", _dp), ("V_C", _dp), ("V_p", C.c_int32), ("V_M", C.c_int32 * 3),
        ("J_q", _dp), ("Lambda_q", _dp), ("nJf", _dp), ("J_f", _dp), ("nJq", _dp),
        ("face_normals", _dp), ("mapP", _lp), ("deterministic", C.c_int32),
    ]


# Names of every entry point declared in include/sse_hip.h.
EXPORTED_SYMBOLS = (
    "sse_create", "sse_residual", "sse_rk_steps", "sse_profile_residual", "sse_set_stream", "sse_set_diffusion",
    "sse_synchronize", "sse_last_error", "sse_destroy", "sse_abi_version", "sse_describe",
    "sse_batch_create", "sse_batch_rk_steps", "sse_batch_describe", "sse_batch_last_error",
    "sse_batch_destroy", "sse_functional", "sse_residual_vjp", "sse_rk_adjoint", "sse_batch_rk_tangent",
    "sse_geometric_factors", "sse_apply_reference_mapping", "sse_project_l2", "sse_l2_error",
    "sse_functional_noalloc", "sse_erk_steps", "sse_ec_flux_batch",
)

SSE_ERK_MAX_STAGES = 16

# Workspace of sse_functional_noalloc, in doubles (SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES).
SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES = 1 << 20

SSE_METRICS_EXACT, SSE_METRICS_CURL = 1, 2


class SseGeometryOps(C.Structure):
    _fields_ = [
        ("d", C.c_int32), ("metrics", C.c_int32), ("N_p", C.c_int32), ("N_q", C.c_int32),
        ("N_f", C.c_int32), ("N_fac", C.c_int32),
        ("D", _dp), ("Vq", _dp), ("Vf", _dp), ("nrstJ", _dp), ("P_J", _dp),
        ("N_p1", C.c_int32), ("I1", _dp), ("D1", _dp), ("Vq1", _dp), ("Vf1", _dp),
    ]
SSE_MASS, SSE_ENERGY, SSE_ENTROPY = 0, 1, 2
