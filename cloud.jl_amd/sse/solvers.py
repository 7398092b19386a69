"""Residual forms, ``Solver``, ``semidiscretize`` and ``semi_discrete_residual``.

Mirrors ``src/Solvers/Solvers.jl`` (forms ``:90-115``, ``Solver`` ``:259-272``,
``project_function``/``initialize``/``semidiscretize`` ``:388-452``,
``semi_discrete_residual!`` ``:474-564``), ``src/Solvers/operators.jl:39-81``
(``FluxDifferencingOperators``) and ``src/Solvers/mass_matrix.jl:19-24``
(default weight-adjusted mass solver).

The reference dispatches the residual on ``Solver``'s parallelism type
(``Serial``/``Threaded``). Here the parallelism is ``MI355X()``: the residual
runs in libsse_hip.so on the GPU, reached through its C-ABI
(``include/sse_hip.h``). There is no CPU residual in this package.
"""

from __future__ import annotations

import copy
import ctypes as C
from dataclasses import dataclass, field
from typing import Any, Optional, Tuple

import numpy as np

from . import abi
from .conservation_laws import (BR1, CentralNumericalFlux, ConservativeFlux,
                                EntropyConservativeFlux, EntropyConservativeNumericalFlux,
                                EulerEquations, LaxFriedrichsNumericalFlux,
                                LinearAdvectionDiffusionEquation, LinearAdvectionEquation)
from .geometry import SpatialDiscretization, apply_reference_mapping
from .reference_approximation import (HEX, LINE, TET, TRI, face_reference_normals,
                                      flux_differencing_operators, num_faces)


# ----------------------------------------------------------------------------- forms & strategies
class SkewSymmetricMapping:
    pass


class StandardMapping:
    pass


@dataclass
class StandardForm:
    mapping_form: Any = field(default_factory=SkewSymmetricMapping)
    inviscid_numerical_flux: Any = field(default_factory=LaxFriedrichsNumericalFlux)
    viscous_numerical_flux: Any = field(default_factory=BR1)


@dataclass
class FluxDifferencingForm:
    mapping_form: Any = field(default_factory=SkewSymmetricMapping)
    inviscid_numerical_flux: Any = field(default_factory=LaxFriedrichsNumericalFlux)
    viscous_numerical_flux: Any = field(default_factory=BR1)
    two_point_flux: Any = field(default_factory=EntropyConservativeFlux)


class ReferenceOperator:
    pass


class PhysicalOperator:
    pass


class MI355X:
    """Parallelism: the residual runs as HIP kernels on one MI355X (gfx950).

    ``deterministic=True`` makes every sum of the line-structured kernels (ModalTensor
    Tri/Tet Euler: time derivative and VJP) run in a fixed order, so repeated residuals
    are bitwise equal; the default accumulates some sums with LDS atomics (faster)."""

    def __init__(self, deterministic: bool = False):
        self.deterministic = bool(deterministic)


@dataclass
class WeightAdjustedSolver:
    """WADG with M⁻¹ = I for orthonormal bases (``mass_matrix.jl:11-17, 59-75``)."""
    Minv: Optional[np.ndarray] = None


class DiagonalSolver:
    """W J diagonal mass matrix for nodal V = I (``mass_matrix.jl:7-9, 108-115``)."""


def default_mass_matrix_solver(sd: SpatialDiscretization):
    """(``mass_matrix.jl:19-24, 41-48``): diagonal for nodal V = I, else WADG."""
    if sd.reference_approximation.V is None:
        return DiagonalSolver()
    return WeightAdjustedSolver()


# ----------------------------------------------------------------------------- descriptor
_ELEM = {LINE: abi.SSE_LINE, TRI: abi.SSE_TRI, TET: abi.SSE_TET, HEX: abi.SSE_HEX}


def _fptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _lptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class DescArrays:
    """Host arrays behind one ``sse_desc`` (kept alive while the descriptor is used)."""

    def __init__(self, law, sd: SpatialDiscretization, form, mass_solver, strategy=None):
        ref = sd.reference_approximation
        gf = sd.geometric_factors
        d = ref.d
        self.law = law
        self.sd = sd
        self.form = form
        standard = isinstance(form, StandardForm)
        if not (standard or isinstance(form, FluxDifferencingForm)):
            raise NotImplementedError(f"unknown residual form {form!r}")
        f64 = lambda a: None if a is None else np.asfortranarray(a, dtype=np.float64)
        if standard:
            # ReferenceOperators / PhysicalOperators hold the geometric factors after
            # the reference mapping and the reference derivative operators D
            # (operators.jl:1-37, 83-160)
            S, S_sparse, Cmat, C_kind = None, False, None, abi.SSE_C_NONE
            gf = apply_reference_mapping(gf, ref, getattr(sd, "backend", "host"))
            self.D = np.asfortranarray(np.stack([np.asarray(Dm, dtype=np.float64) for Dm in ref.D], axis=2))
        else:
            S, S_sparse, Cmat, C_kind = flux_differencing_operators(ref)
            self.D = None
        self.V = f64(ref.V)
        self.R = f64(ref.R)
        self.R_selection = (None if ref.R_selection is None
                            else np.ascontiguousarray(ref.R_selection + 1, dtype=np.int64))
        self.W = f64(ref.W)
        self.B = f64(ref.B)
        # S[:, :, m] column-major  <->  C-order array (d, N_q, N_q) holding S_m transposed
        self.S = (None if S is None else
                  np.ascontiguousarray(np.stack([np.asfortranarray(S[m]).T for m in range(d)])))
        self.C = f64(Cmat)
        self.Minv = f64(mass_solver.Minv) if isinstance(mass_solver, WeightAdjustedSolver) else None
        wp = ref.warped
        self.V_A = f64(wp.A) if wp is not None else None
        self.V_B = f64(wp.B) if wp is not None else None
        self.V_C = f64(wp.C) if (wp is not None and wp.C is not None) else None
        self.J_q = f64(gf.J_q)
        self.Lambda_q = f64(gf.Lambda_q)
        self.nJf = f64(gf.nJf)
        self.J_f = f64(gf.J_f)
        self.nJq = None if standard else f64(gf.nJq)  # only the facet correction uses nJq
        self.face_normals = f64(face_reference_normals(ref))
        self.mapP = np.asfortranarray(sd.mesh.mapP, dtype=np.int64)
        desc = abi.SseDesc()
        desc.abi_version = abi.SSE_ABI_VERSION
        desc.element = _ELEM[ref.element_type]
        desc.d = d
        desc.N_p, desc.N_q, desc.N_f = ref.N_p, ref.N_q, ref.N_f
        desc.N_fac = num_faces(ref.element_type)
        desc.N_c = law.N_c
        desc.N_e = sd.N_e
        if isinstance(law, EulerEquations):
            desc.law = abi.SSE_EULER
            desc.gamma = law.gamma
        elif isinstance(law, LinearAdvectionDiffusionEquation):
            desc.law = abi.SSE_LINEAR_ADVECTION_DIFFUSION
            desc.b = law.b
        else:
            desc.law = abi.SSE_LINEAR_ADVECTION
        if hasattr(law, "a"):
            for m, am in enumerate(law.a):
                desc.a[m] = am
        desc.form = abi.SSE_STANDARD if standard else abi.SSE_FLUX_DIFFERENCING
        if standard:
            # the ReferenceOperator algorithm is the skew-symmetric one whatever the
            # mapping form (standard_form_first_order.jl:16-63); PhysicalOperators
            # follow the form's mapping (operators.jl:83-160)
            skew = (isinstance(form.mapping_form, SkewSymmetricMapping)
                    or (isinstance(strategy, ReferenceOperator) and not law.second_order))
            desc.mapping_form = abi.SSE_SKEW_SYMMETRIC_MAPPING if skew else abi.SSE_STANDARD_MAPPING
            if law.second_order:
                if not isinstance(form.viscous_numerical_flux, BR1):
                    raise NotImplementedError("second-order laws use the BR1 viscous flux")
                desc.viscous_flux = abi.SSE_BR1
        flux = form.inviscid_numerical_flux
        if isinstance(flux, LaxFriedrichsNumericalFlux):
            desc.inviscid_flux = abi.SSE_LAX_FRIEDRICHS
            desc.half_lambda = flux.half_lambda
        elif isinstance(flux, EntropyConservativeNumericalFlux):
            desc.inviscid_flux = abi.SSE_ENTROPY_CONSERVATIVE_NUMERICAL_FLUX
        else:
            desc.inviscid_flux = abi.SSE_CENTRAL
        # StandardForm's interface flux uses the conservative two-point flux
        # (numerical_flux!'s default, ConservationLaws.jl:75-86)
        desc.two_point_flux = (abi.SSE_TWO_POINT_CONSERVATIVE
                               if standard or isinstance(form.two_point_flux, ConservativeFlux)
                               else abi.SSE_TWO_POINT_EC)
        desc.mass_solver = (abi.SSE_MASS_DIAGONAL if isinstance(mass_solver, DiagonalSolver)
                            else abi.SSE_MASS_WEIGHT_ADJUSTED)
        desc.S_sparse = int(S_sparse)
        desc.C_kind = C_kind
        desc.V, desc.R, desc.R_selection = _fptr(self.V), _fptr(self.R), _lptr(self.R_selection)
        desc.W, desc.B, desc.S, desc.C = _fptr(self.W), _fptr(self.B), _fptr(self.S), _fptr(self.C)
        desc.Minv = _fptr(self.Minv)
        desc.D = _fptr(self.D)
        desc.V_A, desc.V_B, desc.V_C = _fptr(self.V_A), _fptr(self.V_B), _fptr(self.V_C)
        if wp is not None:
            desc.V_p = wp.p
            for m, Mm in enumerate(wp.M):
                desc.V_M[m] = Mm
        desc.J_q, desc.Lambda_q, desc.nJf = _fptr(self.J_q), _fptr(self.Lambda_q), _fptr(self.nJf)
        desc.J_f, desc.nJq = _fptr(self.J_f), _fptr(self.nJq)
        desc.face_normals = _fptr(self.face_normals)
        desc.mapP = _lptr(self.mapP)
        self.desc = desc

    @property
    def shape(self) -> Tuple[int, int, int]:
        return (self.desc.N_p, self.desc.N_c, self.desc.N_e)

    def with_law(self, law) -> "DescArrays":
        """The same operators and geometry (shared, not copied) under another law of
        the same kind, e.g. another diffusion coefficient."""
        if type(law) is not type(self.law) or law.N_c != self.law.N_c:
            raise ValueError("with_law keeps the kind of conservation law")
        other = copy.copy(self)
        other.law = law
        desc = abi.SseDesc.from_buffer_copy(self.desc)
        if isinstance(law, EulerEquations):
            desc.gamma = law.gamma
        if isinstance(law, LinearAdvectionDiffusionEquation):
            desc.b = law.b
        if hasattr(law, "a"):
            for m, am in enumerate(law.a):
                desc.a[m] = am
        other.desc = desc
        return other


# ----------------------------------------------------------------------------- Solver
class Solver:
    """``Solver`` (``Solvers.jl:259-272``) whose residual runs on the GPU.

    Holds the descriptor arrays and, on first use, a libsse_hip handle.
    """

    def __init__(self, conservation_law, spatial_discretization: SpatialDiscretization, form,
                 strategy=None, mass_solver=None, parallelism=None):
        self.conservation_law = conservation_law
        self.spatial_discretization = spatial_discretization
        self.form = form
        self.strategy = strategy or ReferenceOperator()
        self.mass_solver = mass_solver or default_mass_matrix_solver(spatial_discretization)
        self.parallelism = parallelism or MI355X()
        self.arrays = DescArrays(conservation_law, spatial_discretization, form, self.mass_solver,
                                 self.strategy)
        self.arrays.desc.deterministic = int(getattr(self.parallelism, "deterministic", False))
        self._handle = None
        self._lib = None

    def with_law(self, conservation_law) -> "Solver":
        """A solver for the same discretisation and mesh under another law of the
        same kind (a parameter study: e.g. another diffusion coefficient). The host
        arrays are shared; a BatchedSolver over such solvers keeps one device copy
        of the geometry."""
        other = object.__new__(Solver)
        other.__dict__.update(self.__dict__)
        other.conservation_law = conservation_law
        other.arrays = self.arrays.with_law(conservation_law)
        other._handle = None
        other._lib = None
        return other

    @property
    def connectivity(self) -> np.ndarray:
        return self.spatial_discretization.mesh.mapP

    def size(self) -> Tuple[int, int, int]:
        """(N_p, N_c, N_e) (``Solvers.jl:276-285``)."""
        return self.arrays.shape

    # -- device handle
    @property
    def handle(self):
        if self._handle is None:
            from .hip import check, load_library
            self._lib = load_library()
            h = C.c_void_p()
            code = self._lib.sse_create(C.byref(self.arrays.desc), C.byref(h))
            check(self._lib, None, code)
            self._handle = h
        return self._handle

    @property
    def lib(self):
        self.handle
        return self._lib

    def describe(self) -> str:
        """Which kernels the handle runs (tensor/generic, launch variant, table sizes)."""
        buf = C.create_string_buffer(256)
        self.lib.sse_describe(self.handle, buf, 256)
        return buf.value.decode()

    def close(self):
        if self._handle is not None:
            self._lib.sse_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def semi_discrete_residual(dudt, u, solver: Solver, t: float = 0.0):
    """``semi_discrete_residual!(dudt, u, solver, t)`` (``Solvers.jl:474-514``).

    ``u``/``dudt`` are either torch CUDA tensors of shape (N_e, N_c, N_p) (the
    reference's column-major (N_p, N_c, N_e) memory) or Fortran-ordered NumPy
    arrays of shape (N_p, N_c, N_e) (copied through host memory).
    """
    from .hip import check
    lib, h = solver.lib, solver.handle
    if isinstance(u, np.ndarray):
        if not (u.flags.f_contiguous and dudt.flags.f_contiguous):
            raise ValueError("host arrays must be Fortran-ordered (N_p, N_c, N_e)")
        if u.shape != solver.size() or dudt.shape != solver.size():
            raise ValueError(f"expected shape {solver.size()}")
        check(lib, h, lib.sse_residual(h, u.ctypes.data, dudt.ctypes.data, float(t), abi.SSE_HOST))
        return dudt
    _check_device_tensor(u, solver)
    _check_device_tensor(dudt, solver)
    check(lib, h, lib.sse_residual(h, u.data_ptr(), dudt.data_ptr(), float(t), abi.SSE_DEVICE))
    return dudt


def _check_device_tensor(x, solver: Solver):
    import torch
    N_p, N_c, N_e = solver.size()
    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64
            and x.is_contiguous() and tuple(x.shape) == (N_e, N_c, N_p)):
        raise ValueError(f"device state must be a contiguous float64 CUDA tensor of shape {(N_e, N_c, N_p)}")


# ----------------------------------------------------------------------------- initial condition
def project_function(initial_data, sd: SpatialDiscretization) -> np.ndarray:
    """(``Solvers.jl:388-418``): nodal evaluation for V = I, else per-element
    L2 projection with the Cholesky factor of V' W J V."""
    ref = sd.reference_approximation
    xq = sd.mesh.xyzq
    uq = np.asarray(initial_data(tuple(xq), 0.0))  # N_q x N_e x N_c
    uq = np.transpose(uq, (0, 2, 1))  # N_q x N_c x N_e
    if ref.V is None:
        return np.asfortranarray(uq)
    V = ref.V
    if getattr(sd, "backend", "host") == "hip":
        return _project_l2_hip(V, ref.W, sd.geometric_factors.J_q, uq)
    WJ = ref.W[:, None] * sd.geometric_factors.J_q  # N_q x N_e
    N_q, N_p = V.shape
    N_e = WJ.shape[1]
    u0 = np.empty((N_e, N_p, uq.shape[1]))
    step = max(1, (1 << 25) // (N_q * N_p))  # elements per batch (bounded temporaries)
    for k0 in range(0, N_e, step):
        k1 = min(N_e, k0 + step)
        VW = V.T[None, :, :] * WJ[:, k0:k1].T[:, None, :]          # K x N_p x N_q
        M = VW @ V                                                  # K x N_p x N_p
        rhs = VW @ np.transpose(uq[:, :, k0:k1], (2, 0, 1))         # K x N_p x N_c
        L = np.linalg.cholesky(M)
        y = np.linalg.solve(L, rhs)
        u0[k0:k1] = np.linalg.solve(np.transpose(L, (0, 2, 1)), y)
    return np.asfortranarray(np.transpose(u0, (1, 2, 0)))


def _project_l2_hip(V, W, J_q, uq) -> np.ndarray:
    """The per-element projection of `project_function` on the GPU (``sse_project_l2``:
    Cholesky of V' W J V in LDS, one workgroup per element)."""
    from .hip import check, load_library
    lib = load_library()
    N_q, N_p = V.shape
    _, N_c, N_e = uq.shape
    Vf = np.asfortranarray(V, dtype=np.float64)
    Wc = np.ascontiguousarray(W, dtype=np.float64)
    Jf = np.asfortranarray(J_q, dtype=np.float64)
    uqf = np.asfortranarray(uq, dtype=np.float64)
    u0 = np.empty((N_p, N_c, N_e), order="F")
    check(lib, None, lib.sse_project_l2(N_q, N_p, N_c, N_e, Vf.ctypes.data, Wc.ctypes.data, Jf.ctypes.data,
                                        uqf.ctypes.data, u0.ctypes.data, abi.SSE_HOST))
    return u0


def initialize(initial_data, sd: SpatialDiscretization) -> np.ndarray:
    return project_function(initial_data, sd)


@dataclass
class ODEProblem:
    """The ``ODEProblem(semi_discrete_residual!, u0, tspan, solver)`` of ``Solvers.jl:451``."""
    f: Any
    u0: np.ndarray
    tspan: Tuple[float, float]
    p: Solver


def semidiscretize(conservation_law, spatial_discretization, initial_data, form, tspan,
                   strategy=None, mass_matrix_solver=None, parallelism=None) -> ODEProblem:
    """(``Solvers.jl:429-452``)."""
    u0 = initialize(initial_data, spatial_discretization)
    solver = Solver(conservation_law, spatial_discretization, form, strategy,
                    mass_matrix_solver, parallelism)
    return ODEProblem(semi_discrete_residual, u0, tuple(tspan), solver)
