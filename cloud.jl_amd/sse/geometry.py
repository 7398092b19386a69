"""Geometric factors and the ``SpatialDiscretization`` container.

Restates ``src/SpatialDiscretizations/mesh.jl:211-506`` (``metrics``,
``GeometricFactors`` for exact and conservative-curl metrics) and
``src/SpatialDiscretizations/SpatialDiscretizations.jl:283-438``
(``GeometricFactors``, ``project_jacobian!``, ``SpatialDiscretization``,
``apply_reference_mapping``, ``check_normals``, ``check_facet_nodes``).
The curl-form metric terms follow StartUpDG's ``geometric_factors``
(Kopriva 2006; Chan & Wilcox 2019): J dr/dx = d/dt((d/ds y) z) - d/ds((d/dt y) z)
etc., evaluated on the mapping nodes.

Array shapes follow the reference exactly (column-major semantics, stored as
Fortran-ordered NumPy arrays):
J_q (N_q, N_e); Lambda_q (N_q, d, d, N_e) with Lambda_q[i, l, m, k] = J d xi_l/d x_m;
J_f (N_f, N_e); nJf (d, N_f, N_e); nJq (d, N_fac, N_q, N_e).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import simplex as S
from .mesh import MeshData
from .reference_approximation import HEX, LINE, TET, TRI, ReferenceApproximation, num_faces


class ExactMetrics:
    pass


class ConservativeCurlMetrics:
    pass


ChanWilcoxMetrics = ConservativeCurlMetrics


@dataclass
class GeometricFactors:
    J_q: np.ndarray
    Lambda_q: np.ndarray
    J_f: np.ndarray
    nJf: np.ndarray
    nJq: np.ndarray


def _f(a):
    return np.asfortranarray(a)


def _normals(ref: ReferenceApproximation, Lambda_q, Lambda_f):
    """nJq[n, f, i, k] = sum_m Lambda_q[i, m, n, k] n_ref_f[m];
    nJf[m, i, k] = sum_n Lambda_f[i, n, m, k] nrstJ[n][i]; J_f = |nJf|."""
    rd = ref.reference_element
    d = ref.d
    nfac = num_faces(rd.element_type)
    npf = ref.N_f // nfac
    n_ref = np.array([[rd.nrstJ[m][npf * f] for m in range(d)] for f in range(nfac)])  # nfac x d
    nJq = np.einsum("imnk,fm->nfik", Lambda_q, n_ref)
    nrst = np.stack(rd.nrstJ, axis=0)  # d x N_f
    nJf = np.einsum("inmk,ni->mik", Lambda_f, nrst)
    J_f = np.sqrt((nJf ** 2).sum(0))
    return _f(nJq), _f(nJf), _f(J_f)


def _metrics_exact(dxdr):
    """``metrics`` (``mesh.jl:211-227``): dxdr[..., m, n] = d x_m / d r_n."""
    d = dxdr.shape[-1]
    if d == 1:
        return dxdr[..., 0, 0], np.ones_like(dxdr)
    if d == 2:
        J = dxdr[..., 0, 0] * dxdr[..., 1, 1] - dxdr[..., 0, 1] * dxdr[..., 1, 0]
        L = np.empty_like(dxdr)
        L[..., 0, 0] = dxdr[..., 1, 1]
        L[..., 0, 1] = -dxdr[..., 0, 1]
        L[..., 1, 0] = -dxdr[..., 1, 0]
        L[..., 1, 1] = dxdr[..., 0, 0]
        return J, L
    J = np.linalg.det(dxdr)
    return J, J[..., None, None] * np.linalg.inv(dxdr)


def geometric_factors_exact(mesh: MeshData, ref: ReferenceApproximation) -> GeometricFactors:
    """(``mesh.jl:229-282``)."""
    rd = ref.reference_element
    d = ref.d
    dxdr = np.stack([np.stack([rd.Drst[n] @ mesh.xyz[m] for n in range(d)], -1) for m in range(d)], -2)
    # dxdr: N_nodes x N_e x d(m) x d(n)
    dq = np.einsum("qa,akmn->qkmn", rd.Vq, dxdr)
    df = np.einsum("qa,akmn->qkmn", rd.Vf, dxdr)
    J_q, Lq = _metrics_exact(dq)          # N_q x N_e, N_q x N_e x d x d
    _, Lf = _metrics_exact(df)
    Lambda_q = np.transpose(Lq, (0, 2, 3, 1))
    Lambda_f = np.transpose(Lf, (0, 2, 3, 1))
    nJq, nJf, J_f = _normals(ref, Lambda_q, Lambda_f)
    return GeometricFactors(_f(J_q), _f(Lambda_q), J_f, nJf, nJq)


def _gf2d(x, y, Dr, Ds):
    xr, xs = Dr @ x, Ds @ x
    yr, ys = Dr @ y, Ds @ y
    J = -xs * yr + xr * ys
    return ys, -yr, -xs, xr, J  # rxJ, sxJ, ryJ, syJ, J


def _gf3d(x, y, z, Dr, Ds, Dt):
    xr, xs, xt = Dr @ x, Ds @ x, Dt @ x
    yr, ys, yt = Dr @ y, Ds @ y, Dt @ y
    zr, zs, zt = Dr @ z, Ds @ z, Dt @ z
    Fr, Fs, Ft = (Dr @ y) * z, (Ds @ y) * z, (Dt @ y) * z
    rxJ = Dt @ Fs - Ds @ Ft
    sxJ = Dr @ Ft - Dt @ Fr
    txJ = Ds @ Fr - Dr @ Fs
    Fr, Fs, Ft = (Dr @ x) * z, (Ds @ x) * z, (Dt @ x) * z
    ryJ = -(Dt @ Fs - Ds @ Ft)
    syJ = -(Dr @ Ft - Dt @ Fr)
    tyJ = -(Ds @ Fr - Dr @ Fs)
    Fr, Fs, Ft = (Dr @ y) * x, (Ds @ y) * x, (Dt @ y) * x
    rzJ = -(Dt @ Fs - Ds @ Ft)
    szJ = -(Dr @ Ft - Dt @ Fr)
    tzJ = -(Ds @ Fr - Dr @ Fs)
    J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt)
    return rxJ, sxJ, txJ, ryJ, syJ, tyJ, rzJ, szJ, tzJ, J


def _tet_curl_operators(rd):
    """The degree-(N+1) nodal operators of the tetrahedral curl metrics
    (``mesh.jl:416-433``): degree-N -> degree-(N+1) interpolation, the three
    derivative matrices at the degree-(N+1) nodes, and Vq, Vf composed with the
    degree-(N+1) -> degree-N interpolation."""
    N = rd.N
    r1, s1, t1 = S.nodes_tet(N + 1)
    V1p, Vr1, Vs1, Vt1 = S.basis_tet(N + 1, r1, s1, t1)
    V1pinv = np.linalg.inv(V1p)
    N_to_N1 = S.vandermonde_tet(N, r1, s1, t1) @ np.linalg.inv(rd.VDM)
    N1_to_N = S.vandermonde_tet(N + 1, *rd.rst) @ V1pinv
    D1 = (Vr1 @ V1pinv, Vs1 @ V1pinv, Vt1 @ V1pinv)
    return N_to_N1, D1, rd.Vq @ N1_to_N, rd.Vf @ N1_to_N


def geometric_factors_curl(mesh: MeshData, ref: ReferenceApproximation) -> GeometricFactors:
    """Conservative-curl (Chan-Wilcox) metrics: Tri ``mesh.jl:284-339``,
    Hex ``mesh.jl:341-408`` (the curl form on the mapping nodes, then interpolated),
    Tet ``mesh.jl:410-506``."""
    rd = ref.reference_element
    elem = rd.element_type
    if elem == TRI:
        Dr, Ds = rd.Drst
        rxJ, sxJ, ryJ, syJ, J = _gf2d(mesh.xyz[0], mesh.xyz[1], Dr, Ds)
        comps = {(0, 0): rxJ, (1, 0): sxJ, (0, 1): ryJ, (1, 1): syJ}
        N_q, N_e = rd.Vq.shape[0], J.shape[1]
        Lq = np.empty((N_q, 2, 2, N_e))
        Lf = np.empty((rd.Vf.shape[0], 2, 2, N_e))
        for (l, m), v in comps.items():
            Lq[:, l, m, :] = rd.Vq @ v
            Lf[:, l, m, :] = rd.Vf @ v
        J_q = rd.Vq @ J
    elif elem == HEX:
        g = _gf3d(*mesh.xyz, *rd.Drst)
        rxJ, sxJ, txJ, ryJ, syJ, tyJ, rzJ, szJ, tzJ, J = g
        comps = {(0, 0): rxJ, (1, 0): sxJ, (2, 0): txJ, (0, 1): ryJ, (1, 1): syJ, (2, 1): tyJ,
                 (0, 2): rzJ, (1, 2): szJ, (2, 2): tzJ}
        N_e = J.shape[1]
        Lq = np.empty((rd.Vq.shape[0], 3, 3, N_e))
        Lf = np.empty((rd.Vf.shape[0], 3, 3, N_e))
        for (l, m), v in comps.items():
            Lq[:, l, m, :] = rd.Vq @ v
            Lf[:, l, m, :] = rd.Vf @ v
        J_q = rd.Vq @ J
    elif elem == TET:
        N_to_N1, D1, Vq, Vf = _tet_curl_operators(rd)
        J = _gf3d(*mesh.xyz, *rd.Drst)[-1]
        J_q = rd.Vq @ J
        g = _gf3d(N_to_N1 @ mesh.xyz[0], N_to_N1 @ mesh.xyz[1], N_to_N1 @ mesh.xyz[2], *D1)
        rxJ, sxJ, txJ, ryJ, syJ, tyJ, rzJ, szJ, tzJ = g[:9]
        comps = {(0, 0): rxJ, (1, 0): sxJ, (2, 0): txJ, (0, 1): ryJ, (1, 1): syJ, (2, 1): tyJ,
                 (0, 2): rzJ, (1, 2): szJ, (2, 2): tzJ}
        N_e = J.shape[1]
        Lq = np.empty((Vq.shape[0], 3, 3, N_e))
        Lf = np.empty((Vf.shape[0], 3, 3, N_e))
        for (l, m), v in comps.items():
            Lq[:, l, m, :] = Vq @ v
            Lf[:, l, m, :] = Vf @ v
    else:
        raise NotImplementedError(elem)
    nJq, nJf, J_f = _normals(ref, Lq, Lf)
    return GeometricFactors(_f(J_q), _f(Lq), J_f, nJf, nJq)


def _jacobian_projector(ref: ReferenceApproximation) -> np.ndarray:
    V = ref.Vdense()
    W = ref.W
    return V @ np.linalg.solve(V.T @ (W[:, None] * V), V.T * W[None, :])


def project_jacobian(J_q: np.ndarray, ref: ReferenceApproximation) -> np.ndarray:
    """(``SpatialDiscretizations.jl:310-317``)."""
    return _f(_jacobian_projector(ref) @ J_q)


@dataclass
class SpatialDiscretization:
    """(``SpatialDiscretizations.jl:299-392``)."""
    mesh: MeshData
    reference_approximation: ReferenceApproximation
    geometric_factors: GeometricFactors
    backend: str = "host"   # where set-up steps run ("hip": sse_* set-up entries on the GPU)

    @property
    def N_e(self) -> int:
        return self.mesh.N_e

    @property
    def d(self) -> int:
        return self.reference_approximation.d


def _is_exact(metric_type) -> bool:
    return metric_type is None or isinstance(metric_type, ExactMetrics) or metric_type is ExactMetrics


def make_spatial_discretization(mesh: MeshData, ref: ReferenceApproximation, metric_type=None,
                                project_jacobian_flag: bool = True, backend: str = "host") -> SpatialDiscretization:
    """`backend="hip"` computes the geometric factors on the GPU (``sse_geometric_factors``)
    and marks the discretisation so that the later set-up steps (the reference mapping
    of StandardForm, the projection of the initial data) run there too."""
    if backend == "hip":
        gf = geometric_factors_hip(mesh, ref, metric_type, project_jacobian_flag)
    elif backend != "host":
        raise ValueError(f"unknown set-up backend {backend!r}")
    elif _is_exact(metric_type):
        gf = geometric_factors_exact(mesh, ref)
        if project_jacobian_flag:
            gf.J_q = project_jacobian(gf.J_q, ref)
    else:
        gf = geometric_factors_curl(mesh, ref)
    return SpatialDiscretization(mesh, ref, gf, backend)


def geometric_factors_hip(mesh: MeshData, ref: ReferenceApproximation, metric_type=None,
                          project_jacobian_flag: bool = True) -> GeometricFactors:
    """Exact (``mesh.jl:229-282``) or conservative-curl (``mesh.jl:284-339``,
    ``:410-506``) geometric factors computed by ``sse_geometric_factors`` on the GPU,
    one workgroup per element; the Jacobian of exact metrics projected as by
    ``project_jacobian!``. Same arrays as the host functions above."""
    from . import abi
    from .hip import check, load_library
    lib = load_library()
    rd = ref.reference_element
    d = ref.d
    exact = _is_exact(metric_type)
    if not exact and rd.element_type not in (TRI, TET):
        raise NotImplementedError(f"curl metrics on {rd.element_type}")
    keep = []

    def ptr(a):
        if a is None:
            return None
        a = np.asfortranarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double))

    N_p, N_e = mesh.xyz[0].shape
    N_q, N_f = rd.Vq.shape[0], rd.Vf.shape[0]
    N_fac = num_faces(rd.element_type)
    ops = abi.SseGeometryOps()
    ops.d, ops.N_p, ops.N_q, ops.N_f, ops.N_fac = d, N_p, N_q, N_f, N_fac
    ops.metrics = abi.SSE_METRICS_EXACT if exact else abi.SSE_METRICS_CURL
    ops.D = ptr(np.stack(rd.Drst, axis=2))
    ops.Vq, ops.Vf = ptr(rd.Vq), ptr(rd.Vf)
    ops.nrstJ = ptr(np.stack(rd.nrstJ, axis=1))
    ops.P_J = ptr(_jacobian_projector(ref)) if (exact and project_jacobian_flag) else None
    if not exact and rd.element_type == TET:
        N_to_N1, D1, Vq1, Vf1 = _tet_curl_operators(rd)
        ops.N_p1 = N_to_N1.shape[0]
        ops.I1, ops.D1 = ptr(N_to_N1), ptr(np.stack(D1, axis=2))
        ops.Vq1, ops.Vf1 = ptr(Vq1), ptr(Vf1)
    xyz = [np.asfortranarray(c, dtype=np.float64) for c in mesh.xyz] + [None] * (3 - d)
    J_q = np.empty((N_q, N_e), order="F")
    Lq = np.empty((N_q, d, d, N_e), order="F")
    J_f = np.empty((N_f, N_e), order="F")
    nJf = np.empty((d, N_f, N_e), order="F")
    nJq = np.empty((d, N_fac, N_q, N_e), order="F")
    vp = lambda a: None if a is None else a.ctypes.data
    check(lib, None, lib.sse_geometric_factors(C.byref(ops), N_e, vp(xyz[0]), vp(xyz[1]), vp(xyz[2]),
                                               vp(J_q), vp(Lq), vp(J_f), vp(nJf), vp(nJq), abi.SSE_HOST))
    return GeometricFactors(J_q, Lq, J_f, nJf, nJq)


def apply_reference_mapping(gf: GeometricFactors, ref: ReferenceApproximation,
                            backend: str = "host") -> GeometricFactors:
    """Lambda_eta[i, m, n, k] = sum_l Lambda_ref[i, m, l] Lambda_q[i, l, n, k] / J_ref[i]
    (``SpatialDiscretizations.jl:395-411``); on the GPU with `backend="hip"`
    (``sse_apply_reference_mapping``)."""
    if ref.reference_mapping is None:
        return gf
    Jr = ref.reference_mapping.J_ref
    Lr = ref.reference_mapping.Lambda_ref
    if backend == "hip":
        from . import abi
        from .hip import check, load_library
        lib = load_library()
        N_q, d, _, N_e = gf.Lambda_q.shape
        Lr_f = np.asfortranarray(Lr, dtype=np.float64)
        Jr_f = np.ascontiguousarray(Jr, dtype=np.float64)
        Lq = np.asfortranarray(gf.Lambda_q, dtype=np.float64)
        Le = np.empty((N_q, d, d, N_e), order="F")
        check(lib, None, lib.sse_apply_reference_mapping(N_q, d, N_e, Lr_f.ctypes.data, Jr_f.ctypes.data,
                                                         Lq.ctypes.data, Le.ctypes.data, abi.SSE_HOST))
        return GeometricFactors(gf.J_q, Le, gf.J_f, gf.nJf, gf.nJq)
    Le = np.einsum("iml,ilnk->imnk", Lr, gf.Lambda_q) / Jr[:, None, None, None]
    return GeometricFactors(gf.J_q, _f(Le), gf.J_f, gf.nJf, gf.nJq)


def check_normals(sd: SpatialDiscretization):
    """max |nJf(k) + nJf(mapP)| per direction (``SpatialDiscretizations.jl:425-430``)."""
    gf = sd.geometric_factors
    mp = sd.mesh.mapP.reshape(-1, order="F") - 1
    out = []
    for m in range(sd.d):
        a = gf.nJf[m].reshape(-1, order="F")
        out.append(float(np.abs(a + a[mp]).max()))
    return tuple(out)


def check_facet_nodes(sd: SpatialDiscretization):
    """max periodic-wrapped |xf(k) - xf(mapP)| (``SpatialDiscretizations.jl:433-438``)."""
    mp = sd.mesh.mapP.reshape(-1, order="F") - 1
    out = []
    for m in range(sd.d):
        a = sd.mesh.xyzf[m].reshape(-1, order="F")
        dif = a - a[mp]
        L = sd.mesh.periodic_L[m]
        dif = dif - L * np.round(dif / L)
        out.append(float(np.abs(dif).max()))
    return tuple(out)
