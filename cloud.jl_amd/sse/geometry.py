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

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import simplex as S
from .mesh import MeshData
from .reference_approximation import LINE, TET, TRI, ReferenceApproximation, num_faces


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


def geometric_factors_curl(mesh: MeshData, ref: ReferenceApproximation) -> GeometricFactors:
    """Conservative-curl (Chan-Wilcox) metrics: Tri ``mesh.jl:284-339``,
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
    elif elem == TET:
        N = rd.N
        r1, s1, t1 = S.nodes_tet(N + 1)
        V1p, Vr1, Vs1, Vt1 = S.basis_tet(N + 1, r1, s1, t1)
        V1pinv = np.linalg.inv(V1p)
        N_to_N1 = S.vandermonde_tet(N, r1, s1, t1) @ np.linalg.inv(rd.VDM)
        N1_to_N = S.vandermonde_tet(N + 1, *rd.rst) @ V1pinv
        Vq = rd.Vq @ N1_to_N
        Vf = rd.Vf @ N1_to_N
        J = _gf3d(*mesh.xyz, *rd.Drst)[-1]
        J_q = rd.Vq @ J
        g = _gf3d(N_to_N1 @ mesh.xyz[0], N_to_N1 @ mesh.xyz[1], N_to_N1 @ mesh.xyz[2],
                  Vr1 @ V1pinv, Vs1 @ V1pinv, Vt1 @ V1pinv)
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


def project_jacobian(J_q: np.ndarray, ref: ReferenceApproximation) -> np.ndarray:
    """(``SpatialDiscretizations.jl:310-317``)."""
    V = ref.Vdense()
    W = ref.W
    proj = V @ np.linalg.solve(V.T @ (W[:, None] * V), V.T * W[None, :])
    return _f(proj @ J_q)


@dataclass
class SpatialDiscretization:
    """(``SpatialDiscretizations.jl:299-392``)."""
    mesh: MeshData
    reference_approximation: ReferenceApproximation
    geometric_factors: GeometricFactors

    @property
    def N_e(self) -> int:
        return self.mesh.N_e

    @property
    def d(self) -> int:
        return self.reference_approximation.d


def make_spatial_discretization(mesh: MeshData, ref: ReferenceApproximation, metric_type=None,
                                project_jacobian_flag: bool = True) -> SpatialDiscretization:
    if metric_type is None or isinstance(metric_type, ExactMetrics) or metric_type is ExactMetrics:
        gf = geometric_factors_exact(mesh, ref)
        if project_jacobian_flag:
            gf.J_q = project_jacobian(gf.J_q, ref)
    else:
        gf = geometric_factors_curl(mesh, ref)
    return SpatialDiscretization(mesh, ref, gf)


def apply_reference_mapping(gf: GeometricFactors, ref: ReferenceApproximation) -> GeometricFactors:
    """Lambda_eta[i, m, n, k] = sum_l Lambda_ref[i, m, l] Lambda_q[i, l, n, k] / J_ref[i]
    (``SpatialDiscretizations.jl:395-411``)."""
    if ref.reference_mapping is None:
        return gf
    Jr = ref.reference_mapping.J_ref
    Lr = ref.reference_mapping.Lambda_ref
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
