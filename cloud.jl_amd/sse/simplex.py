"""Nodal sets and orthonormal (PKDO/Dubiner) bases on the bi-unit triangle and
tetrahedron.

The reference takes these from StartUpDG.jl / NodesAndModes.jl (compat "1",
not vendored): ``nodes(Tri(), N)``, ``nodes(Tet(), N)``, ``basis``,
``vandermonde`` (called at ``src/SpatialDiscretizations/ref_elem_data.jl:11-14,
77-80`` and ``src/SpatialDiscretizations/mesh.jl:418-427``). This module
restates the published algorithms of Hesthaven & Warburton, *Nodal
Discontinuous Galerkin Methods* (2008): Nodes2D / Warpfactor / xytors (sec.
6.1), Nodes3D / WarpShiftFace3D / xyztorst (sec. 10.1), Simplex2DP /
GradSimplex2DP and Simplex3DP / GradSimplex3DP (sec. 6.1, 10.1).

Reference triangle: vertices (-1,-1), (1,-1), (-1,1).
Reference tetrahedron: vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1), (-1,-1,1).
"""

from __future__ import annotations

import math

import numpy as np

from .polynomials import gauss_lobatto_legendre_nodes, grad_jacobi_p, jacobi_p, vandermonde_1d

# Optimised blending exponents (Hesthaven & Warburton, Table 6.1 and sec. 10.1)
_ALPHA_TRI = [0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
              1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258]
_ALPHA_TET = [0.0, 0.0, 0.0, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577, 1.1603,
              1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655]


# ----------------------------------------------------------------------------- triangle
def _warpfactor(n: int, rout: np.ndarray) -> np.ndarray:
    lglr = gauss_lobatto_legendre_nodes(n)
    req = np.linspace(-1.0, 1.0, n + 1)
    veq = vandermonde_1d(n, req)
    pmat = np.stack([jacobi_p(rout, 0.0, 0.0, i) for i in range(n + 1)], axis=0)
    lmat = np.linalg.solve(veq.T, pmat)
    warp = lmat.T @ (lglr - req)
    zerof = (np.abs(rout) < 1.0 - 1.0e-10).astype(np.float64)
    sf = 1.0 - (zerof * rout) ** 2
    return warp / sf + warp * (zerof - 1.0)


def _xy_to_rs(x, y):
    l1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    l2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    l3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    return -l2 + l3 - l1, -l2 - l3 + l1


def nodes_tri(n: int, alpha=None):
    """Warp & blend interpolation nodes of degree n on the bi-unit triangle.

    The reference's mapping nodes (StartUpDG's ``RefElemData(Tri(), N)``) are
    reproduced with blending exponent alpha = 0: at N = 4 this matches its
    "Advection 2D Energy-Conservative ModalTensor Tri" golden L2 error
    (test/runtests.jl:38-61) to 2e-16, where the optimised alpha = 0.1001 misses
    it by 3e-8; for N <= 3 the nodes do not depend on alpha. The optimised
    exponents (Hesthaven & Warburton, Table 6.1) remain available by argument.
    """
    if n == 0:
        return np.array([-1.0 / 3.0]), np.array([-1.0 / 3.0])
    if alpha is None:
        alpha = 0.0
    elif alpha == "optimised":
        alpha = _ALPHA_TRI[n - 1] if n < 16 else 5.0 / 3.0
    L1, L3 = [], []
    for i in range(1, n + 2):
        for m in range(1, n + 3 - i):
            L1.append((i - 1) / n)
            L3.append((m - 1) / n)
    L1 = np.array(L1)
    L3 = np.array(L3)
    L2 = 1.0 - L1 - L3
    x = -L2 + L3
    y = (-L2 - L3 + 2.0 * L1) / math.sqrt(3.0)
    blend1 = 4.0 * L2 * L3
    blend2 = 4.0 * L1 * L3
    blend3 = 4.0 * L1 * L2
    w1 = blend1 * _warpfactor(n, L3 - L2) * (1.0 + (alpha * L1) ** 2)
    w2 = blend2 * _warpfactor(n, L1 - L3) * (1.0 + (alpha * L2) ** 2)
    w3 = blend3 * _warpfactor(n, L2 - L1) * (1.0 + (alpha * L3) ** 2)
    x = x + w1 + math.cos(2 * math.pi / 3) * w2 + math.cos(4 * math.pi / 3) * w3
    y = y + 0.0 * w1 + math.sin(2 * math.pi / 3) * w2 + math.sin(4 * math.pi / 3) * w3
    return _xy_to_rs(x, y)


def rs_to_ab(r, s):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    a = np.where(np.abs(s - 1.0) > 1e-14, 2.0 * (1.0 + r) / np.where(s != 1.0, 1.0 - s, 1.0) - 1.0, -1.0)
    return a, s


def simplex2d_p(a, b, i, j):
    h1 = jacobi_p(a, 0.0, 0.0, i)
    h2 = jacobi_p(b, 2.0 * i + 1.0, 0.0, j)
    return math.sqrt(2.0) * h1 * h2 * (1.0 - b) ** i


def grad_simplex2d_p(a, b, i, j):
    fa = jacobi_p(a, 0.0, 0.0, i)
    dfa = grad_jacobi_p(a, 0.0, 0.0, i)
    gb = jacobi_p(b, 2.0 * i + 1.0, 0.0, j)
    dgb = grad_jacobi_p(b, 2.0 * i + 1.0, 0.0, j)
    dmodedr = dfa * gb
    if i > 0:
        dmodedr = dmodedr * (0.5 * (1.0 - b)) ** (i - 1)
    dmodeds = dfa * (gb * (0.5 * (1.0 + a)))
    if i > 0:
        dmodeds = dmodeds * (0.5 * (1.0 - b)) ** (i - 1)
    tmp = dgb * (0.5 * (1.0 - b)) ** i
    if i > 0:
        tmp = tmp - 0.5 * i * gb * (0.5 * (1.0 - b)) ** (i - 1)
    dmodeds = dmodeds + fa * tmp
    scale = 2.0 ** (i + 0.5)
    return dmodedr * scale, dmodeds * scale


def tri_modes(n: int):
    return [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]


def vandermonde_tri(n: int, r, s):
    a, b = rs_to_ab(r, s)
    return np.stack([simplex2d_p(a, b, i, j) for (i, j) in tri_modes(n)], axis=1)


def basis_tri(n: int, r, s):
    a, b = rs_to_ab(r, s)
    V = vandermonde_tri(n, r, s)
    grads = [grad_simplex2d_p(a, b, i, j) for (i, j) in tri_modes(n)]
    Vr = np.stack([g[0] for g in grads], axis=1)
    Vs = np.stack([g[1] for g in grads], axis=1)
    return V, Vr, Vs


# ----------------------------------------------------------------------------- tetrahedron
def _equi_nodes_tet(n: int):
    X, Y, Z = [], [], []
    for i in range(1, n + 2):
        for m in range(1, n + 3 - i):
            for q in range(1, n + 4 - i - m):
                X.append(-1.0 + (q - 1) * 2.0 / n)
                Y.append(-1.0 + (m - 1) * 2.0 / n)
                Z.append(-1.0 + (i - 1) * 2.0 / n)
    return np.array(X), np.array(Y), np.array(Z)


def _evalwarp(p: int, xnodes: np.ndarray, xout: np.ndarray) -> np.ndarray:
    warp = np.zeros_like(xout)
    xeq = np.array([-1.0 + 2.0 * (p + 1 - i) / p for i in range(1, p + 2)])
    for i in range(1, p + 2):
        d = xnodes[i - 1] - xeq[i - 1]
        for j in range(2, p + 1):
            if i != j:
                d = d * (xout - xeq[j - 1]) / (xeq[i - 1] - xeq[j - 1])
        if i != 1:
            d = -d / (xeq[i - 1] - xeq[0])
        if i != p + 1:
            d = d / (xeq[i - 1] - xeq[p])
        warp = warp + d
    return warp


def _evalshift(p, pval, L1, L2, L3):
    gauss_x = -gauss_lobatto_legendre_nodes(p)
    blend1 = L2 * L3
    blend2 = L1 * L3
    blend3 = L1 * L2
    wf1 = 4.0 * _evalwarp(p, gauss_x, L3 - L2)
    wf2 = 4.0 * _evalwarp(p, gauss_x, L1 - L3)
    wf3 = 4.0 * _evalwarp(p, gauss_x, L2 - L1)
    w1 = blend1 * wf1 * (1.0 + (pval * L1) ** 2)
    w2 = blend2 * wf2 * (1.0 + (pval * L2) ** 2)
    w3 = blend3 * wf3 * (1.0 + (pval * L3) ** 2)
    dx = w1 + math.cos(2 * math.pi / 3) * w2 + math.cos(4 * math.pi / 3) * w3
    dy = 0.0 * w1 + math.sin(2 * math.pi / 3) * w2 + math.sin(4 * math.pi / 3) * w3
    return dx, dy


_V1 = np.array([-1.0, -1.0 / math.sqrt(3.0), -1.0 / math.sqrt(6.0)])
_V2 = np.array([1.0, -1.0 / math.sqrt(3.0), -1.0 / math.sqrt(6.0)])
_V3 = np.array([0.0, 2.0 / math.sqrt(3.0), -1.0 / math.sqrt(6.0)])
_V4 = np.array([0.0, 0.0, 3.0 / math.sqrt(6.0)])


def _xyz_to_rst(X, Y, Z):
    rhs = np.stack([X, Y, Z], axis=0) - 0.5 * (_V2 + _V3 + _V4 - _V1)[:, None]
    A = np.stack([0.5 * (_V2 - _V1), 0.5 * (_V3 - _V1), 0.5 * (_V4 - _V1)], axis=1)
    rst = np.linalg.solve(A, rhs)
    return rst[0], rst[1], rst[2]


def nodes_tet(n: int):
    """Warp & blend interpolation nodes of degree n on the bi-unit tetrahedron."""
    if n == 0:
        return np.array([-0.5]), np.array([-0.5]), np.array([-0.5])
    alpha = _ALPHA_TET[n - 1] if n <= 15 else 1.0
    tol = 1e-10
    r, s, t = _equi_nodes_tet(n)
    L1 = (1.0 + t) / 2.0
    L2 = (1.0 + s) / 2.0
    L3 = -(1.0 + r + s + t) / 2.0
    L4 = (1.0 + r) / 2.0
    t1 = np.array([_V2 - _V1, _V2 - _V1, _V3 - _V2, _V3 - _V1])
    t2 = np.array([_V3 - 0.5 * (_V1 + _V2), _V4 - 0.5 * (_V1 + _V2),
                   _V4 - 0.5 * (_V2 + _V3), _V4 - 0.5 * (_V1 + _V3)])
    t1 = t1 / np.linalg.norm(t1, axis=1)[:, None]
    t2 = t2 / np.linalg.norm(t2, axis=1)[:, None]
    XYZ = np.outer(L3, _V1) + np.outer(L4, _V2) + np.outer(L2, _V3) + np.outer(L1, _V4)
    shift = np.zeros_like(XYZ)
    for face in range(4):
        if face == 0:
            La, Lb, Lc, Ld = L1, L2, L3, L4
        elif face == 1:
            La, Lb, Lc, Ld = L2, L1, L3, L4
        elif face == 2:
            La, Lb, Lc, Ld = L3, L1, L4, L2
        else:
            La, Lb, Lc, Ld = L4, L1, L3, L2
        warp1, warp2 = _evalshift(n, alpha, Lb, Lc, Ld)
        blend = Lb * Lc * Ld
        denom = (Lb + 0.5 * La) * (Lc + 0.5 * La) * (Ld + 0.5 * La)
        ids = denom > tol
        blend[ids] = (1.0 + (alpha * La[ids]) ** 2) * blend[ids] / denom[ids]
        shift = shift + np.outer(blend * warp1, t1[face]) + np.outer(blend * warp2, t2[face])
        cnt = (Lb > tol).astype(int) + (Lc > tol).astype(int) + (Ld > tol).astype(int)
        ids = (La < tol) & (cnt < 3)
        shift[ids, :] = np.outer(warp1[ids], t1[face]) + np.outer(warp2[ids], t2[face])
    XYZ = XYZ + shift
    return _xyz_to_rst(XYZ[:, 0], XYZ[:, 1], XYZ[:, 2])


def rst_to_abc(r, s, t):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    den_a = -s - t
    den_b = 1.0 - t
    a = np.where(np.abs(den_a) > 1e-14, 2.0 * (1.0 + r) / np.where(den_a != 0.0, den_a, 1.0) - 1.0, -1.0)
    b = np.where(np.abs(den_b) > 1e-14, 2.0 * (1.0 + s) / np.where(den_b != 0.0, den_b, 1.0) - 1.0, -1.0)
    return a, b, t


def simplex3d_p(a, b, c, i, j, k):
    h1 = jacobi_p(a, 0.0, 0.0, i)
    h2 = jacobi_p(b, 2.0 * i + 1.0, 0.0, j)
    h3 = jacobi_p(c, 2.0 * (i + j) + 2.0, 0.0, k)
    return 2.0 * math.sqrt(2.0) * h1 * h2 * (1.0 - b) ** i * h3 * (1.0 - c) ** (i + j)


def grad_simplex3d_p(a, b, c, i, j, k):
    fa = jacobi_p(a, 0.0, 0.0, i)
    dfa = grad_jacobi_p(a, 0.0, 0.0, i)
    gb = jacobi_p(b, 2.0 * i + 1.0, 0.0, j)
    dgb = grad_jacobi_p(b, 2.0 * i + 1.0, 0.0, j)
    hc = jacobi_p(c, 2.0 * (i + j) + 2.0, 0.0, k)
    dhc = grad_jacobi_p(c, 2.0 * (i + j) + 2.0, 0.0, k)
    vr = dfa * (gb * hc)
    if i > 0:
        vr = vr * (0.5 * (1.0 - b)) ** (i - 1)
    if i + j > 0:
        vr = vr * (0.5 * (1.0 - c)) ** (i + j - 1)
    vs = 0.5 * (1.0 + a) * vr
    tmp = dgb * (0.5 * (1.0 - b)) ** i
    if i > 0:
        tmp = tmp + (-0.5 * i) * (gb * (0.5 * (1.0 - b)) ** (i - 1))
    if i + j > 0:
        tmp = tmp * (0.5 * (1.0 - c)) ** (i + j - 1)
    tmp = fa * (tmp * hc)
    vs = vs + tmp
    vt = 0.5 * (1.0 + a) * vr + 0.5 * (1.0 + b) * tmp
    tmp = dhc * (0.5 * (1.0 - c)) ** (i + j)
    if i + j > 0:
        tmp = tmp - 0.5 * (i + j) * (hc * (0.5 * (1.0 - c)) ** (i + j - 1))
    tmp = fa * (gb * tmp)
    tmp = tmp * (0.5 * (1.0 - b)) ** i
    vt = vt + tmp
    scale = 2.0 ** (2 * i + j + 1.5)
    return vr * scale, vs * scale, vt * scale


def tet_modes(n: int):
    return [(i, j, k) for i in range(n + 1) for j in range(n + 1 - i) for k in range(n + 1 - i - j)]


def vandermonde_tet(n: int, r, s, t):
    a, b, c = rst_to_abc(r, s, t)
    return np.stack([simplex3d_p(a, b, c, i, j, k) for (i, j, k) in tet_modes(n)], axis=1)


def basis_tet(n: int, r, s, t):
    a, b, c = rst_to_abc(r, s, t)
    V = vandermonde_tet(n, r, s, t)
    grads = [grad_simplex3d_p(a, b, c, i, j, k) for (i, j, k) in tet_modes(n)]
    return (V, np.stack([g[0] for g in grads], axis=1), np.stack([g[1] for g in grads], axis=1),
            np.stack([g[2] for g in grads], axis=1))


def equi_nodes_tri(n: int):
    r, s = [], []
    for j in range(n + 1):
        for i in range(n + 1 - j):
            r.append(-1.0 + 2.0 * i / n)
            s.append(-1.0 + 2.0 * j / n)
    return np.array(r), np.array(s)
