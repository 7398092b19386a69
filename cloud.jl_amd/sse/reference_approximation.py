"""Reference-element operators: quadrature rules, collapsed-coordinate maps, the
warped tensor-product (sum-factorised PKDO) Vandermonde, and tensor-product
SBP operators on the Line, Tri and Tet.

Restates ``src/SpatialDiscretizations/quadrature_rules.jl``,
``src/SpatialDiscretizations/tensor_simplex.jl``,
``src/SpatialDiscretizations/tensor_cartesian.jl:1-32, 85-131`` and
``src/SpatialDiscretizations/ref_elem_data.jl`` of the reference.
All matrices are dense NumPy arrays (the reference builds LinearMaps; the dense
form of each map is what crosses the C-ABI, and the library recovers the
sparsity).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

from . import polynomials as P
from . import simplex as S

# ----------------------------------------------------------------------------- element shapes
LINE, TRI, TET, QUAD, HEX = "Line", "Tri", "Tet", "Quad", "Hex"
_DIM = {LINE: 1, TRI: 2, QUAD: 2, TET: 3, HEX: 3}
_NUM_FACES = {LINE: 2, TRI: 3, QUAD: 4, TET: 4, HEX: 6}


def dim(elem: str) -> int:
    return _DIM[elem]


def num_faces(elem: str) -> int:
    return _NUM_FACES[elem]


# ----------------------------------------------------------------------------- quadrature rules
@dataclass(frozen=True)
class QuadratureRule:
    """``GaussQuadrature``/``GaussLobattoQuadrature``/``GaussRadauQuadrature``
    (``src/SpatialDiscretizations/quadrature_rules.jl:1-23``): q+1 points,
    Jacobi weight (1-x)^a (1+x)^b."""
    kind: str  # "gauss", "lobatto", "radau"
    q: int
    a: int = 0
    b: int = 0


def LGQuadrature(q: int) -> QuadratureRule:
    return QuadratureRule("gauss", q, 0, 0)


def LGLQuadrature(q: int) -> QuadratureRule:
    return QuadratureRule("lobatto", q, 0, 0)


def LGRQuadrature(q: int) -> QuadratureRule:
    return QuadratureRule("radau", q, 0, 0)


def GaussQuadrature(q: int, a: int, b: int) -> QuadratureRule:
    return QuadratureRule("gauss", q, a, b)


def quadrature_line(rule: QuadratureRule):
    """``quadrature(::Line, rule)`` (``quadrature_rules.jl:76-92``)."""
    if rule.kind == "gauss":
        return P.gauss_jacobi(rule.q + 1, rule.a, rule.b)
    if rule.kind == "lobatto":
        return P.gauss_lobatto_jacobi(rule.q + 1, rule.a, rule.b)
    if rule.kind == "radau":
        return P.gauss_radau_jacobi_m(rule.q + 1, rule.a, rule.b)
    raise ValueError(rule.kind)


def _grid2(x1, x2):
    """Tensor grid with the second coordinate fastest (``meshgrid`` then
    column-major flattening, ``quadrature_rules.jl:37-40, 102-108``)."""
    a = np.repeat(np.asarray(x1), len(x2))
    b = np.tile(np.asarray(x2), len(x1))
    return a, b


def _grid3(x1, x2, x3):
    """Tensor grid with the third coordinate fastest (``quadrature_rules.jl:42-46,
    111-118``)."""
    a = np.repeat(np.asarray(x1), len(x2) * len(x3))
    b = np.tile(np.repeat(np.asarray(x2), len(x3)), len(x1))
    c = np.tile(np.asarray(x3), len(x1) * len(x2))
    return a, b, c


def chi_tri(eta1, eta2):
    """Duffy map square -> triangle (``tensor_simplex.jl:2-4``)."""
    return 0.5 * (1.0 + eta1) * (1.0 - eta2) - 1.0, eta2


def chi_tet(eta1, eta2, eta3):
    """Collapsed map cube -> tetrahedron (``tensor_simplex.jl:7-11``)."""
    xi_pri1 = 0.5 * (1.0 + eta1) * (1.0 - eta3) - 1.0
    xi_pyr2 = 0.5 * (1.0 + eta2) * (1.0 - eta3) - 1.0
    return 0.5 * (1.0 + xi_pri1) * (1.0 - eta2) - 1.0, xi_pyr2, eta3


def _ab(rule):
    return (rule.a, rule.b)


def quadrature_tri(rules):
    """``quadrature(::Tri, (rule1, rule2))`` (``quadrature_rules.jl:125-141``)."""
    r1, w1 = quadrature_line(rules[0])
    r2, w2 = quadrature_line(rules[1])
    e1, e2 = _grid2(r1, r2)
    wa, wb = _grid2(w1, w2)
    w = wa * wb
    if _ab(rules[0]) == (0, 0) and _ab(rules[1]) == (0, 0):
        return (*chi_tri(e1, e2), 0.5 * (1.0 - e2) * w)
    if _ab(rules[0]) == (0, 0) and _ab(rules[1]) == (1, 0):
        return (*chi_tri(e1, e2), 0.5 * w)
    raise ValueError("Chosen Jacobi weight not supported.")


def quadrature_tet(rules):
    """``quadrature(::Tet, (r1, r2, r3))`` (``quadrature_rules.jl:143-166``)."""
    r1, w1 = quadrature_line(rules[0])
    r2, w2 = quadrature_line(rules[1])
    r3, w3 = quadrature_line(rules[2])
    e1, e2, e3 = _grid3(r1, r2, r3)
    wa, wb, wc = _grid3(w1, w2, w3)
    w = wa * wb * wc
    if all(_ab(r) == (0, 0) for r in rules):
        return (*chi_tet(e1, e2, e3), 0.125 * (1.0 - e2) * (1.0 - e3) ** 2 * w)
    if _ab(rules[0]) == (0, 0) and _ab(rules[1]) == (0, 0) and _ab(rules[2]) == (1, 0):
        return (*chi_tet(e1, e2, e3), 0.125 * (1.0 - e2) * (1.0 - e3) * w)
    raise ValueError("Chosen Jacobi weight not supported")


# ----------------------------------------------------------------------------- reference mapping
def reference_geometric_factors_tri(rules):
    """J_ref and Lambda_ref = J_ref * d eta / d xi (``tensor_simplex.jl:13-40``)."""
    r1, _ = quadrature_line(rules[0])
    r2, _ = quadrature_line(rules[1])
    e1, e2 = _grid2(r1, r2)
    n = len(e1)
    L = np.zeros((n, 2, 2))
    if _ab(rules[0]) == (0, 0) and _ab(rules[1]) == (0, 0):
        J = 0.5 * (1.0 - e2)
        L[:, 0, 0] = 1.0
        L[:, 0, 1] = 0.5 * (1.0 + e1)
        L[:, 1, 0] = 0.0
        L[:, 1, 1] = 0.5 * (1.0 - e2)
    elif _ab(rules[0]) == (0, 0) and _ab(rules[1]) == (1, 0):
        J = 0.5 * np.ones(n)
        L[:, 0, 0] = 1.0 / (1.0 - e2)
        L[:, 0, 1] = 0.5 * (1.0 + e1) / (1.0 - e2)
        L[:, 1, 0] = 0.0
        L[:, 1, 1] = 0.5
    else:
        raise ValueError("Chosen Jacobi weight not supported")
    return J, L


def reference_geometric_factors_tet(rules):
    """(``tensor_simplex.jl:42-82``)."""
    r1, _ = quadrature_line(rules[0])
    r2, _ = quadrature_line(rules[1])
    r3, _ = quadrature_line(rules[2])
    e1, e2, e3 = _grid3(r1, r2, r3)
    n = len(e1)
    L = np.zeros((n, 3, 3))
    if all(_ab(r) == (0, 0) for r in rules):
        J = 0.5 * (1.0 - e2) * (0.5 * (1.0 - e3)) ** 2
        L[:, 0, 0] = 0.5 * (1.0 - e3)
        L[:, 0, 1] = 0.5 * (1.0 + e1) * 0.5 * (1.0 - e3)
        L[:, 0, 2] = 0.5 * (1.0 + e1) * 0.5 * (1.0 - e3)
        L[:, 1, 1] = 0.5 * (1.0 - e2) * 0.5 * (1.0 - e3)
        L[:, 1, 2] = 0.5 * (1.0 + e2) * 0.5 * (1.0 - e2) * 0.5 * (1.0 - e3)
        L[:, 2, 2] = 0.5 * (1.0 - e2) * (0.5 * (1.0 - e3)) ** 2
    elif _ab(rules[0]) == (0, 0) and _ab(rules[1]) == (0, 0) and _ab(rules[2]) == (1, 0):
        J = 0.125 * (1.0 - e2) * (1.0 - e3)
        L[:, 0, 0] = 0.5
        L[:, 0, 1] = 0.25 * (1.0 + e1)
        L[:, 0, 2] = 0.25 * (1.0 + e1)
        L[:, 1, 1] = 0.25 * (1.0 - e2)
        L[:, 1, 2] = 0.125 * (1.0 + e2) * (1.0 - e2)
        L[:, 2, 2] = 0.125 * (1.0 - e2) * (1.0 - e3)
    else:
        raise ValueError("Chosen Jacobi weight not supported")
    return J, L


# ----------------------------------------------------------------------------- warped product
@dataclass
class WarpedTensorProduct:
    """Sum-factorised PKDO Vandermonde (``tensor_simplex.jl:84-140``,
    ``src/MatrixFreeOperators/warped_product_{2d,3d}.jl``).

    A[a1, i], B[a2, i, j], C[a3, i, j, k] (C only in 3D); ``sigma_in`` maps a
    mode multi-index (i, j[, k]) to its 0-based mode number (-1 if absent);
    output node index = a1*M2*M3 + a2*M3 + a3 (last coordinate fastest).
    """
    A: np.ndarray
    B: np.ndarray
    C: Optional[np.ndarray]
    sigma_in: np.ndarray
    M: Tuple[int, ...]
    p: int

    def dense(self) -> np.ndarray:
        p = self.p
        if self.C is None:
            M1, M2 = self.M
            N_p = int((self.sigma_in >= 0).sum())
            V = np.zeros((M1 * M2, N_p))
            for i in range(p + 1):
                for j in range(p + 1 - i):
                    col = self.sigma_in[i, j]
                    V[:, col] = np.outer(self.A[:, i], self.B[:, i, j]).reshape(-1)
            return V
        M1, M2, M3 = self.M
        N_p = int((self.sigma_in >= 0).sum())
        V = np.zeros((M1 * M2 * M3, N_p))
        for i in range(p + 1):
            for j in range(p + 1 - i):
                for k in range(p + 1 - i - j):
                    col = self.sigma_in[i, j, k]
                    V[:, col] = np.einsum("a,b,c->abc", self.A[:, i], self.B[:, i, j],
                                          self.C[:, i, j, k]).reshape(-1)
        return V


def warped_product_tri(p: int, eta1d) -> WarpedTensorProduct:
    M1, M2 = len(eta1d[0]), len(eta1d[1])
    sig = -np.ones((p + 1, p + 1), dtype=np.int64)
    A = np.zeros((M1, p + 1))
    B = np.zeros((M2, p + 1, p + 1))
    k = 0
    for i in range(p + 1):
        for j in range(p + 1 - i):
            sig[i, j] = k
            k += 1
            A[:, i] = np.sqrt(2.0) * P.jacobi_p(eta1d[0], 0.0, 0.0, i)
            B[:, i, j] = (1.0 - eta1d[1]) ** i * P.jacobi_p(eta1d[1], 2.0 * i + 1.0, 0.0, j)
    return WarpedTensorProduct(A, B, None, sig, (M1, M2), p)


def warped_product_tet(p: int, eta1d) -> WarpedTensorProduct:
    M1, M2, M3 = (len(e) for e in eta1d)
    sig = -np.ones((p + 1, p + 1, p + 1), dtype=np.int64)
    A = np.zeros((M1, p + 1))
    B = np.zeros((M2, p + 1, p + 1))
    C = np.zeros((M3, p + 1, p + 1, p + 1))
    l = 0
    for i in range(p + 1):
        for j in range(p + 1 - i):
            for k in range(p + 1 - i - j):
                sig[i, j, k] = l
                l += 1
                A[:, i] = np.sqrt(2.0) * P.jacobi_p(eta1d[0], 0.0, 0.0, i)
                B[:, i, j] = (1.0 - eta1d[1]) ** i * P.jacobi_p(eta1d[1], 2.0 * i + 1.0, 0.0, j)
                C[:, i, j, k] = 2.0 * (1.0 - eta1d[2]) ** (i + j) * P.jacobi_p(
                    eta1d[2], 2.0 * i + 2.0 * j + 2.0, 0.0, k)
    return WarpedTensorProduct(A, B, C, sig, (M1, M2, M3), p)


# ----------------------------------------------------------------------------- 1D operators
def operators_1d(rules):
    """(``tensor_simplex.jl:142-156``)."""
    out = []
    for rule in rules:
        eta, _ = quadrature_line(rule)
        q = len(eta) - 1
        V1 = P.vandermonde_1d(q, eta)
        V1inv = np.linalg.inv(V1)
        D = P.grad_vandermonde_1d(q, eta) @ V1inv
        RL = P.vandermonde_1d(q, [-1.0]) @ V1inv
        RR = P.vandermonde_1d(q, [1.0]) @ V1inv
        out.append(dict(eta=eta, q=q, V=V1, D=D, I=np.eye(q + 1), RL=RL, RR=RR))
    return out


def _interp_1d(rule_from: QuadratureRule, rule_to: QuadratureRule):
    if rule_from == rule_to:
        n = rule_from.q + 1
        return np.eye(n)
    eta, _ = quadrature_line(rule_from)
    eta_to, _ = quadrature_line(rule_to)
    q = len(eta) - 1
    return P.vandermonde_1d(q, eta_to) @ np.linalg.inv(P.vandermonde_1d(q, eta))


# ----------------------------------------------------------------------------- RefElemData
@dataclass
class RefElemData:
    """Mapping-side reference element (``ref_elem_data.jl``; StartUpDG
    ``RefElemData`` for the Line)."""
    element_type: str
    N: int
    rst: Tuple[np.ndarray, ...]
    VDM: np.ndarray
    Drst: Tuple[np.ndarray, ...]
    V1: np.ndarray
    rstq: Tuple[np.ndarray, ...]
    wq: np.ndarray
    Vq: np.ndarray
    rstf: Tuple[np.ndarray, ...]
    wf: np.ndarray
    Vf: np.ndarray
    nrstJ: Tuple[np.ndarray, ...]
    face_vertices: Tuple[Tuple[int, ...], ...]

    @property
    def num_faces(self) -> int:
        return num_faces(self.element_type)


def _tri_face_vertices():
    # face 1: s = -1, face 2: r + s = 0, face 3: r = -1 (ref_elem_data.jl:25-27)
    return ((0, 1), (1, 2), (0, 2))


def _tet_face_vertices():
    # face 1: s = -1, face 2: r+s+t = -1, face 3: r = -1, face 4: t = -1 (ref_elem_data.jl:92-99)
    return ((0, 1, 3), (1, 2, 3), (0, 2, 3), (0, 1, 2))


def _hex_face_vertices():
    # vertices (r, s, t) in {-1, 1}^3 numbered with r fastest; faces r = -1, r = 1,
    # s = -1, s = 1, t = -1, t = 1 (StartUpDG's face order for the Hex)
    return ((0, 2, 4, 6), (1, 3, 5, 7), (0, 1, 4, 5), (2, 3, 6, 7), (0, 1, 2, 3), (4, 5, 6, 7))


def hex_vertices():
    """Reference vertices of the Hex, r fastest: (8, 3)."""
    return np.array([[-1.0 + 2 * (v & 1), -1.0 + 2 * ((v >> 1) & 1), -1.0 + 2 * ((v >> 2) & 1)]
                     for v in range(8)])


def ref_elem_data_line(N: int, rq: np.ndarray, wq: np.ndarray) -> RefElemData:
    r = P.gauss_lobatto_legendre_nodes(N)
    VDM = P.vandermonde_1d(N, r)
    Vinv = np.linalg.inv(VDM)
    Dr = P.grad_vandermonde_1d(N, r) @ Vinv
    V1 = P.vandermonde_1d(1, r) @ np.linalg.inv(P.vandermonde_1d(1, np.array([-1.0, 1.0])))
    rf = np.array([-1.0, 1.0])
    return RefElemData(LINE, N, (r,), VDM, (Dr,), V1, (rq,), wq,
                       P.vandermonde_1d(N, rq) @ Vinv, (rf,), np.ones(2),
                       P.vandermonde_1d(N, rf) @ Vinv, (np.array([-1.0, 1.0]),), ((0,), (1,)))


def ref_elem_data_tri(N: int, volume_rule, facet_rule) -> RefElemData:
    r, s = S.nodes_tri(N)
    VDM, Vr, Vs = S.basis_tri(N, r, s)
    Vinv = np.linalg.inv(VDM)
    r1, s1 = S.nodes_tri(1)
    V1 = S.vandermonde_tri(1, r, s) @ np.linalg.inv(S.vandermonde_tri(1, r1, s1))
    r1d, w1d = quadrature_line(facet_rule)
    one = np.ones_like(r1d)
    wf = np.concatenate([w1d, w1d, w1d])
    rf = np.concatenate([r1d, -r1d, -one])
    sf = np.concatenate([-one, r1d, r1d])
    nrJ = np.concatenate([0 * one, one, -one])
    nsJ = np.concatenate([-one, one, 0 * one])
    rq, sq, wq = quadrature_tri(volume_rule)
    Vq = S.vandermonde_tri(N, rq, sq) @ Vinv
    Vf = S.vandermonde_tri(N, rf, sf) @ Vinv
    return RefElemData(TRI, N, (r, s), VDM, (Vr @ Vinv, Vs @ Vinv), V1, (rq, sq), wq, Vq,
                       (rf, sf), wf, Vf, (nrJ, nsJ), _tri_face_vertices())


def ref_elem_data_tet(N: int, volume_rule, facet_rule) -> RefElemData:
    r, s, t = S.nodes_tet(N)
    VDM, Vr, Vs, Vt = S.basis_tet(N, r, s, t)
    Vinv = np.linalg.inv(VDM)
    r1, s1, t1 = S.nodes_tet(1)
    V1 = S.vandermonde_tet(1, r, s, t) @ np.linalg.inv(S.vandermonde_tet(1, r1, s1, t1))
    r2, s2, w2 = quadrature_tri(facet_rule)
    ee, zz = np.ones_like(r2), np.zeros_like(r2)
    rf = np.concatenate([r2, -(ee + r2 + s2), -ee, r2])
    sf = np.concatenate([-ee, r2, r2, s2])
    tf = np.concatenate([s2, s2, s2, -ee])
    wf = np.concatenate([w2, w2, w2, w2])
    nrJ = np.concatenate([zz, ee, -ee, zz])
    nsJ = np.concatenate([-ee, ee, zz, zz])
    ntJ = np.concatenate([zz, ee, zz, -ee])
    rq, sq, tq, wq = quadrature_tet(volume_rule)
    Vq = S.vandermonde_tet(N, rq, sq, tq) @ Vinv
    Vf = S.vandermonde_tet(N, rf, sf, tf) @ Vinv
    return RefElemData(TET, N, (r, s, t), VDM, (Vr @ Vinv, Vs @ Vinv, Vt @ Vinv), V1,
                       (rq, sq, tq), wq, Vq, (rf, sf, tf), wf, Vf, (nrJ, nsJ, ntJ),
                       _tet_face_vertices())


def ref_elem_data_hex(N: int, volume_rule, facet_rule) -> RefElemData:
    """StartUpDG's ``RefElemData(Hex(), N; quad_rule_vol, quad_rule_face)``: mapping
    nodes the tensor Gauss-Lobatto grid of degree N (the first coordinate slowest,
    as the volume quadrature grid), nodal derivative and interpolation operators
    built from 1D ones (exact for the tensor-product polynomial space), facet nodes
    on the faces r = -1, 1, s = -1, 1, t = -1, 1 from the Quad rule (its first and
    second coordinates are (s, t), (r, t) and (r, s) on those faces)."""
    r1 = P.gauss_lobatto_legendre_nodes(N)
    V1d = P.vandermonde_1d(N, r1)
    V1inv = np.linalg.inv(V1d)
    D1 = P.grad_vandermonde_1d(N, r1) @ V1inv
    I1 = np.eye(N + 1)
    r, s, t = _grid3(r1, r1, r1)
    VDM = _kron(V1d, V1d, V1d)
    Drst = (_kron(D1, I1, I1), _kron(I1, D1, I1), _kron(I1, I1, D1))
    vol = volume_rule if isinstance(volume_rule, tuple) else (volume_rule,) * 3
    x1, w1 = quadrature_line(vol[0])
    x2, w2 = quadrature_line(vol[1])
    x3, w3 = quadrature_line(vol[2])
    rq, sq, tq = _grid3(x1, x2, x3)
    wa, wb, wc = _grid3(w1, w2, w3)
    wq = wa * wb * wc
    interp = lambda x: P.vandermonde_1d(N, x) @ V1inv  # 1D mapping-node interpolation to points x
    Vq = _kron(interp(x1), interp(x2), interp(x3))
    fr = facet_rule if isinstance(facet_rule, tuple) else (facet_rule,) * 2
    y1, v1 = quadrature_line(fr[0])
    y2, v2 = quadrature_line(fr[1])
    a, b = _grid2(y1, y2)
    va, vb = _grid2(v1, v2)
    wq2 = va * vb
    e = np.ones_like(a)
    z = np.zeros_like(a)
    rf = np.concatenate([-e, e, a, a, a, a])
    sf = np.concatenate([a, a, -e, e, b, b])
    tf = np.concatenate([b, b, b, b, -e, e])
    wf = np.concatenate([wq2] * 6)
    nrJ = np.concatenate([-e, e, z, z, z, z])
    nsJ = np.concatenate([z, z, -e, e, z, z])
    ntJ = np.concatenate([z, z, z, z, -e, e])
    L, R = interp([-1.0]), interp([1.0])
    A, B = interp(y1), interp(y2)
    Vf = np.vstack([_kron(L, A, B), _kron(R, A, B), _kron(A, L, B), _kron(A, R, B), _kron(A, B, L),
                    _kron(A, B, R)])
    hv = hex_vertices()
    V1 = np.stack([np.prod([0.5 * (1.0 + hv[v, m] * np.array((r, s, t)[m])) for m in range(3)], axis=0)
                   for v in range(8)], axis=1)
    return RefElemData(HEX, N, (r, s, t), VDM, Drst, V1, (rq, sq, tq), wq, Vq, (rf, sf, tf), wf, Vf,
                       (nrJ, nsJ, ntJ), _hex_face_vertices())


# ----------------------------------------------------------------------------- ReferenceApproximation
@dataclass
class NodalTensor:
    p: int


@dataclass
class ModalTensor:
    p: int


@dataclass
class ModalMulti:
    """Modal basis of total degree p on a volume quadrature of degree 2p, with dense
    operators (``src/SpatialDiscretizations/multidimensional.jl:1-41``)."""
    p: int


@dataclass
class NodalMulti:
    """Nodal (collocated) values at a volume quadrature of degree 2p, with dense
    degree-p operators (``src/SpatialDiscretizations/multidimensional.jl:43-75``)."""
    p: int


@dataclass
class ReferenceMapping:
    J_ref: np.ndarray
    Lambda_ref: np.ndarray  # N_q x d x d


@dataclass
class ReferenceApproximation:
    """(``src/SpatialDiscretizations/SpatialDiscretizations.jl:191-251``).

    ``V`` is None for nodal schemes (the reference's ``UniformScalingMap``);
    ``R_selection`` holds the volume-node index of each facet node for diagonal-E
    operators (the reference's ``SelectionMap``), else None.
    """
    approx_type: object
    reference_element: RefElemData
    D: Tuple[np.ndarray, ...]
    V: Optional[np.ndarray]
    R: np.ndarray
    W: np.ndarray
    B: np.ndarray
    reference_mapping: Optional[ReferenceMapping]
    warped: Optional[WarpedTensorProduct] = None
    R_selection: Optional[np.ndarray] = None
    volume_quadrature_rule: tuple = ()
    facet_quadrature_rule: tuple = ()

    @property
    def N_p(self) -> int:
        return self.D[0].shape[0] if self.V is None else self.V.shape[1]

    @property
    def N_q(self) -> int:
        return self.D[0].shape[0]

    @property
    def N_f(self) -> int:
        return self.R.shape[0]

    @property
    def d(self) -> int:
        return len(self.D)

    @property
    def element_type(self) -> str:
        return self.reference_element.element_type

    @property
    def nodes_per_face(self) -> int:
        return self.N_f // num_faces(self.element_type)

    def Vdense(self) -> np.ndarray:
        return np.eye(self.N_q) if self.V is None else self.V


def _kron(*mats):
    out = mats[0]
    for m in mats[1:]:
        out = np.kron(out, m)
    return out


def ReferenceApproximationLine(approx_type, mapping_degree: int = 1,
                               volume_quadrature_rule: Optional[QuadratureRule] = None):
    """NodalTensor (``tensor_cartesian.jl:1-32``) or ModalMulti
    (``multidimensional.jl:1-41``) on the Line."""
    if isinstance(approx_type, ModalMulti):
        p = approx_type.p
        # DefaultQuadrature(2p) on the Line is LG of degree ceil((2p - 1) / 2)
        # (``quadrature_rules.jl:48-50``)
        rule = volume_quadrature_rule or LGQuadrature(-(-(2 * p - 1) // 2))
        rq, wq = quadrature_line(rule)
        rd = ref_elem_data_line(mapping_degree, rq, wq)
        VDM = P.vandermonde_1d(p, rq)
        dVDM = P.grad_vandermonde_1d(p, rq)
        Vf = P.vandermonde_1d(p, rd.rstf[0])
        Pr = np.linalg.inv(VDM.T @ (wq[:, None] * VDM)) @ (VDM.T * wq[None, :])
        return ReferenceApproximation(approx_type, rd, (dVDM @ Pr,), VDM, Vf @ Pr, wq.copy(),
                                      rd.wf.copy(), None, None, None, (rule,), ())
    rule = volume_quadrature_rule or LGLQuadrature(approx_type.p)
    rq, wq = quadrature_line(rule)
    rd = ref_elem_data_line(mapping_degree, rq, wq)
    q = len(rq) - 1
    VDM = P.vandermonde_1d(q, rq)
    Vinv = np.linalg.inv(VDM)
    D = P.grad_vandermonde_1d(q, rq) @ Vinv
    rf = rd.rstf[0]
    sel = None
    if rule.kind == "lobatto":
        sel = np.array([int(np.argmin(np.abs(rq - x))) for x in rf], dtype=np.int64)
        R = np.zeros((len(rf), q + 1))
        R[np.arange(len(rf)), sel] = 1.0
    else:
        R = P.vandermonde_1d(q, rf) @ Vinv
    return ReferenceApproximation(NodalTensor(q), rd, (D,), None, R, wq.copy(), rd.wf.copy(),
                                  None, None, sel, (rule,), ())


def _multi_approximation(approx_type, rd: RefElemData, vrule, frule):
    """ModalMulti / NodalMulti operators on a Tri or Tet (``multidimensional.jl:1-75``):
    P = (V' W V)^{-1} V' W projects quadrature values onto the degree-p modes, D_m =
    dV_m P, R = Vf P; ModalMulti keeps V (the modal state), NodalMulti the identity.
    There is no reference mapping: D_m is already the derivative in r, s (, t)."""
    p = approx_type.p
    basis = S.basis_tri if rd.element_type == TRI else S.basis_tet
    vand = S.vandermonde_tri if rd.element_type == TRI else S.vandermonde_tet
    VDM, *dV = basis(p, *rd.rstq)
    Vf = vand(p, *rd.rstf)
    wq = rd.wq
    Pr = np.linalg.solve(VDM.T @ (wq[:, None] * VDM), VDM.T * wq[None, :])
    D = tuple(dVm @ Pr for dVm in dV)
    V = VDM if isinstance(approx_type, ModalMulti) else None
    return ReferenceApproximation(approx_type, rd, D, V, Vf @ Pr, wq.copy(), rd.wf.copy(), None,
                                  None, None, tuple(vrule), tuple(frule) if isinstance(frule, tuple)
                                  else (frule,))


def ReferenceApproximationTri(approx_type, mapping_degree: int = 1, volume_quadrature_rule=None,
                              facet_quadrature_rule=None):
    """ModalTensor/NodalTensor on the Tri (``tensor_simplex.jl:158-219``);
    ModalMulti/NodalMulti (``multidimensional.jl:1-75``).

    The multidimensional types default to quadratures of degree 2p. On the facets
    that is the reference's own rule, LG with p + 1 nodes (``quadrature_rules.jl:48-50``).
    In the volume the reference takes StartUpDG's Xiao-Gimbutas table, which is not
    available to this package: the default here is the collapsed rule LG(p) x LG(p),
    also exact to degree 2p. Pass `volume_quadrature_rule` to choose another."""
    p = approx_type.p
    vrule = volume_quadrature_rule or (LGQuadrature(p), LGQuadrature(p))
    frule = facet_quadrature_rule or LGQuadrature(p)
    if isinstance(approx_type, (ModalMulti, NodalMulti)):
        return _multi_approximation(approx_type, ref_elem_data_tri(mapping_degree, vrule, frule),
                                    vrule, frule)
    ops = operators_1d(vrule)
    J_ref, L_ref = reference_geometric_factors_tri(vrule)
    e1_to_f = _interp_1d(vrule[0], frule)
    e2_to_f = _interp_1d(vrule[1], frule)
    R = np.vstack([_kron(e1_to_f, ops[1]["RL"]), _kron(ops[0]["RR"], e2_to_f),
                   _kron(ops[0]["RL"], e2_to_f)])
    rd = ref_elem_data_tri(mapping_degree, vrule, frule)
    D = (_kron(ops[0]["D"], ops[1]["I"]), _kron(ops[0]["I"], ops[1]["D"]))
    if isinstance(approx_type, ModalTensor):
        wp = warped_product_tri(p, (ops[0]["eta"], ops[1]["eta"]))
        V = wp.dense()
    else:
        wp, V = None, None
    return ReferenceApproximation(approx_type, rd, D, V, R, rd.wq.copy(), rd.wf.copy(),
                                  ReferenceMapping(J_ref, L_ref), wp, None, tuple(vrule), (frule,))


def ReferenceApproximationTet(approx_type, mapping_degree: int = 1, volume_quadrature_rule=None,
                              facet_quadrature_rule=None):
    """ModalTensor/NodalTensor on the Tet (``tensor_simplex.jl:221-306``);
    ModalMulti/NodalMulti (``multidimensional.jl:1-75``).

    The multidimensional types default to quadratures of degree 2p. The reference
    takes the Jaskowiec-Sukumar (volume) and Xiao-Gimbutas (facet) tables from
    StartUpDG, which are not available to this package: the defaults here are the
    collapsed rules LG(p) x LG(p) x GJ(p; 1, 0) and LG(p) x GJ(p; 1, 0), exact to
    degree 2p as well."""
    p = approx_type.p
    vrule = volume_quadrature_rule or (LGQuadrature(p), LGQuadrature(p), GaussQuadrature(p, 1, 0))
    frule = facet_quadrature_rule or (LGQuadrature(p), GaussQuadrature(p, 1, 0))
    if isinstance(approx_type, (ModalMulti, NodalMulti)):
        return _multi_approximation(approx_type, ref_elem_data_tet(mapping_degree, vrule, frule),
                                    vrule, frule)
    ops = operators_1d(vrule)
    J_ref, L_ref = reference_geometric_factors_tet(vrule)
    e1_f1 = _interp_1d(vrule[0], frule[0])
    e2_f1 = _interp_1d(vrule[1], frule[0])
    e2_f2 = _interp_1d(vrule[1], frule[1])
    e3_f2 = _interp_1d(vrule[2], frule[1])
    R = np.vstack([_kron(e1_f1, ops[1]["RL"], e3_f2),
                   _kron(ops[0]["RR"], e2_f1, e3_f2),
                   _kron(ops[0]["RL"], e2_f1, e3_f2),
                   _kron(e1_f1, e2_f2, ops[2]["RL"])])
    rd = ref_elem_data_tet(mapping_degree, vrule, frule)
    I1, I2, I3 = ops[0]["I"], ops[1]["I"], ops[2]["I"]
    D = (_kron(ops[0]["D"], I2, I3), _kron(I1, ops[1]["D"], I3), _kron(I1, I2, ops[2]["D"]))
    if isinstance(approx_type, ModalTensor):
        wp = warped_product_tet(p, (ops[0]["eta"], ops[1]["eta"], ops[2]["eta"]))
        V = wp.dense()
    else:
        wp, V = None, None
    return ReferenceApproximation(approx_type, rd, D, V, R, rd.wq.copy(), rd.wf.copy(),
                                  ReferenceMapping(J_ref, L_ref), wp, None, tuple(vrule), tuple(frule))


def ReferenceApproximationHex(approx_type, mapping_degree: int = 1, volume_quadrature_rule=None,
                              facet_quadrature_rule=None):
    """NodalTensor on the Hex (``tensor_cartesian.jl:85-131``): Gauss-Lobatto volume
    and facet rules by default, D = (D1 x I x I, I x D1 x I, I x I x D1), V = I, and
    R a selection of the volume nodes on the faces when both rules are the same
    Gauss-Lobatto rule (diagonal E), else the interpolation to the facet nodes."""
    if not isinstance(approx_type, NodalTensor):
        raise NotImplementedError("the Hex is implemented for NodalTensor")
    p = approx_type.p
    vrule = volume_quadrature_rule or LGLQuadrature(p)
    frule = facet_quadrature_rule or LGLQuadrature(p)
    rd = ref_elem_data_hex(mapping_degree, vrule, frule)
    x1, _ = quadrature_line(vrule)
    q = len(x1) - 1
    V1 = P.vandermonde_1d(q, x1)
    V1inv = np.linalg.inv(V1)
    D1 = P.grad_vandermonde_1d(q, x1) @ V1inv
    I1 = np.eye(q + 1)
    D = (_kron(D1, I1, I1), _kron(I1, D1, I1), _kron(I1, I1, D1))
    rq = np.stack(rd.rstq, axis=1)
    rf = np.stack(rd.rstf, axis=1)
    sel = None
    if vrule == frule and vrule.kind == "lobatto":
        dist = np.abs(rf[:, None, :] - rq[None, :, :]).max(axis=2)
        sel = np.argmin(dist, axis=1).astype(np.int64)
        if dist[np.arange(len(sel)), sel].max() > 1e-12:
            raise ValueError("facet nodes are not volume nodes")
        R = np.zeros((rf.shape[0], rq.shape[0]))
        R[np.arange(len(sel)), sel] = 1.0
    else:
        interp = lambda x: P.vandermonde_1d(q, [x]) @ V1inv  # 1 x (q + 1)
        R = np.vstack([_kron(interp(rf[j, 0]), interp(rf[j, 1]), interp(rf[j, 2])) for j in range(rf.shape[0])])
    return ReferenceApproximation(NodalTensor(q), rd, D, None, R, rd.wq.copy(), rd.wf.copy(), None, None, sel,
                                  (vrule,) * 3, (frule,) * 2)


def ReferenceApproximationFor(approx_type, element_type: str, **kw):
    if element_type == LINE:
        return ReferenceApproximationLine(approx_type, **kw)
    if element_type == TRI:
        return ReferenceApproximationTri(approx_type, **kw)
    if element_type == TET:
        return ReferenceApproximationTet(approx_type, **kw)
    if element_type == HEX:
        return ReferenceApproximationHex(approx_type, **kw)
    raise NotImplementedError(element_type)


# ----------------------------------------------------------------------------- derived operators
def reference_derivative_operators(ref: ReferenceApproximation):
    """D_xi_m = sum_l diag(Lambda_ref[:, l, m] / J_ref) D_eta_l
    (``SpatialDiscretizations.jl:414-422``)."""
    if ref.reference_mapping is None:
        return ref.D
    J = ref.reference_mapping.J_ref
    L = ref.reference_mapping.Lambda_ref
    d = ref.d
    return tuple(sum((L[:, l, m] / J)[:, None] * ref.D[l] for l in range(d)) for m in range(d))


def flux_differencing_operators(ref: ReferenceApproximation):
    """S_m = 1/2 (W D_xi_m - D_xi_m^T W), C = R^T B (``src/Solvers/operators.jl:162-221``).

    Returns (S, S_sparse, C, C_kind) with C_kind 0 = none (diagonal-E), 1 = dense,
    2 = sparse. Tensor-product types with d > 1 use the sparse algorithms.
    """
    D_xi = reference_derivative_operators(ref)
    W = np.diag(ref.W)
    S = np.stack([0.5 * (W @ Dm - Dm.T @ W) for Dm in D_xi])
    sparse = ref.d > 1 and isinstance(ref.approx_type, (ModalTensor, NodalTensor))
    if ref.R_selection is not None:
        return S, sparse, None, 0
    C = ref.R.T @ np.diag(ref.B)
    return S, sparse, C, (2 if sparse else 1)


def check_sbp_property(ref: ReferenceApproximation):
    """max |W D_xi_m + D_xi_m^T W - R^T B diag(n_m) R| (``SpatialDiscretizations.jl:441-455``)."""
    D_xi = reference_derivative_operators(ref)
    W = np.diag(ref.W)
    B = np.diag(ref.B)
    out = []
    for m in range(ref.d):
        E = ref.R.T @ B @ np.diag(ref.reference_element.nrstJ[m]) @ ref.R
        out.append(float(np.abs(W @ D_xi[m] + D_xi[m].T @ W - E).max()))
    return tuple(out)


def face_reference_normals(ref: ReferenceApproximation) -> np.ndarray:
    """n_ref for each face (d x N_fac), taken at the first node of each face as
    in ``src/SpatialDiscretizations/mesh.jl:265``."""
    npf = ref.nodes_per_face
    nf = num_faces(ref.element_type)
    return np.array([[ref.reference_element.nrstJ[m][npf * f] for f in range(nf)] for m in range(ref.d)])
