"""Periodic uniform meshes, mesh warping and face connectivity (``mapP``).

Restates ``src/SpatialDiscretizations/mesh.jl:1-209`` of the reference, plus
the StartUpDG.jl pieces it calls (compat "1", not vendored): ``uniform_mesh``,
``MeshData`` (mapping nodes x = V1 * VX[EToV]', quadrature/facet node
positions) and ``make_periodic`` (facet-node matching across periodic
boundaries). The tetrahedral split of a hexahedron is the 6-tet Kuhn
(Freudenthal) split along the main diagonal; element vertex order then follows
the reference's collapsed orientation (``mesh.jl:150-169``: vertices sorted by
descending global index, first two swapped if the frame is left-handed).

``mapP`` keeps the reference convention: a 1-based linear index into the
column-major (N_f, N_e) facet array, stored as int64 with shape (N_f, N_e).
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from itertools import permutations
from typing import Optional, Tuple

import numpy as np

from .reference_approximation import HEX, LINE, TET, TRI, RefElemData


@dataclass
class MeshData:
    element_type: str
    VXYZ: Tuple[np.ndarray, ...]
    EToV: np.ndarray                     # N_e x N_vert, 0-based
    xyz: Tuple[np.ndarray, ...]          # mapping nodes, N_nodes x N_e each
    xyzq: Tuple[np.ndarray, ...]         # N_q x N_e
    xyzf: Tuple[np.ndarray, ...]         # N_f x N_e
    mapP: np.ndarray                     # N_f x N_e, int64, 1-based linear index
    FToF: np.ndarray                     # N_fac x N_e, 0-based neighbour element
    periodic_L: Tuple[float, ...]

    @property
    def N_e(self) -> int:
        return self.EToV.shape[0]

    @property
    def d(self) -> int:
        return len(self.xyz)


# ----------------------------------------------------------------------------- vertex meshes
def uniform_mesh_line(K: int):
    VX = np.linspace(-1.0, 1.0, K + 1)
    EToV = np.stack([np.arange(K), np.arange(1, K + 1)], axis=1)
    return (VX,), EToV


def uniform_mesh_tri(Kx: int, Ky: int, diagonal: str = "bl-tr"):
    """Each square split into two counter-clockwise triangles. Vertex id of
    (ix, iy) is ix + (Kx+1) iy. ``diagonal`` selects the splitting diagonal."""
    xs = np.linspace(-1.0, 1.0, Kx + 1)
    ys = np.linspace(-1.0, 1.0, Ky + 1)
    VX = np.tile(xs, Ky + 1)
    VY = np.repeat(ys, Kx + 1)
    tris = []
    for ey in range(Ky):
        for ex in range(Kx):
            i1 = ex + (Kx + 1) * ey
            i2 = i1 + 1
            i4 = i1 + (Kx + 1)
            i3 = i4 + 1
            if diagonal == "bl-tr":
                tris.append((i1, i2, i3))
                tris.append((i3, i4, i1))
            else:
                tris.append((i1, i2, i4))
                tris.append((i3, i4, i2))
    return (VX, VY), np.array(tris, dtype=np.int64)


def uniform_mesh_tet(Kx: int, Ky: int, Kz: int):
    """Hexahedra split into 6 tetrahedra (Kuhn split along the (0,0,0)-(1,1,1)
    diagonal); vertex id of (ix, iy, iz) is ix + (Kx+1)(iy + (Ky+1) iz)."""
    xs = np.linspace(-1.0, 1.0, Kx + 1)
    ys = np.linspace(-1.0, 1.0, Ky + 1)
    zs = np.linspace(-1.0, 1.0, Kz + 1)
    IZ, IY, IX = np.meshgrid(np.arange(Kz + 1), np.arange(Ky + 1), np.arange(Kx + 1), indexing="ij")
    VX = xs[IX.reshape(-1)]
    VY = ys[IY.reshape(-1)]
    VZ = zs[IZ.reshape(-1)]

    def vid(ix, iy, iz):
        return ix + (Kx + 1) * (iy + (Ky + 1) * iz)

    ez, ey, ex = np.meshgrid(np.arange(Kz), np.arange(Ky), np.arange(Kx), indexing="ij")
    ex, ey, ez = ex.reshape(-1), ey.reshape(-1), ez.reshape(-1)
    tets = []
    for perm in permutations(range(3)):
        off = np.zeros((len(ex), 3), dtype=np.int64)
        verts = [vid(ex, ey, ez)]
        for axis in perm:
            off[:, axis] += 1
            verts.append(vid(ex + off[:, 0], ey + off[:, 1], ez + off[:, 2]))
        tets.append(np.stack(verts, axis=1))
    EToV = np.stack(tets, axis=1).reshape(-1, 4)  # 6 consecutive tets per hex
    return (VX, VY, VZ), EToV


def uniform_mesh_hex(Kx: int, Ky: int, Kz: int):
    """Hexahedra of a uniform grid on [-1, 1]^3; vertex id of (ix, iy, iz) is
    ix + (Kx+1)(iy + (Ky+1) iz), element vertices in the reference order (r fastest)."""
    xs = np.linspace(-1.0, 1.0, Kx + 1)
    ys = np.linspace(-1.0, 1.0, Ky + 1)
    zs = np.linspace(-1.0, 1.0, Kz + 1)
    IZ, IY, IX = np.meshgrid(np.arange(Kz + 1), np.arange(Ky + 1), np.arange(Kx + 1), indexing="ij")
    VX, VY, VZ = xs[IX.reshape(-1)], ys[IY.reshape(-1)], zs[IZ.reshape(-1)]
    ez, ey, ex = np.meshgrid(np.arange(Kz), np.arange(Ky), np.arange(Kx), indexing="ij")
    ex, ey, ez = ex.reshape(-1), ey.reshape(-1), ez.reshape(-1)
    verts = [(ex + (v & 1)) + (Kx + 1) * ((ey + ((v >> 1) & 1)) + (Ky + 1) * (ez + ((v >> 2) & 1)))
             for v in range(8)]
    return (VX, VY, VZ), np.stack(verts, axis=1).astype(np.int64)


def collapsed_orientation(VXYZ, EToV):
    """Second algorithm of Warburton's thesis as used in ``mesh.jl:150-169``."""
    E = np.sort(EToV, axis=1)[:, ::-1].copy()
    X = np.stack([np.stack([VXYZ[c][E[:, m]] - VXYZ[c][E[:, 0]] for c in range(3)], axis=1)
                  for m in (1, 2, 3)], axis=2)
    det = np.linalg.det(X)
    swap = det < 0
    E[swap, 0], E[swap, 1] = E[swap, 1].copy(), E[swap, 0].copy()
    return E


# ----------------------------------------------------------------------------- MeshData
def _node_positions(rd: RefElemData, VXYZ, EToV):
    return tuple(rd.V1 @ VXYZ[m][EToV].T for m in range(len(VXYZ)))


def _periodic_key(coords, L, tol=1e-8):
    """Integer key of a point modulo the periodic box."""
    c = np.stack([np.mod(coords[m], L[m]) for m in range(len(L))], axis=-1)
    c = np.where(np.abs(c - np.array(L)) < tol, 0.0, c)
    return np.round(c / tol).astype(np.int64)


def make_mesh(rd: RefElemData, VXYZ, EToV, limits) -> MeshData:
    """Periodic ``MeshData`` on the box ``limits`` (((lo, hi), ...) per dimension)."""
    d = len(VXYZ)
    VXYZ = tuple(limits[m][0] + 0.5 * (limits[m][1] - limits[m][0]) * (VXYZ[m] + 1.0) for m in range(d))
    L = tuple(float(limits[m][1] - limits[m][0]) for m in range(d))
    lo = tuple(float(limits[m][0]) for m in range(d))
    xyz = _node_positions(rd, VXYZ, EToV)
    xyzq = tuple(rd.Vq @ xyz[m] for m in range(d))
    xyzf = tuple(rd.Vf @ xyz[m] for m in range(d))
    N_e = EToV.shape[0]
    nfac = len(rd.face_vertices)
    N_f = xyzf[0].shape[0]
    npf = N_f // nfac

    # face matching on the face centroid taken modulo the periodic box
    fv = np.array(rd.face_vertices, dtype=np.int64)          # nfac x nv_face
    cen = [VXYZ[m][EToV[:, fv]].mean(axis=2).reshape(-1) - lo[m] for m in range(d)]
    # centroids of distinct faces are a fraction of an element apart; the key's cell
    # (1e-6 of the box) is far above round-off and far below that separation
    flat = _periodic_key(cen, L, tol=1e-6 * max(L))
    _, inv, counts = np.unique(flat, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    if not np.all(counts == 2):
        raise ValueError("mesh is not periodic-conforming (a face has %d owners)" % counts.max())
    order = np.argsort(inv, kind="stable")
    pair_a = order[0::2]
    pair_b = order[1::2]
    FToF = np.empty(N_e * nfac, dtype=np.int64)
    FToF[pair_a] = pair_b // nfac
    FToF[pair_b] = pair_a // nfac

    # facet-node matching (periodic distance) for each matched face pair
    mapP = np.empty((N_f, N_e), dtype=np.int64)
    fx = np.stack([xyzf[m].T.reshape(N_e, nfac, npf) for m in range(d)], axis=-1)  # N_e x nfac x npf x d
    Larr = np.array(L)
    # The two faces of a pair carry the same node set. Shift face b by its periodic
    # image onto face a, order both faces' nodes along a generic direction and pair
    # them in that order; the distances are then checked.
    wdir = np.array([1.0, 2.0 ** 0.5, 3.0 ** 0.5][:d])
    for (ga, gb) in ((pair_a, pair_b), (pair_b, pair_a)):
        ka, fa = ga // nfac, ga % nfac
        kb, fb = gb // nfac, gb % nfac
        xa = fx[ka, fa]               # P x npf x d
        xb = fx[kb, fb]
        shift = Larr * np.round((xa.mean(axis=1) - xb.mean(axis=1)) / Larr)
        xb = xb + shift[:, None, :]
        ia = np.argsort(xa @ wdir, axis=1, kind="stable")
        ib = np.argsort(xb @ wdir, axis=1, kind="stable")
        jb = np.empty_like(ia)
        np.put_along_axis(jb, ia, ib, axis=1)  # node ia[t] of face a pairs with node ib[t] of face b
        gap = np.sqrt(((xa - np.take_along_axis(xb, jb[:, :, None], axis=1)) ** 2).sum(-1))
        if gap.max() > 1e-8 * max(L):
            raise ValueError("facet nodes do not conform across a face")
        rows = fa[:, None] * npf + np.arange(npf)[None, :]
        nbr = (fb[:, None] * npf + jb) + N_f * kb[:, None] + 1
        mapP[rows, ka[:, None]] = nbr
    return MeshData(rd.element_type, VXYZ, EToV, xyz, xyzq, xyzf, mapP,
                    FToF.reshape(N_e, nfac).T.copy(), L)


def uniform_periodic_mesh(ref, limits, M, tri_diagonal: str = "bl-tr") -> MeshData:
    """(``src/SpatialDiscretizations/mesh.jl:122-179, 508-545``)."""
    rd = ref.reference_element
    elem = rd.element_type
    if elem == LINE:
        VXYZ, EToV = uniform_mesh_line(M if np.isscalar(M) else M[0])
        limits = (limits,) if np.isscalar(limits[0]) else limits
    elif elem == TRI:
        VXYZ, EToV = uniform_mesh_tri(M[0], M[1], tri_diagonal)
    elif elem == TET:
        VXYZ, EToV = uniform_mesh_tet(M[0], M[1], M[2])
        EToV = collapsed_orientation(VXYZ, EToV)
    elif elem == HEX:
        VXYZ, EToV = uniform_mesh_hex(M[0], M[1], M[2])
    else:
        raise NotImplementedError(elem)
    return make_mesh(rd, VXYZ, EToV, limits)


# ----------------------------------------------------------------------------- warping
@dataclass
class UniformWarping:
    """(``mesh.jl:17-21, 78-104``)."""
    factor: float
    L: Tuple[float, ...]


@dataclass
class ChanWarping:
    """(``mesh.jl:10-15, 43-76``)."""
    factor: float
    L: Tuple[float, ...]


@dataclass
class DelReyWarping:
    """(``mesh.jl:6-9, 30-41, 106-120``)."""
    factor: float
    L: Tuple[float, ...]


def _warp_coords(xyz, w):
    f, L = w.factor, w.L
    pi = np.pi
    if isinstance(w, UniformWarping):
        if len(xyz) == 2:
            x, y = xyz
            eps = f * np.sin(2 * pi * (x - L[0] / 2) / L[0]) * np.sin(2 * pi * (y - L[1] / 2) / L[1])
            return (x + L[0] * eps, y + L[1] * eps)
        x, y, z = xyz
        eps = (f * np.sin(2 * pi * (x - L[0] / 2) / L[0]) * np.sin(2 * pi * (y - L[1] / 2) / L[1])
               * np.sin(2 * pi * (z - L[2] / 2) / L[2]))
        return (x + L[0] * eps, y + L[1] * eps, z + L[2] * eps)
    if isinstance(w, ChanWarping):
        if len(xyz) == 2:
            x, y = xyz
            xn = x + L[0] * f * np.cos(pi / L[0] * (x - 0.5 * L[0])) * np.cos(3 * pi / L[1] * (y - 0.5 * L[1]))
            yn = y + L[1] * f * np.sin(4 * pi / L[0] * (xn - 0.5 * L[0])) * np.cos(pi / L[1] * (y - 0.5 * L[1]))
            return (xn, yn)
        x, y, z = xyz
        yn = y + L[1] * f * (np.cos(3 * pi / L[0] * (x - 0.5 * L[0])) * np.cos(pi / L[1] * (y - 0.5 * L[1]))
                             * np.cos(pi / L[2] * (z - 0.5 * L[2])))
        xn = x + L[0] * f * (np.cos(pi / L[0] * (x - 0.5 * L[0])) * np.sin(4 * pi / L[1] * (yn - 0.5 * L[1]))
                             * np.cos(pi / L[2] * (z - 0.5 * L[2])))
        zn = z + L[2] * f * (np.cos(pi / L[0] * (xn - 0.5 * L[0])) * np.cos(2 * pi / L[1] * (yn - 0.5 * L[1]))
                             * np.cos(pi / L[2] * (z - 0.5 * L[2])))
        return (xn, yn, zn)
    if isinstance(w, DelReyWarping):
        if len(xyz) == 2:
            x, y = xyz
            xn = x + L[0] * f * np.sin(pi * x / L[0]) * np.sin(pi * y / L[1])
            yn = y + L[1] * f * np.exp(1.0 - y / L[1]) * np.sin(pi * x / L[0]) * np.sin(pi * y / L[1])
            return (xn, yn)
        x, y, z = xyz
        xn = x + L[0] * f * np.sin(pi * x / L[0]) * np.sin(pi * y / L[1])
        yn = y + L[1] * f * np.exp((1.0 - y) / L[1]) * np.sin(pi * x / L[0]) * np.sin(pi * y / L[1])
        zn = z + 0.25 * L[2] * f * (np.sin(2 * pi * x / L[0]) * np.sin(2 * pi * y / L[1])) * np.sin(2 * pi * z / L[2])
        return (xn, yn, zn)
    raise TypeError(w)


def warp_mesh(mesh: MeshData, ref, warping, L: float = 1.0) -> MeshData:
    """Displace the mapping nodes; connectivity is kept (``mesh.jl:23-120``). A
    number as ``warping`` is ``DelReyWarping(factor, (L,) * d)`` (``mesh.jl:23-28``)."""
    if isinstance(warping, (int, float)):
        warping = DelReyWarping(float(warping), (float(L),) * len(mesh.xyz))
    rd = ref.reference_element
    xyz = _warp_coords(mesh.xyz, warping)
    return replace(mesh, xyz=xyz, xyzq=tuple(rd.Vq @ x for x in xyz),
                   xyzf=tuple(rd.Vf @ x for x in xyz))
