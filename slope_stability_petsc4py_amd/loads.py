"""Surface loads: a uniform pressure on a strip of the top surface (a strip footing).

The reference's limit-load continuation scales the gravity load only (continuation/limit_load.py:
66-374 with f = f_V); a footing pressure is an additional load vector for the same continuation
(F(U) = t f with f the footing load), optionally on top of a constant gravity load.

Nodal forces of a uniform pressure q on one P2 face (exact integrals of the P2 shape functions over
the face):
  * 2D, 3-node edge of length L: end nodes q L / 6, midside node 2 q L / 3;
  * 3D, 6-node triangle of area A: corner nodes 0, midside nodes q A / 3;
  * P1 faces: q L / 2 per end node, q A / 3 per corner node.
"""

from __future__ import annotations

import numpy as np

from .mesh import SlopeMesh

_EDGES = {2: ((1, 2), (0, 2), (0, 1)), 3: ((0, 1), (1, 2), (0, 2), (1, 3), (2, 3), (0, 3))}
_TET_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


def _midside(a: int, b: int, dim: int) -> int:
    """Local index of the midside node of edge (a, b) in the element's P2 numbering."""
    nv = dim + 1
    for k, (i, j) in enumerate(_EDGES[dim]):
        if (i, j) == (a, b) or (i, j) == (b, a):
            return nv + k
    raise KeyError((a, b))


def strip_footing_load(mesh: SlopeMesh, x_range: tuple[float, float], q: float = 1.0, tol: float = 1e-9):
    """(dim, n_nodes) nodal force array of a downward pressure q (force per unit area, along -y) on
    every top-surface face (all vertices at the largest y) whose vertices lie in x_range; constrained
    dofs are zeroed.  Returns (forces, loaded width along x) -- in 3D the strip spans the mesh's
    whole z extent (a strip footing in plane strain).  P1 and P2 elements."""
    dim = mesh.dim
    nv = dim + 1
    n_p = mesh.elem.shape[0]
    if mesh.elem_type not in ("P1", "P2") or n_p not in ((3, 6) if dim == 2 else (4, 10)):
        raise NotImplementedError("strip_footing_load: P1/P2 elements only")
    p2 = n_p > nv
    coord = mesh.coord
    y_top = coord[1].max()
    x0, x1 = x_range
    vert = mesh.elem[:nv]
    on_top = np.abs(coord[1] - y_top) <= tol * max(1.0, abs(y_top))
    in_x = (coord[0] >= x0 - tol) & (coord[0] <= x1 + tol)
    sel = on_top & in_x
    F = np.zeros((dim, mesh.n_nodes))
    xs_loaded = []
    faces = [(0, 1), (1, 2), (0, 2)] if dim == 2 else _TET_FACES
    for face in faces:
        fv = vert[list(face)]                                # (dim, n_elem) vertex ids of this local face
        hit = np.all(sel[fv], axis=0)
        if not hit.any():
            continue
        ids = np.flatnonzero(hit)
        P = coord[:, fv[:, ids]]                             # (dim, dim, n_hit) coordinates
        if dim == 2:
            meas = np.abs(P[0, 1] - P[0, 0])                 # edge length (the edge lies at y = y_top)
        else:
            ex, ez = P[0, 1] - P[0, 0], P[2, 1] - P[2, 0]
            fx, fz = P[0, 2] - P[0, 0], P[2, 2] - P[2, 0]
            meas = 0.5 * np.abs(ex * fz - ez * fx)           # area in the x-z plane
        xs_loaded.append(P[0].reshape(-1))
        if dim == 2:
            a, b = face
            if p2:
                m = mesh.elem[_midside(a, b, 2), ids]
                np.add.at(F[1], fv[0, ids], -q * meas / 6.0)
                np.add.at(F[1], fv[1, ids], -q * meas / 6.0)
                np.add.at(F[1], m, -2.0 * q * meas / 3.0)
            else:
                np.add.at(F[1], fv[0, ids], -q * meas / 2.0)
                np.add.at(F[1], fv[1, ids], -q * meas / 2.0)
        else:
            if p2:
                for (a, b) in ((face[0], face[1]), (face[1], face[2]), (face[0], face[2])):
                    np.add.at(F[1], mesh.elem[_midside(a, b, 3), ids], -q * meas / 3.0)
            else:
                for c in range(3):
                    np.add.at(F[1], fv[c, ids], -q * meas / 3.0)
    if not xs_loaded:
        raise ValueError("strip_footing_load: no top-surface face inside x_range")
    xs = np.concatenate(xs_loaded)
    F[~np.asarray(mesh.q_mask, dtype=bool)] = 0.0
    return F, float(xs.max() - xs.min())
