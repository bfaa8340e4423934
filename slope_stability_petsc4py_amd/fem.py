"""Finite-element geometry on simplices: quadrature rules, Lagrange bases and the per-point
physical gradients (dphi) and weights that every assembly kernel consumes.

Setup-time host code (runs once per mesh).  The rules, node orderings and formulas are the
reference's, so that dphi/w agree with it to rounding:

* quadrature tables: fem/quadrature.py:7-165 (same decimal values);
* P1/P2 bases: fem/basis.py:9-264; P4 tetrahedra: core/simplex_lagrange.py:49-161 (barycentric
  node tuples: vertices, edges in (01,12,02,13,23,03) order, faces, interior);
* geometry: fem/assembly.py:50-136 (2D), :139-245 (3D cofactor inverse, w = wf*|det J|).

Output layout (what the device kernels take): ``dphi`` is (dim, n_int, n_p) float64 and ``weight``
is (n_int,), with point index g = e*n_q + q.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np

# --------------------------------------------------------------------------------------------
# Quadrature (fem/quadrature.py)
# --------------------------------------------------------------------------------------------

_Q2D_P2_A = 0.1012865073235
_Q2D_P2_B = 0.7974269853531
_Q2D_P2_C = 0.4701420641051
_Q2D_P2_D = 0.0597158717898


def quadrature_2d(elem_type: str) -> tuple[np.ndarray, np.ndarray]:
    """Triangle rules on the unit reference triangle: (xi (2, n_q), wf (n_q,))."""
    t = elem_type.upper()
    if t == "P1":
        return np.array([[1 / 3], [1 / 3]]), np.array([0.5])
    if t == "P2":
        a, b, c, d = _Q2D_P2_A, _Q2D_P2_B, _Q2D_P2_C, _Q2D_P2_D
        xi = np.array([[a, b, a, c, c, d, 1 / 3], [a, a, b, d, c, c, 1 / 3]], dtype=np.float64)
        wf = np.array([0.1259391805448] * 3 + [0.1323941527885] * 3 + [0.225], dtype=np.float64) / 2
        return xi, wf
    if t == "P4":
        p = [0.063089014491502, 0.873821971016996, 0.249286745170910, 0.501426509658179,
             0.310352451033785, 0.053145049844816, 0.636502499121399]
        xi = np.array(
            [[p[0], p[0], p[1], p[2], p[2], p[3], p[4], p[4], p[5], p[5], p[6], p[6]],
             [p[0], p[1], p[0], p[2], p[3], p[2], p[5], p[6], p[4], p[6], p[4], p[5]]],
            dtype=np.float64,
        )
        wf = np.array([0.050844906370207] * 3 + [0.116786275726379] * 3 + [0.082851075618374] * 6) / 2
        return xi, wf
    raise ValueError(f"Unsupported elem_type={elem_type}")


def quadrature_3d(elem_type: str) -> tuple[np.ndarray, np.ndarray]:
    """Tetrahedron rules on the unit reference tetrahedron: (xi (3, n_q), wf (n_q,))."""
    t = elem_type.upper()
    if t == "P1":
        return np.array([[0.25], [0.25], [0.25]]), np.array([1 / 6])
    if t == "P2":
        # 11-point rule with one negative weight (Keast), exact to degree 4.
        a, b = 0.0714285714285714, 0.785714285714286
        c, d = 0.399403576166799, 0.100596423833201
        xi = np.array(
            [[0.25, a, b, a, a, c, d, d, c, c, d],
             [0.25, a, a, b, a, d, c, d, c, d, c],
             [0.25, a, a, a, b, d, d, c, d, c, c]],
            dtype=np.float64,
        )
        wf = np.array([-0.013155555555555] + [0.007622222222222] * 4 + [0.024888888888888] * 6, dtype=np.float64)
        return xi, wf
    if t == "P4":
        # Keast order-7 rule: 24 points, exact to degree 6.
        a, b = 0.3561913862225449, 0.2146028712591517
        c, d = 0.8779781243961660, 0.0406739585346113
        e, f = 0.0329863295731731, 0.3223378901422757
        g, h, k = 0.2696723314583159, 0.0636610018750175, 0.6030056647916491
        xi = np.array(
            [[a, b, b, b, c, d, d, d, e, f, f, f, g, h, h, k, h, h, h, g, k, h, g, k],
             [b, b, b, a, d, d, d, c, f, f, f, e, h, g, h, h, k, h, g, k, h, k, h, g],
             [b, b, a, b, d, d, c, d, f, f, e, f, h, h, g, h, h, k, k, h, g, g, k, h]],
            dtype=np.float64,
        )
        wf = np.array([0.0399227502581679] * 4 + [0.0100772110553207] * 4 + [0.0553571815436544] * 4
                      + [0.0482142857142857] * 12, dtype=np.float64) / 6.0
        return xi, wf
    raise ValueError(f"Unsupported elem_type={elem_type}")


# --------------------------------------------------------------------------------------------
# Lagrange bases
# --------------------------------------------------------------------------------------------

_TRI_EDGES = ((0, 1), (1, 2), (0, 2))
_TET_EDGES = ((0, 1), (1, 2), (0, 2), (1, 3), (2, 3), (0, 3))
_TET_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


@lru_cache(maxsize=None)
def triangle_node_tuples(p: int) -> tuple[tuple[int, int, int], ...]:
    """Barycentric multi-indices of the order-p triangle nodes (vertices, edges, interior)."""
    out = [(p, 0, 0), (0, p, 0), (0, 0, p)]
    for i, j in _TRI_EDGES:
        for k in range(1, p):
            c = [0, 0, 0]
            c[i], c[j] = p - k, k
            out.append(tuple(c))
    for a in range(p - 2, 0, -1):
        for b in range(p - a - 1, 0, -1):
            if p - a - b > 0:
                out.append((a, b, p - a - b))
    return tuple(out)


@lru_cache(maxsize=None)
def tetra_node_tuples(p: int) -> tuple[tuple[int, int, int, int], ...]:
    """Barycentric multi-indices of the order-p tetrahedron nodes (core/simplex_lagrange.py:49-83)."""
    out = [(p, 0, 0, 0), (0, p, 0, 0), (0, 0, p, 0), (0, 0, 0, p)]
    for i, j in _TET_EDGES:
        for k in range(1, p):
            c = [0, 0, 0, 0]
            c[i], c[j] = p - k, k
            out.append(tuple(c))
    interior_tri = triangle_node_tuples(p)[3 + 3 * (p - 1):] if p >= 3 else ()
    for face in _TET_FACES:
        for t in interior_tri:
            c = [0, 0, 0, 0]
            for li, gi in enumerate(face):
                c[gi] = t[li]
            out.append(tuple(c))
    for a in range(p - 3, 0, -1):
        for b in range(p - a - 2, 0, -1):
            for c3 in range(p - a - b - 1, 0, -1):
                if p - a - b - c3 > 0:
                    out.append((a, b, c3, p - a - b - c3))
    return tuple(out)


def _factor(p: int, m: int, lam: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """prod_{k<m} (p*lam - k)/(k+1) and its derivative in lam."""
    if m == 0:
        return np.ones_like(lam), np.zeros_like(lam)
    fs = [(p * lam - k) / (k + 1) for k in range(m)]
    val = np.ones_like(lam)
    for f in fs:
        val = val * f
    der = np.zeros_like(lam)
    for k in range(m):
        term = np.full_like(lam, p / (k + 1))
        for n, f in enumerate(fs):
            if n != k:
                term = term * f
        der = der + term
    return val, der


def tetra_lagrange_basis(p: int, xi: np.ndarray):
    """(hatp, dhat1, dhat2, dhat3), each (n_p, n_q), for the order-p tetrahedron
    (core/simplex_lagrange.py:116-161)."""
    x1, x2, x3 = xi
    bary = (1.0 - x1 - x2 - x3, x1, x2, x3)
    tuples = tetra_node_tuples(p)
    n_p, n_q = len(tuples), xi.shape[1]
    hatp = np.empty((n_p, n_q))
    d = [np.empty((n_p, n_q)) for _ in range(3)]
    for idx, counts in enumerate(tuples):
        vals, ders = zip(*(_factor(p, m, lam) for lam, m in zip(bary, counts)))
        dl = []
        for a in range(4):
            t = ders[a]
            for b in range(4):
                if b != a:
                    t = t * vals[b]
            dl.append(t)
        hatp[idx] = vals[0] * vals[1] * vals[2] * vals[3]
        for c in range(3):
            d[c][idx] = dl[c + 1] - dl[0]
    return hatp, d[0], d[1], d[2]


def basis_3d(elem_type: str, xi: np.ndarray):
    t = elem_type.upper()
    x1, x2, x3 = xi
    n_q = xi.shape[1]
    if t == "P1":
        hatp = np.array([1 - x1 - x2 - x3, x1, x2, x3])
        z = np.zeros(n_q)
        return (hatp, np.array([-1 + z, 1 + z, z, z]), np.array([-1 + z, z, 1 + z, z]),
                np.array([-1 + z, z, z, 1 + z]))
    if t == "P2":
        x0 = 1 - x1 - x2 - x3
        z = np.zeros(n_q)
        hatp = np.array([x0 * (2 * x0 - 1), x1 * (2 * x1 - 1), x2 * (2 * x2 - 1), x3 * (2 * x3 - 1),
                         4 * x0 * x1, 4 * x1 * x2, 4 * x0 * x2, 4 * x1 * x3, 4 * x2 * x3, 4 * x0 * x3])
        d1 = np.array([-4 * x0 + 1, 4 * x1 - 1, z, z, 4 * (x0 - x1), 4 * x2, -4 * x2, 4 * x3, z, -4 * x3])
        d2 = np.array([-4 * x0 + 1, z, 4 * x2 - 1, z, -4 * x1, 4 * x1, 4 * (x0 - x2), z, 4 * x3, -4 * x3])
        d3 = np.array([-4 * x0 + 1, z, z, 4 * x3 - 1, -4 * x1, z, -4 * x2, 4 * x1, 4 * x2, 4 * (x0 - x3)])
        return hatp, d1, d2, d3
    if t == "P4":
        return tetra_lagrange_basis(4, xi)
    raise ValueError(f"Unsupported 3D elem_type={elem_type}")


def basis_2d(elem_type: str, xi: np.ndarray):
    t = elem_type.upper()
    x1, x2 = xi
    n_q = xi.shape[1]
    z = np.zeros(n_q)
    if t == "P1":
        return np.array([1 - x1 - x2, x1, x2]), np.array([-1 + z, 1 + z, z]), np.array([-1 + z, z, 1 + z])
    if t == "P2":
        x0 = 1 - x1 - x2
        hatp = np.array([x0 * (2 * x0 - 1), x1 * (2 * x1 - 1), x2 * (2 * x2 - 1), 4 * x1 * x2, 4 * x0 * x2,
                         4 * x0 * x1])
        d1 = np.array([-4 * x0 + 1, 4 * x1 - 1, z, 4 * x2, -4 * x2, 4 * (x0 - x1)])
        d2 = np.array([-4 * x0 + 1, z, 4 * x2 - 1, 4 * x1, 4 * (x0 - x2), -4 * x1])
        return hatp, d1, d2
    raise ValueError(f"Unsupported 2D elem_type={elem_type}")


NODES_PER_ELEMENT = {(2, "P1"): 3, (2, "P2"): 6, (2, "P4"): 15, (3, "P1"): 4, (3, "P2"): 10, (3, "P4"): 35}


@dataclass
class Geometry:
    """Per-point geometry of a simplex mesh."""

    dim: int
    elem_type: str
    n_elem: int
    n_q: int
    n_p: int
    dphi: np.ndarray      # (dim, n_int, n_p)
    weight: np.ndarray    # (n_int,)
    hatp: np.ndarray      # (n_p, n_q) reference basis values at the points

    @property
    def n_int(self) -> int:
        return self.n_elem * self.n_q


def geometry(coord: np.ndarray, elem: np.ndarray, elem_type: str, quad: tuple | None = None) -> Geometry:
    """dphi and weights for every (element, point); coord (dim, n_nodes), elem (n_p, n_elem).
    ``quad`` = (xi, wf) overrides the element's own rule (a coarser basis at a finer level's points)."""
    coord = np.asarray(coord, dtype=np.float64)
    elem = np.asarray(elem, dtype=np.int64)
    dim = coord.shape[0]
    n_p, n_e = elem.shape
    if dim == 3:
        xi, wf = quadrature_3d(elem_type) if quad is None else quad
        hatp, h1, h2, h3 = basis_3d(elem_type, xi)
        n_q = xi.shape[1]
        cx, cy, cz = coord[0, elem], coord[1, elem], coord[2, elem]  # (n_p, n_e)
        # J_ab = sum_i x_b(node i) * dhat_a(i, q): shapes (n_e, n_q)
        j11, j12, j13 = cx.T @ h1, cy.T @ h1, cz.T @ h1
        j21, j22, j23 = cx.T @ h2, cy.T @ h2, cz.T @ h2
        j31, j32, j33 = cx.T @ h3, cy.T @ h3, cz.T @ h3
        det = j11 * (j22 * j33 - j23 * j32) - j12 * (j21 * j33 - j23 * j31) + j13 * (j21 * j32 - j22 * j31)
        inv = 1.0 / det
        H1, H2, H3 = h1.T[None], h2.T[None], h3.T[None]  # (1, n_q, n_p)
        c = lambda a: a[:, :, None]  # noqa: E731
        dx = (c(j22 * j33 - j23 * j32) * H1 - c(j12 * j33 - j13 * j32) * H2 + c(j12 * j23 - j13 * j22) * H3) * c(inv)
        dy = (-c(j21 * j33 - j23 * j31) * H1 + c(j11 * j33 - j13 * j31) * H2 - c(j11 * j23 - j13 * j21) * H3) * c(inv)
        dz = (c(j21 * j32 - j22 * j31) * H1 - c(j11 * j32 - j12 * j31) * H2 + c(j11 * j22 - j12 * j21) * H3) * c(inv)
        dphi = np.stack([dx.reshape(n_e * n_q, n_p), dy.reshape(n_e * n_q, n_p), dz.reshape(n_e * n_q, n_p)])
        weight = (np.abs(det) * wf[None, :]).reshape(-1)
    elif dim == 2:
        xi, wf = quadrature_2d(elem_type)
        hatp, h1, h2 = basis_2d(elem_type, xi)
        n_q = xi.shape[1]
        cx, cy = coord[0, elem], coord[1, elem]
        j11, j12 = cx.T @ h1, cy.T @ h1
        j21, j22 = cx.T @ h2, cy.T @ h2
        det = j11 * j22 - j12 * j21
        H1, H2 = h1.T[None], h2.T[None]
        c = lambda a: a[:, :, None]  # noqa: E731
        dx = (c(j22) * H1 - c(j12) * H2) / c(det)
        dy = (-c(j21) * H1 + c(j11) * H2) / c(det)
        dphi = np.stack([dx.reshape(n_e * n_q, n_p), dy.reshape(n_e * n_q, n_p)])
        weight = (np.abs(det) * wf[None, :]).reshape(-1)
    else:
        raise ValueError("dim must be 2 or 3")
    return Geometry(dim=dim, elem_type=elem_type.upper(), n_elem=n_e, n_q=n_q, n_p=n_p,
                    dphi=np.ascontiguousarray(dphi), weight=np.ascontiguousarray(weight), hatp=hatp)


@dataclass
class AffineGeometry:
    """Constant-Jacobian form of a 3D mesh's geometry: with G = gm[e],
    dphi_d(e, q, n) = sum_k G[3 d + k] h_k(q, n) and w(e, q) = G[9] wf[q]."""

    gm: np.ndarray    # (n_elem, 10): (J^-1)[d, k] at 3 d + k (cofactors / det J), |det J| at 9
    ref: np.ndarray   # (3 n_q n_p + n_q,): h_k(q, n) at (k n_q + q) n_p + n, then wf


def affine_geometry(coord: np.ndarray, elem: np.ndarray, elem_type: str, rtol: float = 1e-10) -> AffineGeometry | None:
    """Per-element factors of a 3D mesh whose elements are affine images of the reference
    tetrahedron (straight edges, edge nodes at their midpoints, as the P1 -> P2 elevation places
    them).  The Jacobian is then the same at every point, so the per-point gradients and weights of
    ``geometry`` follow from 10 numbers per element and the rule's reference gradients.  The
    Jacobian is taken at the rule's first point by ``geometry``'s formula (cofactor inverse,
    fem/assembly.py:139-245); None when some element's Jacobian varies over its points by more than
    ``rtol`` of its largest entry (curved elements keep the per-point form)."""
    coord = np.asarray(coord, dtype=np.float64)
    elem = np.asarray(elem, dtype=np.int64)
    if coord.shape[0] != 3:
        return None
    xi, wf = quadrature_3d(elem_type)
    _, h1, h2, h3 = basis_3d(elem_type, xi)          # (n_p, n_q) each
    cx, cy, cz = coord[0, elem].T, coord[1, elem].T, coord[2, elem].T   # (n_e, n_p)

    def jac(q):
        return [[c @ h[:, q] for c in (cx, cy, cz)] for h in (h1, h2, h3)]   # J[a][b] = dx_b / dxi_a

    J = jac(0)
    scale = np.max(np.abs(np.array(J)), axis=(0, 1))
    for q in range(1, xi.shape[1]):
        Jq = jac(q)
        dev = np.max(np.abs(np.array(Jq) - np.array(J)), axis=(0, 1))
        if np.any(dev > rtol * scale):
            return None
    (j11, j12, j13), (j21, j22, j23), (j31, j32, j33) = J
    det = j11 * (j22 * j33 - j23 * j32) - j12 * (j21 * j33 - j23 * j31) + j13 * (j21 * j32 - j22 * j31)
    inv = 1.0 / det
    gm = np.empty((elem.shape[1], 10))
    cof = [(j22 * j33 - j23 * j32), -(j12 * j33 - j13 * j32), (j12 * j23 - j13 * j22),
           -(j21 * j33 - j23 * j31), (j11 * j33 - j13 * j31), -(j11 * j23 - j13 * j21),
           (j21 * j32 - j22 * j31), -(j11 * j32 - j12 * j31), (j11 * j22 - j12 * j21)]
    for k, c in enumerate(cof):
        gm[:, k] = c * inv
    gm[:, 9] = np.abs(det)
    ref = np.concatenate([h1.T.reshape(-1), h2.T.reshape(-1), h3.T.reshape(-1), wf])
    return AffineGeometry(gm=np.ascontiguousarray(gm), ref=np.ascontiguousarray(ref, dtype=np.float64))


def volume_load(geom: Geometry, elem: np.ndarray, n_nodes: int, f_int: np.ndarray) -> np.ndarray:
    """Nodal load f[c, node] = sum_{e,q} hatp[i,q] w_g f_int[c,g] (fem/assembly.py:337-364)."""
    elem = np.asarray(elem, dtype=np.int64)
    n_q = geom.n_q
    out = np.zeros((geom.dim, n_nodes))
    for c in range(geom.dim):
        wf = (geom.weight * f_int[c]).reshape(geom.n_elem, n_q)  # (n_e, n_q)
        contrib = wf @ geom.hatp.T                                 # (n_e, n_p)
        np.add.at(out[c], elem.T.reshape(-1), contrib.reshape(-1))
    return out
