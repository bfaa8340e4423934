"""Slope meshes: the reference's procedural 2D P2 slope, a structured 3D slope of tetrahedra with
the same profile, tet4 -> tet10 elevation, xyz node reordering and a Gmsh 2.2 reader.

Setup-time host code.  Numbering rules that decide the CSR pattern are the reference's:

* 2D P2 slope: mesh/slope_2d.py:173-348 (``mesh_p2_2d``): structured P2 grid, foundation block of
  2*nx+1 columns, slope block of 2*nx12+1 columns stretched to the face, two triangles per cell,
  node order [v0, v1, v2, m12, m02, m01]; Dirichlet rule mesh/slope_2d.py:21-26.
* tet4 -> tet10: io.py:235-276: vertices keep their numbers, edge midpoints are appended in
  first-encounter order over elements, local edges (01, 12, 02, 13, 23, 03).
* xyz reordering: mesh/reorder.py:49-59 (lexsort with x primary, then y, then z).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np


@dataclass
class SlopeMesh:
    coord: np.ndarray          # (dim, n_nodes)
    elem: np.ndarray           # (n_p, n_elem)
    q_mask: np.ndarray         # (dim, n_nodes) True = free
    material: np.ndarray       # (n_elem,) material id
    elem_type: str
    surf: np.ndarray = field(default_factory=lambda: np.empty((0, 0), dtype=np.int64))

    @property
    def dim(self) -> int:
        return int(self.coord.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.coord.shape[1])

    @property
    def n_elem(self) -> int:
        return int(self.elem.shape[1])


# --------------------------------------------------------------------------------------------
# 2D procedural slope (mesh/slope_2d.py:173)
# --------------------------------------------------------------------------------------------

def slope_2d_p2(h: float = 1.0, x1: float = 15.0, x2: float = 10.0, x3: float = 15.0, y1: float = 10.0,
                y2: float = 10.0) -> SlopeMesh:
    nx12 = int(round((x1 + x2) / h))
    nx3 = int(round(x3 / h))
    nx = nx12 + nx3
    ny1 = int(round(y1 / h))
    ny2 = int(round(y2 / h))
    xs = np.concatenate((np.linspace(0.0, x1 + x2, 2 * nx12 + 1), np.linspace(x1 + x2, x1 + x2 + x3, 2 * nx3 + 1)[1:]))
    ys = np.concatenate((np.linspace(0.0, y1, 2 * ny1 + 1), np.linspace(y1, y1 + y2, 2 * ny2 + 1)[1:]))
    ncol_lo, ncol_hi = 2 * nx + 1, 2 * nx12 + 1
    nrow_lo, nrow_hi = 2 * ny1 + 1, 2 * ny2
    n_lo = nrow_lo * ncol_lo
    n_nodes = n_lo + nrow_hi * ncol_hi
    coord = np.zeros((2, n_nodes))
    # foundation block: row-major over (row j, column i)
    coord[0, :n_lo] = np.tile(xs, nrow_lo)
    coord[1, :n_lo] = np.repeat(ys[:nrow_lo], ncol_lo)
    # slope block: each row stretched to [0, x_max(y)]
    for r in range(nrow_hi):
        y = ys[nrow_lo + r]
        x_max = x1 + x2 * (y1 + y2 - y) / y2
        sl = slice(n_lo + r * ncol_hi, n_lo + (r + 1) * ncol_hi)
        coord[0, sl] = np.linspace(0.0, x_max, ncol_hi)
        coord[1, sl] = y

    def node(i, j):  # grid index -> node number (i column, j row), vectorised
        i = np.asarray(i)
        j = np.asarray(j)
        return np.where(j < nrow_lo, j * ncol_lo + i, n_lo + (j - nrow_lo) * ncol_hi + i)

    blocks = []
    for (j0, j1, ncx) in ((0, ny1, nx), (ny1, ny1 + ny2, nx12)):
        jj, ii = np.meshgrid(np.arange(j0, j1), np.arange(ncx), indexing="ij")
        ii, jj = ii.ravel(), jj.ravel()  # cells in (row-major over j, then i) order
        i2, j2 = 2 * ii, 2 * jj
        t1 = np.stack([node(i2, j2), node(i2 + 2, j2), node(i2, j2 + 2), node(i2 + 1, j2 + 1), node(i2, j2 + 1),
                       node(i2 + 1, j2)])
        t2 = np.stack([node(i2 + 2, j2 + 2), node(i2, j2 + 2), node(i2 + 2, j2), node(i2 + 1, j2 + 1),
                       node(i2 + 2, j2 + 1), node(i2 + 1, j2 + 2)])
        pair = np.empty((6, 2 * ii.size), dtype=np.int64)
        pair[:, 0::2] = t1
        pair[:, 1::2] = t2
        blocks.append(pair)
    elem = np.hstack(blocks)
    x_max = coord[0].max() - 1e-9
    q = np.zeros((2, n_nodes), dtype=bool)
    q[0] = (coord[0] > 0.0) & (coord[1] > 0.0) & (coord[0] < x_max)
    q[1] = coord[1] > 0.0
    return SlopeMesh(coord=coord, elem=elem, q_mask=q, material=np.zeros(elem.shape[1], dtype=np.int64),
                     elem_type="P2")


# --------------------------------------------------------------------------------------------
# Elevation and reordering
# --------------------------------------------------------------------------------------------

_TET_EDGES = ((0, 1), (1, 2), (0, 2), (1, 3), (2, 3), (0, 3))


def elevate_tet4_to_tet10(coord: np.ndarray, elem4: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Append edge midpoints in first-encounter order (elements ascending, local edges in
    (01, 12, 02, 13, 23, 03) order); returns (coord10, elem10)."""
    elem4 = np.asarray(elem4, dtype=np.int64)
    n_v = coord.shape[1]
    n_e = elem4.shape[1]
    a = np.stack([elem4[i] for i, _ in _TET_EDGES], axis=1)  # (n_e, 6)
    b = np.stack([elem4[j] for _, j in _TET_EDGES], axis=1)
    lo, hi = np.minimum(a, b).ravel(), np.maximum(a, b).ravel()  # element-major, local-edge minor
    key = lo * n_v + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")          # unique edges by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    edge_id = rank[inv]                               # edge number per (element, local edge)
    lo_u, hi_u = lo[first[order]], hi[first[order]]
    mid = 0.5 * (coord[:, lo_u] + coord[:, hi_u])
    coord10 = np.hstack((coord, mid))
    elem10 = np.vstack((elem4, (n_v + edge_id).reshape(n_e, 6).T))
    return coord10, elem10


def xyz_order(coord: np.ndarray) -> np.ndarray:
    keys = (coord[1], coord[0]) if coord.shape[0] == 2 else (coord[2], coord[1], coord[0])
    return np.lexsort(keys).astype(np.int64)


def reorder_nodes(mesh: SlopeMesh, perm: np.ndarray) -> SlopeMesh:
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    surf = inv[mesh.surf] if mesh.surf.size else mesh.surf
    return SlopeMesh(coord=mesh.coord[:, perm], elem=inv[mesh.elem], q_mask=mesh.q_mask[:, perm],
                     material=mesh.material, elem_type=mesh.elem_type, surf=surf)


# --------------------------------------------------------------------------------------------
# Structured 3D slope (own mesher; the reference's 3D meshes are Gmsh files)
# --------------------------------------------------------------------------------------------

_KUHN = ((0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7))


def slope_3d(h: float = 1.0, x1: float = 15.0, x2: float = 10.0, x3: float = 15.0, y1: float = 10.0,
             y2: float = 10.0, z_len: float = 40.0, nz: int | None = None, elem_type: str = "P2",
             layers: tuple[float, ...] | None = None, reorder: bool = True) -> SlopeMesh:
    """Slope of height y2 over a foundation of depth y1, crest bench x1, face run x2 (45 deg when
    x2 == y2), toe bench x3, extruded over z_len.  Hexahedral cells split into 6 Kuhn tetrahedra
    (a conforming split), elevated to P2 and reordered xyz.  Base fixed in all directions, x and z
    side planes on rollers.  ``layers``: increasing y levels splitting materials 0, 1, ... from the
    top down (heterogeneous case)."""
    nx12 = int(round((x1 + x2) / h))
    nx3 = int(round(x3 / h))
    nx = nx12 + nx3
    ny1 = int(round(y1 / h))
    ny2 = int(round(y2 / h))
    nz = int(nz if nz is not None else round(z_len / h))
    xs_lo = np.concatenate((np.linspace(0.0, x1 + x2, nx12 + 1), np.linspace(x1 + x2, x1 + x2 + x3, nx3 + 1)[1:]))
    ys = np.concatenate((np.linspace(0.0, y1, ny1 + 1), np.linspace(y1, y1 + y2, ny2 + 1)[1:]))
    zs = np.linspace(0.0, z_len, nz + 1)
    # 2D vertex grid (same layout as the P1 2D slope): rows j, columns i
    pts2 = []
    idx2 = {}
    for j in range(ny1 + ny2 + 1):
        if j <= ny1:
            xs = xs_lo
        else:
            x_max = x1 + x2 * (y1 + y2 - ys[j]) / y2
            xs = np.linspace(0.0, x_max, nx12 + 1)
        for i, x in enumerate(xs):
            idx2[(i, j)] = len(pts2)
            pts2.append((x, ys[j]))
    pts2 = np.asarray(pts2)
    n2 = pts2.shape[0]
    coord = np.zeros((3, n2 * (nz + 1)))
    coord[0] = np.tile(pts2[:, 0], nz + 1)
    coord[1] = np.tile(pts2[:, 1], nz + 1)
    coord[2] = np.repeat(zs, n2)
    quads = []
    for j in range(ny1 + ny2):
        ncx = nx if j < ny1 else nx12
        for i in range(ncx):
            quads.append((idx2[(i, j)], idx2[(i + 1, j)], idx2[(i, j + 1)], idx2[(i + 1, j + 1)]))
    quads = np.asarray(quads, dtype=np.int64)  # corners (i,j), (i+1,j), (i,j+1), (i+1,j+1)
    tets = []
    for k in range(nz):
        base = k * n2
        # hex corner c = bx + 2*by + 4*bz
        corner = np.stack([quads[:, 0] + base, quads[:, 1] + base, quads[:, 2] + base, quads[:, 3] + base,
                           quads[:, 0] + base + n2, quads[:, 1] + base + n2, quads[:, 2] + base + n2,
                           quads[:, 3] + base + n2], axis=1)
        for t in _KUHN:
            tets.append(corner[:, list(t)])
    elem4 = np.concatenate([t[:, None, :] for t in tets], axis=1).reshape(-1, 4).T  # hex-major, tet-minor
    # orient positively
    v = coord[:, elem4]
    det = np.einsum("ij,ij->j", np.cross((v[:, 1] - v[:, 0]).T, (v[:, 2] - v[:, 0]).T).T, v[:, 3] - v[:, 0])
    neg = det < 0
    elem4[[1, 2]] = np.where(neg, elem4[[2, 1]], elem4[[1, 2]])
    if elem_type.upper() == "P1":
        coord_e, elem = coord, elem4
    elif elem_type.upper() == "P2":
        coord_e, elem = elevate_tet4_to_tet10(coord, elem4)
    else:
        raise ValueError("slope_3d supports P1 and P2")
    q = np.ones((3, coord_e.shape[1]), dtype=bool)
    tol = 1e-9
    xmax, zmax = coord_e[0].max(), coord_e[2].max()
    q[0, (coord_e[0] < tol) | (coord_e[0] > xmax - tol)] = False
    q[2, (coord_e[2] < tol) | (coord_e[2] > zmax - tol)] = False
    q[:, coord_e[1] < tol] = False
    material = np.zeros(elem.shape[1], dtype=np.int64)
    if layers:
        yc = coord_e[1, elem[:4]].mean(axis=0)
        levels = sorted(layers, reverse=True)
        for m, lev in enumerate(levels):
            material[yc < lev] = m + 1
    mesh = SlopeMesh(coord=coord_e, elem=elem, q_mask=q, material=material, elem_type=elem_type.upper())
    if reorder:
        mesh = reorder_nodes(mesh, xyz_order(mesh.coord))
    return mesh


# --------------------------------------------------------------------------------------------
# Gmsh 2.2 ASCII reader (tetrahedra + boundary triangles with physical tags)
# --------------------------------------------------------------------------------------------

def read_gmsh22(path: str | Path):
    """Return (coord (3, n), tet4 (4, n_t), tet_phys (n_t,), tri3 (3, n_f), tri_phys (n_f,)) with
    0-based node numbers following the file's node order."""
    lines = Path(path).read_text().splitlines()
    it = iter(range(len(lines)))
    coord = None
    node_index = None
    tets, tet_tag, tris, tri_tag = [], [], [], []
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if s == "$Nodes":
            n = int(lines[i + 1])
            arr = np.array([lines[i + 2 + k].split() for k in range(n)], dtype=np.float64)
            ids = arr[:, 0].astype(np.int64)
            coord = arr[:, 1:4].T.copy()
            node_index = np.full(ids.max() + 1, -1, dtype=np.int64)
            node_index[ids] = np.arange(n)
            i += n + 2
            continue
        if s == "$Elements":
            n = int(lines[i + 1])
            for k in range(n):
                parts = lines[i + 2 + k].split()
                etype, ntags = int(parts[1]), int(parts[2])
                tags = parts[3:3 + ntags]
                nodes = [int(v) for v in parts[3 + ntags:]]
                phys = int(tags[0]) if ntags else 0
                if etype == 4:
                    tets.append(nodes)
                    tet_tag.append(phys)
                elif etype == 2:
                    tris.append(nodes)
                    tri_tag.append(phys)
            i += n + 2
            continue
        i += 1
    del it
    tet4 = node_index[np.asarray(tets, dtype=np.int64)].T
    tri3 = node_index[np.asarray(tris, dtype=np.int64)].T if tris else np.empty((3, 0), dtype=np.int64)
    return coord, tet4, np.asarray(tet_tag, dtype=np.int64), tri3, np.asarray(tri_tag, dtype=np.int64)
