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


def elevate_tet4_to_tet35(coord: np.ndarray, elem4: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Order-4 tetrahedra (io.py:339-398): per edge 3 nodes at (1-k/4) v_lo + (k/4) v_hi (k = 1..3,
    reversed when the element walks the edge from its larger vertex), per face the 3 interior
    nodes (2,1,1), (1,2,1), (1,1,2) in the face's sorted-vertex frame, one centroid per element;
    local order: vertices, edges (01,12,02,13,23,03), faces (012,013,023,123), centroid.
    New nodes are numbered edges, faces, centroids (the caller reorders xyz)."""
    elem4 = np.asarray(elem4, dtype=np.int64)
    n_v = coord.shape[1]
    n_e = elem4.shape[1]
    # edges
    a = np.stack([elem4[i] for i, _ in _TET_EDGES], axis=1)
    b = np.stack([elem4[j] for _, j in _TET_EDGES], axis=1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = (lo * n_v + hi).ravel()
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    eid = rank[inv].reshape(n_e, 6)
    lo_u, hi_u = lo.ravel()[first[order]], hi.ravel()[first[order]]
    n_edges = lo_u.size
    pts = [np.empty((3, 0))]
    for k in (1, 2, 3):
        pts.append(((4 - k) / 4.0) * coord[:, lo_u] + (k / 4.0) * coord[:, hi_u])
    edge_nodes = n_v + np.stack([np.arange(n_edges) * 3 + t for t in range(3)], axis=1)  # (n_edges, 3) lo->hi
    edge_pts = np.stack(pts[1:], axis=2).reshape(3, -1)  # node (edge, k) -> column edge*3 + k-1
    fwd = (a == lo)  # element walks lo -> hi
    loc_edges = []
    for t in range(6):
        en = edge_nodes[eid[:, t]]
        loc_edges.append(np.where(fwd[:, t:t + 1], en, en[:, ::-1]))
    # faces
    faces = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))
    fv = np.stack([np.stack([elem4[i] for i in f], axis=1) for f in faces], axis=1)  # (n_e, 4, 3) local verts
    can = np.sort(fv, axis=2)
    fkey = (can[..., 0] * n_v + can[..., 1]) * n_v + can[..., 2]
    fu, ffirst, finv = np.unique(fkey.ravel(), return_index=True, return_inverse=True)
    forder = np.argsort(ffirst, kind="stable")
    frank = np.empty_like(forder)
    frank[forder] = np.arange(forder.size)
    fid = frank[finv].reshape(n_e, 4)
    can_u = can.reshape(-1, 3)[ffirst[forder]]  # (n_faces, 3)
    n_faces = can_u.shape[0]
    tri_t = ((2, 1, 1), (1, 2, 1), (1, 1, 2))
    base_f = n_v + 3 * n_edges
    face_pts = []
    for t in tri_t:
        p = np.zeros((3, n_faces))
        for cnt, col in zip(t, range(3)):
            p = p + (cnt / 4.0) * coord[:, can_u[:, col]]
        face_pts.append(p)
    face_pts = np.stack(face_pts, axis=2).reshape(3, -1)  # (face, t) -> column face*3 + t
    loc_faces = []
    for fl in range(4):
        verts = fv[:, fl, :]                              # local order
        cf = can[:, fl, :]
        pos_in_can = (verts[:, :, None] == cf[:, None, :]).argmax(axis=2)  # local idx -> canonical idx
        nodes = []
        for t in tri_t:
            cc = np.zeros((n_e, 3), dtype=np.int64)
            for li in range(3):
                cc[np.arange(n_e), pos_in_can[:, li]] = t[li]
            tidx = np.where((cc == np.array([2, 1, 1])).all(1), 0, np.where((cc == np.array([1, 2, 1])).all(1), 1, 2))
            nodes.append(base_f + 3 * fid[:, fl] + tidx)
        loc_faces.append(np.stack(nodes, axis=1))
    cen = 0.25 * (coord[:, elem4[0]] + coord[:, elem4[1]] + coord[:, elem4[2]] + coord[:, elem4[3]])
    base_c = base_f + 3 * n_faces
    elem35 = np.vstack([elem4] + [e.T for e in loc_edges] + [f.T for f in loc_faces]
                       + [(base_c + np.arange(n_e))[None]])
    coord35 = np.hstack((coord, edge_pts, face_pts, cen))
    return coord35, elem35


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
    elif elem_type.upper() == "P4":
        coord_e, elem = elevate_tet4_to_tet35(coord, elem4)
    else:
        raise ValueError("slope_3d supports P1, P2 and P4")
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
    0-based node numbers following the file's node order.  Reads ASCII and binary 2.2 files."""
    data = Path(path).read_bytes()
    pos = data.index(b"$MeshFormat") + len(b"$MeshFormat\n")
    line_end = data.index(b"\n", pos)
    version, ftype, _dsize = data[pos:line_end].split()
    binary = int(ftype) == 1
    pos = line_end + 1
    if binary:
        pos += 4  # endianness marker (int 1)
    nodes_at = data.index(b"$Nodes", pos) + len(b"$Nodes\n")
    line_end = data.index(b"\n", nodes_at)
    n_nodes = int(data[nodes_at:line_end])
    pos = line_end + 1
    if binary:
        rec = np.dtype([("id", "<i4"), ("xyz", "<f8", 3)])
        arr = np.frombuffer(data, dtype=rec, count=n_nodes, offset=pos)
        ids, xyz = arr["id"].astype(np.int64), arr["xyz"]
    else:
        txt = data[pos:data.index(b"$EndNodes", pos)].split()
        a = np.array(txt, dtype=np.float64).reshape(n_nodes, 4)
        ids, xyz = a[:, 0].astype(np.int64), a[:, 1:4]
    node_index = np.full(ids.max() + 1, -1, dtype=np.int64)
    node_index[ids] = np.arange(n_nodes)
    el_at = data.index(b"$Elements", pos) + len(b"$Elements\n")
    line_end = data.index(b"\n", el_at)
    n_el = int(data[el_at:line_end])
    pos = line_end + 1
    nn = {1: 2, 2: 3, 4: 4, 15: 1}
    tets, tet_tag, tris, tri_tag = [], [], [], []
    if binary:
        done = 0
        while done < n_el:
            etype, nfollow, ntags = np.frombuffer(data, dtype="<i4", count=3, offset=pos)
            pos += 12
            width = 1 + int(ntags) + nn[int(etype)]
            blk = np.frombuffer(data, dtype="<i4", count=int(nfollow) * width, offset=pos).reshape(int(nfollow), width)
            pos += 4 * blk.size
            phys = blk[:, 1] if ntags else np.zeros(int(nfollow), dtype=np.int32)
            conn = blk[:, 1 + int(ntags):]
            if etype == 4:
                tets.append(conn)
                tet_tag.append(phys)
            elif etype == 2:
                tris.append(conn)
                tri_tag.append(phys)
            done += int(nfollow)
    else:
        lines = data[pos:data.index(b"$EndElements", pos)].decode().splitlines()
        for ln in lines[:n_el]:
            parts = [int(v) for v in ln.split()]
            etype, ntags = parts[1], parts[2]
            phys = parts[3] if ntags else 0
            conn = parts[3 + ntags:]
            if etype == 4:
                tets.append(np.array([conn]))
                tet_tag.append(np.array([phys]))
            elif etype == 2:
                tris.append(np.array([conn]))
                tri_tag.append(np.array([phys]))
    tet4 = node_index[np.concatenate(tets).astype(np.int64)].T
    tri3 = node_index[np.concatenate(tris).astype(np.int64)].T if tris else np.empty((3, 0), dtype=np.int64)
    return (xyz.T.copy(), tet4, np.concatenate(tet_tag).astype(np.int64),
            tri3, np.concatenate(tri_tag).astype(np.int64) if tri_tag else np.empty(0, dtype=np.int64))


def read_gmsh_physical_names(path: str | Path) -> dict[tuple[int, int], str]:
    """{(dim, physical tag): name} of a Gmsh 2.2 file's $PhysicalNames section ({} if absent)."""
    data = Path(path).read_bytes()
    at = data.find(b"$PhysicalNames")
    if at < 0:
        return {}
    end = data.index(b"$EndPhysicalNames", at)
    lines = data[at:end].decode(errors="replace").splitlines()[2:]
    out = {}
    for ln in lines:
        parts = ln.split(None, 2)
        if len(parts) == 3:
            out[(int(parts[0]), int(parts[1]))] = parts[2].strip().strip('"')
    return out


def logical_ids(tags: np.ndarray, names: dict[tuple[int, int], str], dim: int, prefix: str) -> np.ndarray:
    """Physical tags -> the logical ids the reference uses (io.py:165-189): a group named
    "<prefix>_<k>" of this dimension stands for id k; without such names the tags are the ids."""
    mapping = {}
    for (d, tag), name in names.items():
        low = name.strip().lower()
        if d == dim and low.startswith(prefix + "_") and low[len(prefix) + 1:].isdigit():
            mapping[tag] = int(low[len(prefix) + 1:])
    ids = np.asarray(tags, dtype=np.int64).ravel()
    if not mapping:
        return ids
    missing = sorted(int(v) for v in np.unique(ids) if int(v) not in mapping)
    if missing:
        raise ValueError(f"physical groups {missing} have no '{prefix}_<k>' name")
    lut = np.zeros(max(mapping) + 1, dtype=np.int64)
    for tag, k in mapping.items():
        lut[tag] = k
    return lut[ids]


_P4_FACE_LOCAL = {  # local P4 node indices of each tetrahedron face (vertices, edge nodes, face nodes)
    (0, 1, 2): [0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 22, 23, 24],
    (0, 1, 3): [0, 1, 3, 4, 5, 6, 13, 14, 15, 19, 20, 21, 25, 26, 27],
    (0, 2, 3): [0, 2, 3, 10, 11, 12, 16, 17, 18, 19, 20, 21, 28, 29, 30],
    (1, 2, 3): [1, 2, 3, 7, 8, 9, 13, 14, 15, 16, 17, 18, 31, 32, 33],
}


def _p4_face_nodes(tri3, tet4, elem35, n_v):
    """(15, n_faces) P4 nodes of the boundary triangles, found through an element that owns each face."""
    key_of = lambda a, b, c: (a * n_v + b) * n_v + c  # noqa: E731
    lut = {}
    for f, loc in _P4_FACE_LOCAL.items():
        v = np.sort(tet4[list(f)], axis=0)
        k = key_of(v[0], v[1], v[2])
        for e, kk in enumerate(k.tolist()):
            lut.setdefault(kk, (e, loc))
    t = np.sort(tri3, axis=0)
    out = np.empty((15, tri3.shape[1]), dtype=np.int64)
    for i, kk in enumerate(key_of(t[0], t[1], t[2]).tolist()):
        e, loc = lut[kk]
        out[:, i] = elem35[loc, e]
    return out


def tet_mesh_from_arrays(coord, tet4, material, tri3, boundary, dirichlet_labels: dict, elem_type: str = "P2",
                         reorder: bool = True) -> SlopeMesh:
    """Gmsh-family slope: elevate tet4 (io.py:235-276), build the Dirichlet mask from labelled
    boundary triangles (io.py:61-93: each axis constrained on its label set; triangle midpoints
    inherit the constraint through the elevated faces), reorder xyz (mesh/reorder.py:49-59)."""
    coord = np.asarray(coord, dtype=np.float64)
    if elem_type.upper() == "P2":
        coord_e, elem = elevate_tet4_to_tet10(coord, tet4)
        # face midpoint nodes of the boundary triangles: look the edges up through the element edges
        n_v = coord.shape[1]
        e4 = np.asarray(tet4, dtype=np.int64)
        a = np.stack([e4[i] for i, _ in _TET_EDGES], axis=1).ravel()
        b = np.stack([e4[j] for _, j in _TET_EDGES], axis=1).ravel()
        keys = np.minimum(a, b) * n_v + np.maximum(a, b)
        mids = elem[4:].T.ravel()
        order = np.argsort(keys)
        ks, ms = keys[order], mids[order]
        tri = np.asarray(tri3, dtype=np.int64)
        tri_edges = [(0, 1), (1, 2), (0, 2)]
        surf_nodes = [tri]
        for i, j in tri_edges:
            kk = np.minimum(tri[i], tri[j]) * n_v + np.maximum(tri[i], tri[j])
            surf_nodes.append(ms[np.searchsorted(ks, kk)][None])
        surf = np.vstack(surf_nodes)
    elif elem_type.upper() == "P1":
        coord_e, elem, surf = coord, np.asarray(tet4, dtype=np.int64), np.asarray(tri3, dtype=np.int64)
    elif elem_type.upper() == "P4":
        coord_e, elem = elevate_tet4_to_tet35(coord, tet4)
        # boundary faces: every P4 node of a labelled triangle (vertices, edge and face nodes) is
        # found through the elements' faces
        surf = _p4_face_nodes(np.asarray(tri3, dtype=np.int64), np.asarray(tet4, dtype=np.int64), elem,
                              coord.shape[1])
    else:
        raise ValueError("P1, P2 and P4 supported")
    q = np.ones((3, coord_e.shape[1]), dtype=bool)
    labels = np.asarray(boundary, dtype=np.int64)
    for axis, name in enumerate("xyz"):
        lab = dirichlet_labels.get(name, ())
        if lab:
            sel = np.isin(labels, np.asarray(lab))
            q[axis, surf[:, sel].ravel()] = False
    mesh = SlopeMesh(coord=coord_e, elem=elem, q_mask=q, material=np.asarray(material, dtype=np.int64),
                     elem_type=elem_type.upper(), surf=surf)
    if reorder:
        mesh = reorder_nodes(mesh, xyz_order(mesh.coord))
    return mesh
