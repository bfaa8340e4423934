"""Host construction of the aggregation level below the sparse P1 level (no GPU)."""

import numpy as np
import pytest

from slope_stability_petsc4py_amd import mesh
from slope_stability_petsc4py_amd.linear import rbm_aggregates


def _vertex_mesh(h=2.0):
    m = mesh.slope_3d(h=h, nz=4, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0)
    verts = np.unique(m.elem[:4].reshape(-1))
    cnum = np.full(m.coord.shape[1], -1, dtype=np.int64)
    cnum[verts] = np.arange(verts.size)
    return m.coord[:, verts], cnum[m.elem[:4]], m.q_mask[:, verts].T.reshape(-1)


def _dense(tp, ti, tw, n_rows, n_t):
    T = np.zeros((n_rows, n_t))
    for r in range(n_rows):
        T[r, ti[tp[r]:tp[r + 1]]] = tw[tp[r]:tp[r + 1]]
    return T


@pytest.mark.parametrize("radius", [1, 2])
def test_rbm_aggregates_orthonormal_and_reproduce_rigid_motions(radius):
    coord, elem, free = _vertex_mesh()
    tp, ti, tw, n_t = rbm_aggregates(coord, elem, free, radius=radius)
    n = coord.shape[1]
    T = _dense(tp, ti, tw, 3 * n, n_t)
    # constrained dofs carry no mode
    assert not T[~free].any()
    # orthonormal columns (one aggregate per column, SVD basis within it)
    assert np.abs(T.T @ T - np.eye(n_t)).max() < 1e-12
    # global rigid motions restricted to the free dofs lie in the range of T
    x = coord - coord.mean(axis=1, keepdims=True)
    rng = np.random.default_rng(0)
    for _ in range(3):
        t = rng.standard_normal(3)
        w = rng.standard_normal(3)
        u = (t[:, None] + np.cross(w, x.T).T).T.reshape(-1)
        u[~free] = 0.0
        assert np.linalg.norm(T @ (T.T @ u) - u) <= 1e-10 * np.linalg.norm(u)
    # a non-rigid field (quadratic) is not reproduced: the space is the piecewise rigid one
    q = np.zeros((n, 3))
    q[:, 0] = x[1] ** 2
    q = q.reshape(-1)
    q[~free] = 0.0
    assert np.linalg.norm(T @ (T.T @ q) - q) > 1e-3 * np.linalg.norm(q)


def test_rbm_aggregates_partition_every_node():
    coord, elem, free = _vertex_mesh()
    tp, ti, tw, n_t = rbm_aggregates(coord, elem, free, radius=2)
    n = coord.shape[1]
    # the columns of each node's free dofs are one aggregate's: all dofs of a node share a column set
    cols_by_node = {}
    for node in range(n):
        s = set()
        for c in range(3):
            r = 3 * node + c
            s |= set(ti[tp[r]:tp[r + 1]].tolist())
        if s:
            cols_by_node[node] = frozenset(s)
    # column sets of different aggregates are disjoint
    groups = set(cols_by_node.values())
    seen = set()
    for g in groups:
        assert not (g & seen)
        seen |= g
    assert seen == set(range(n_t))
    # radius 2 gives far fewer columns than radius 1
    assert n_t < rbm_aggregates(coord, elem, free, radius=1)[3]
