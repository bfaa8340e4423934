"""Geometry and mesh setup: quadrature exactness, nodal bases, volumes, the reference's 2D slope
counts, tet10 elevation order, xyz ordering."""

import math
from itertools import product

import numpy as np
import pytest

from slope_stability_petsc4py_amd import fem, mesh


def _mono_tet(a, b, c):
    # integral of x^a y^b z^c over the unit tetrahedron
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


@pytest.mark.parametrize("et,deg", [("P1", 1), ("P2", 3), ("P4", 6)])
def test_tet_quadrature_exact(et, deg):
    xi, wf = fem.quadrature_3d(et)
    # the P2 rule's published decimals carry ~1e-13 truncation
    tol = 1e-12 if et != "P2" else 5e-12
    for a, b, c in product(range(deg + 1), repeat=3):
        if a + b + c <= deg:
            val = np.sum(wf * xi[0] ** a * xi[1] ** b * xi[2] ** c)
            assert abs(val - _mono_tet(a, b, c)) < tol, (a, b, c)


def test_tri_quadrature_exact():
    xi, wf = fem.quadrature_2d("P2")
    for a, b in product(range(6), repeat=2):
        if a + b <= 5:
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(wf * xi[0] ** a * xi[1] ** b) - exact) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 4])
def test_tet_basis_nodal_and_partition(p):
    nodes = np.array([[t[1], t[2], t[3]] for t in fem.tetra_node_tuples(p)], dtype=float).T / p
    if p == 4:
        hatp, d1, d2, d3 = fem.tetra_lagrange_basis(4, nodes)
    else:
        hatp, d1, d2, d3 = fem.basis_3d(f"P{p}", nodes)
    np.testing.assert_allclose(hatp, np.eye(nodes.shape[1]), atol=1e-12)
    xi, _ = fem.quadrature_3d("P4")
    h, g1, g2, g3 = fem.tetra_lagrange_basis(p, xi)
    np.testing.assert_allclose(h.sum(0), 1.0, atol=1e-12)
    for g in (g1, g2, g3):
        np.testing.assert_allclose(g.sum(0), 0.0, atol=1e-11)


def test_p2_closed_form_matches_general_lagrange():
    xi, _ = fem.quadrature_3d("P2")
    a = fem.basis_3d("P2", xi)
    b = fem.tetra_lagrange_basis(2, xi)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, atol=1e-13)


def test_reference_2d_slope_counts():
    # benchmarks/run_2D_homo_SSR_capture run_info.json: 2721 nodes, 1300 elements, 5220 unknowns
    m = mesh.slope_2d_p2(h=1.0, x1=15.0, x2=10.0, x3=15.0, y1=10.0, y2=10.0)
    assert m.coord.shape == (2, 2721)
    assert m.elem.shape == (6, 1300)
    assert int(m.q_mask.sum()) == 5220


def test_volumes_and_weights():
    m = mesh.slope_2d_p2(h=1.0)
    g = fem.geometry(m.coord, m.elem, "P2")
    area = 40 * 10 + 0.5 * (15 + 25) * 10  # foundation + trapezoid
    assert abs(g.weight.sum() - area) < 1e-9
    m3 = mesh.slope_3d(h=2.5, nz=4)
    g3 = fem.geometry(m3.coord, m3.elem, "P2")
    assert abs(g3.weight.sum() - area * 40.0) < 1e-8
    # gradients of a linear field are exact
    u = 2.0 * m3.coord[0] - 3.0 * m3.coord[1] + 0.5 * m3.coord[2]
    grad = np.stack([(g3.dphi[c] * u[m3.elem.T].repeat(g3.n_q, axis=0)).sum(1) for c in range(3)])
    np.testing.assert_allclose(grad, np.array([[2.0], [-3.0], [0.5]]) * np.ones((1, g3.n_int)), atol=1e-10)


def test_affine_geometry_factors():
    """Per-element factors of straight-sided P2 tetrahedra: for x = A xi + x0 the factors are
    A^-T and |det A| (textbook inverse Jacobian), they rebuild geometry()'s per-point gradients and
    weights, and a curved element (one edge node moved off its midpoint) has none."""
    rng = np.random.default_rng(3)
    A = rng.normal(size=(3, 3)) + 2 * np.eye(3)
    x0 = rng.normal(size=3)
    ref_v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]).T        # reference vertices
    coord4 = A @ ref_v + x0[:, None]
    coord, elem = mesh.elevate_tet4_to_tet10(coord4, np.arange(4)[:, None])
    aff = fem.affine_geometry(coord, elem, "P2")
    assert aff is not None
    np.testing.assert_allclose(aff.gm[0, :9].reshape(3, 3), np.linalg.inv(A).T, rtol=0, atol=1e-13)
    assert abs(aff.gm[0, 9] - abs(np.linalg.det(A))) < 1e-13
    # a transposed inverse would not rebuild the gradients: A is not symmetric
    assert np.abs(np.linalg.inv(A) - np.linalg.inv(A).T).max() > 1e-3

    m = mesh.slope_3d(h=2.5, nz=4)
    g = fem.geometry(m.coord, m.elem, "P2")
    aff = fem.affine_geometry(m.coord, m.elem, "P2")
    n_q, n_p = g.n_q, g.n_p
    H = aff.ref[:3 * n_q * n_p].reshape(3, n_q, n_p)
    wf = aff.ref[3 * n_q * n_p:]
    M = aff.gm[:, :9].reshape(-1, 3, 3)
    dphi = np.einsum("edk,kqn->deqn", M, H).reshape(3, -1, n_p)
    w = (aff.gm[:, 9:10] * wf[None]).reshape(-1)
    assert np.abs(dphi - g.dphi).max() < 1e-13 * np.abs(g.dphi).max()
    assert np.abs(w - g.weight).max() < 1e-13 * g.weight.max()
    bent = m.coord.copy()
    bent[:, m.elem[4, 7]] += 0.05 * np.array([1.0, 0.5, -0.25])
    assert fem.affine_geometry(bent, m.elem, "P2") is None
    assert fem.affine_geometry(m.coord, m.elem, "P2", rtol=1e-10) is not None


def test_tet10_elevation_first_encounter_order():
    coord = np.array([[0, 1, 0, 0, 1], [0, 0, 1, 0, 1], [0, 0, 0, 1, 1]], dtype=float)
    elem = np.array([[0, 1], [1, 2], [2, 3], [3, 4]])
    c10, e10 = mesh.elevate_tet4_to_tet10(coord, elem)
    # element 0 edges (01,12,02,13,23,03) -> new nodes 5..10
    np.testing.assert_array_equal(e10[4:, 0], [5, 6, 7, 8, 9, 10])
    # element 1 = (1,2,3,4): edges 12 (seen: 6), 23 (seen: 9), 13 (seen: 8), 24 new, 34 new, 14 new
    np.testing.assert_array_equal(e10[4:, 1], [6, 9, 8, 11, 12, 13])
    np.testing.assert_allclose(c10[:, 11], 0.5 * (coord[:, 2] + coord[:, 4]))


@pytest.mark.parametrize("et,nodes,constrained,free", [("P2", 27605, (395, 722, 1336), 80362),
                                                       ("P4", 208549, (1472, 2783, 5070), 616322)])
def test_hetero_L1_known_answers(et, nodes, constrained, free):
    """Known answers of the reference's loader tests (tests/test_3d_hetero_ssr_gmsh_loader.py:40-93)
    and run sizes (80,362 P2 unknowns of the 3D hetero SSR capture; 616,322 P4 unknowns)."""
    from slope_stability_petsc4py_amd.cases import DATA
    d = np.load(DATA / "ssr_hetero_L1_tet4.npz")
    assert d["coord"].shape == (3, 3845) and d["tet4"].shape == (4, 18419) and d["tri3"].shape == (3, 6325)
    m = mesh.tet_mesh_from_arrays(d["coord"], d["tet4"].astype(np.int64), d["tet_phys"].astype(np.int64) - 1,
                                  d["tri3"].astype(np.int64), d["tri_phys"].astype(np.int64) - 1,
                                  {"x": [1, 2], "y": [5], "z": [3, 4]}, et)
    assert m.coord.shape[1] == nodes
    assert tuple(int(v) for v in (~m.q_mask).sum(axis=1)) == constrained
    assert int(m.q_mask.sum()) == free
    if et == "P4":
        # every local node sits at its barycentric lattice position
        tup = fem.tetra_node_tuples(4)
        e = 77
        v = m.coord[:, m.elem[:4, e]]
        for k, t in enumerate(tup):
            np.testing.assert_allclose(m.coord[:, m.elem[k, e]], sum(t[i] / 4.0 * v[:, i] for i in range(4)),
                                       atol=1e-9)


def test_gmsh_reader_binary_and_ascii_agree(tmp_path):
    coord = np.array([[0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1.0]])
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", "4"]
    lines += [f"{i + 1} {coord[0, i]} {coord[1, i]} {coord[2, i]}" for i in range(4)]
    lines += ["$EndNodes", "$Elements", "2", "1 2 2 3 1 1 2 3", "2 4 2 2 1 1 2 3 4", "$EndElements"]
    p = tmp_path / "a.msh"
    p.write_text("\n".join(lines) + "\n")
    c, t4, tp, tri, trp = mesh.read_gmsh22(p)
    np.testing.assert_allclose(c, coord)
    np.testing.assert_array_equal(t4[:, 0], [0, 1, 2, 3])
    assert tp.tolist() == [2] and trp.tolist() == [3] and tri[:, 0].tolist() == [0, 1, 2]


def test_xyz_order_is_x_major():
    m = mesh.slope_3d(h=5.0, nz=2, x1=10, x2=10, x3=10, y1=10, y2=10, z_len=10)
    c = m.coord
    keys = np.lexsort((c[2], c[1], c[0]))
    np.testing.assert_array_equal(keys, np.arange(c.shape[1]))


def test_node_slots_share_offsets_across_free_rows():
    """One int16 offset table per (node, element) serves every free row of the node: it equals
    slot_pos - row start for each of them (the row length for a constrained column); slots are in
    element order."""
    from slope_stability_petsc4py_amd import pattern
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=3)
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    ns = pattern.node_slots(pat)
    dim = pat.dim
    ip = pat.indptr.astype(np.int64)
    rsp = pat.row_slot_ptr.astype(np.int64)
    checked = 0
    for node in range(pat.n_nodes):
        a, b = ns.ptr[node], ns.ptr[node + 1]
        assert np.all(np.diff(ns.elem[a:b]) > 0)
        for c in range(dim):
            row = dim * node + c
            if not pat.row_free[row]:
                continue
            s0, s1 = rsp[row], rsp[row + 1]
            assert s1 - s0 == b - a
            np.testing.assert_array_equal(pat.slot_elem[s0:s1], ns.elem[a:b])
            np.testing.assert_array_equal(pat.slot_lrow[s0:s1] // dim, ns.lnode[a:b])
            pos = pat.slot_pos[s0:s1].astype(np.int64)
            exp = np.where(pos >= 0, pos - ip[row], ip[row + 1] - ip[row])
            np.testing.assert_array_equal(ns.off[a:b], exp)
            checked += 1
    assert checked > 0 and ns.max_len == int(np.diff(ip)[pat.row_free.astype(bool)].max())


def test_gmsh_fixture_meshes_match_reference_unknowns():
    """Free dof counts of the committed Gmsh meshes equal the reference runs' `unknowns`
    (run_info.json of benchmarks/run_3D_hetero_SSR_capture: 80,362; slope_stability_3D_homo_LL:
    73,459; slope_stability_3D_hetero_LL: 105,910 -- the last needs the physical-name mapping of
    its boundary groups, whose tags are not their logical ids)."""
    from slope_stability_petsc4py_amd import cases
    assert int(cases.gmsh_fixture_mesh("ssr_hetero_L1_tet4.npz").q_mask.sum()) == 80362
    assert int(cases.gmsh_fixture_mesh("ll_homo_L1_tet4.npz").q_mask.sum()) == 73459
    assert int(cases.gmsh_fixture_mesh("ll_hetero_L1_tet4.npz").q_mask.sum()) == 105910


def test_logical_ids_follow_physical_names():
    names = {(2, 1): "boundary_0", (2, 2): "boundary_1", (2, 3): "boundary_3", (2, 4): "boundary_5",
             (3, 1): "material_0"}
    np.testing.assert_array_equal(mesh.logical_ids(np.array([1, 3, 4, 2]), names, 2, "boundary"), [0, 3, 5, 1])
    np.testing.assert_array_equal(mesh.logical_ids(np.array([2, 7]), {}, 2, "boundary"), [2, 7])
    with pytest.raises(ValueError):
        mesh.logical_ids(np.array([9]), names, 2, "boundary")


def test_strip_footing_load_totals_and_p2_weights():
    """Uniform pressure on a top strip: total force = q x loaded width (x depth in 3D); on P2 edges
    the end/midside weights are L/6 and 2L/3, on P2 faces the corners carry nothing (exact integrals
    of the quadratic shape functions)."""
    from slope_stability_petsc4py_amd import loads
    m = mesh.slope_2d_p2(h=0.5, x1=12.0, x2=1.0, x3=3.0, y1=8.0, y2=1.0)
    F, w = loads.strip_footing_load(m, (0.0, 2.0), q=10.0)
    assert np.isclose(F.sum(), -10.0 * w) and w > 1.5 and F[0].sum() == 0.0
    yt = m.coord[1].max()
    top = np.flatnonzero((np.abs(m.coord[1] - yt) < 1e-9) & (m.coord[0] <= w + 1e-9))
    top = top[np.argsort(m.coord[0, top])]
    fy = F[1, top]
    L = m.coord[0, top[2]] - m.coord[0, top[0]]
    assert np.isclose(fy[0], -10.0 * L / 6.0) and np.isclose(fy[1], -10.0 * 2.0 * L / 3.0)
    assert np.isclose(fy[2], -10.0 * L / 3.0)           # shared vertex: two edges
    m3 = mesh.slope_3d(h=1.0, x1=12.0, x2=1.0, x3=3.0, y1=8.0, y2=1.0, z_len=3.0, nz=3)
    F3, w3 = loads.strip_footing_load(m3, (0.0, 2.0), q=10.0)
    assert np.isclose(F3.sum(), -10.0 * w3 * 3.0)
    corners = np.unique(m3.elem[:4].ravel())
    assert np.all(F3[:, corners] == 0.0)
    m1 = mesh.slope_3d(h=1.0, x1=12.0, x2=1.0, x3=3.0, y1=8.0, y2=1.0, z_len=3.0, nz=3, elem_type="P1")
    F1, w1 = loads.strip_footing_load(m1, (0.0, 2.0), q=10.0)
    assert np.isclose(F1.sum(), -10.0 * w1 * 3.0)


def test_prandtl_bearing_capacity_factor():
    """N_c closed form against its textbook values: 2 + pi at phi = 0, 30.14 at 30 deg, 14.83 at 20."""
    from slope_stability_petsc4py_amd.cases import prandtl_nc
    assert np.isclose(prandtl_nc(0.0), 2.0 + np.pi)
    assert abs(prandtl_nc(30.0) - 30.14) < 0.01
    assert abs(prandtl_nc(20.0) - 14.83) < 0.01
    assert abs(prandtl_nc(1e-6) - (2.0 + np.pi)) < 1e-4
