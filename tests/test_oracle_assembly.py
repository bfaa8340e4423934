"""Pin the assembly oracle (strain, force, tangent rows) and the host setup it consumes (geometry,
fixed pattern, row slots) against brute force, the reference's golden values and its compiled C."""

from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from slope_stability_petsc4py_amd import fem, mesh, pattern

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def small3d():
    m = mesh.slope_3d(h=5.0, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=2)
    geom = fem.geometry(m.coord, m.elem, m.elem_type)
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    return m, geom, pat


def test_golden_pattern_and_values(small3d):
    m, geom, pat = small3d
    d = np.load(GOLD / "assembly_3d.npz")
    np.testing.assert_array_equal(m.elem, d["elem"])
    for k in ("indptr", "indices", "row_slot_ptr", "slot_elem", "slot_lrow", "slot_pos", "constrained_diag_positions"):
        np.testing.assert_array_equal(getattr(pat, k), d[k], err_msg=k)
    n_q = int(d["n_q"])
    E = oracle.strain_3d(d["dphi"], d["u"], pat.elem_dofs, n_q)
    np.testing.assert_allclose(E, d["E"], rtol=0, atol=1e-13 * np.abs(d["E"]).max())
    F = oracle.force_3d_rows(d["dphi"], d["S"], d["weight"], pat, n_q)
    np.testing.assert_allclose(F, d["F"], rtol=0, atol=1e-13 * np.abs(d["F"]).max())
    K = oracle.tangent_3d_rows(d["dphi"], d["DS"], d["weight"], pat, n_q)
    np.testing.assert_allclose(K, d["K"], rtol=0, atol=1e-13 * np.abs(d["K"]).max())


def test_pattern_matches_set_rule(small3d):
    m, _, pat = small3d
    ip, ix = oracle.brute_pattern(m.elem, m.q_mask)
    np.testing.assert_array_equal(pat.indptr, ip)
    np.testing.assert_array_equal(pat.indices, ix)


def test_pattern_2d_matches_set_rule():
    m = mesh.slope_2d_p2(h=2.0, x1=6.0, x2=4.0, x3=6.0, y1=4.0, y2=4.0)
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    ip, ix = oracle.brute_pattern(m.elem, m.q_mask)
    np.testing.assert_array_equal(pat.indptr, ip)
    np.testing.assert_array_equal(pat.indices, ix)


def test_slot_metadata_consistent(small3d):
    m, _, pat = small3d
    dim, nld = 3, 30
    free = m.q_mask.T.reshape(-1)
    for r in range(pat.n_rows):
        s0, s1 = pat.row_slot_ptr[r], pat.row_slot_ptr[r + 1]
        if not free[r]:
            assert s0 == s1
            assert pat.indptr[r + 1] - pat.indptr[r] == 1 and pat.indices[pat.indptr[r]] == r
            continue
        elems = pat.slot_elem[s0:s1]
        assert np.all(np.diff(elems) > 0)  # element order
        for s in range(s0, s1):
            e, a = pat.slot_elem[s], pat.slot_lrow[s]
            dofs = pat.elem_dofs[e]
            assert dofs[a] == r
            for b in range(nld):
                p = pat.slot_pos[s, b]
                if free[dofs[b]]:
                    assert pat.indptr[r] <= p < pat.indptr[r + 1] and pat.indices[p] == dofs[b]
                else:
                    assert p == -1


def test_strain_and_force_vs_bruteforce_B(small3d):
    m, geom, pat = small3d
    B = oracle.strain_operator(geom.dphi, m.elem, m.n_nodes, geom.n_q)
    rng = np.random.default_rng(3)
    u = rng.normal(size=3 * m.n_nodes)
    E = oracle.strain_3d(geom.dphi, u, pat.elem_dofs, geom.n_q)
    np.testing.assert_allclose(E.reshape(-1), B @ u, rtol=1e-12, atol=1e-12)
    S = rng.normal(size=(geom.n_int, 6))
    F = oracle.force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q)
    F_b = B.T @ (geom.weight[:, None] * S).reshape(-1)
    F_b[pat.row_free == 0] = 0.0
    np.testing.assert_allclose(F, F_b, rtol=1e-12, atol=1e-12 * np.abs(F_b).max())


def test_tangent_rows_vs_global_BtDB(small3d):
    """Rows kernel == restriction of B^T D B to the pattern (the reference's own gate, 1e-11)."""
    m, geom, pat = small3d
    B = oracle.strain_operator(geom.dphi, m.elem, m.n_nodes, geom.n_q)
    rng = np.random.default_rng(4)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    K = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    blocks = [sp.csr_matrix(geom.weight[g] * DS[g].reshape(6, 6, order="F")) for g in range(geom.n_int)]
    Kg = (B.T @ sp.block_diag(blocks, format="csr") @ B).tocsr()
    free = m.q_mask.T.reshape(-1)
    rows = np.repeat(np.arange(pat.n_rows), np.diff(pat.indptr))
    expect = np.asarray(Kg[rows, pat.indices]).ravel()
    expect[pat.constrained_diag_positions] = 1.0
    assert np.all(~free[rows[pat.constrained_diag_positions]])
    np.testing.assert_allclose(K, expect, rtol=1e-11, atol=1e-11 * np.abs(expect).max())


def test_element_matrices_vs_rows(small3d):
    """Legacy element+scatter rule gives the same values (tests/test_distributed_tangent_rows.py:94-116)."""
    m, geom, pat = small3d
    rng = np.random.default_rng(5)
    DS = rng.normal(size=(geom.n_int, 36))
    K = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    ke = oracle.element_matrices_3d(geom.dphi, DS, geom.weight, m.n_elem, geom.n_q)
    out = np.zeros_like(K)
    for s in range(pat.slot_elem.size):
        e, a = pat.slot_elem[s], pat.slot_lrow[s]
        pos = pat.slot_pos[s]
        ok = pos >= 0
        np.add.at(out, pos[ok], ke[e, a, ok])
    out[pat.constrained_diag_positions] = 1.0
    np.testing.assert_allclose(K, out, rtol=1e-11, atol=1e-11 * np.abs(out).max())


@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built here")
def test_oracle_matches_compiled_reference_assembly(small3d):
    m, geom, pat = small3d
    rng = np.random.default_rng(6)
    u = rng.normal(size=3 * m.n_nodes)
    S = rng.normal(size=(geom.n_int, 6))
    DS = rng.normal(size=(geom.n_int, 36))
    assert np.array_equal(oracle.strain_3d(geom.dphi, u, pat.elem_dofs, geom.n_q),
                          oracle.ref_strain_3d(geom.dphi, u, pat.elem_dofs, geom.n_q))
    assert np.array_equal(oracle.force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q),
                          oracle.ref_force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q))
    np.testing.assert_allclose(oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q),
                               oracle.ref_tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q), rtol=1e-14,
                               atol=1e-14)


def test_int32_overflow_guard():
    with pytest.raises(OverflowError):
        pattern._check_int32("nnz", 2**31)


@pytest.mark.parametrize("n", [1, 7, 8, 9, 1003])
def test_gather_order_is_a_permutation_dealt_over_xcds(n):
    """The block gather's launch order visits every node once; workgroup k of the order sits in
    run k mod 8 of the Morton-sorted nodes, and each run is contiguous along the curve."""
    X = np.random.default_rng(n).random((3, n))
    order = pattern.gather_order(X)
    assert np.array_equal(np.sort(order), np.arange(n))
    # runs per XCD: consecutive workgroups of one XCD take consecutive Morton positions
    q = np.minimum((X - X.min(axis=1, keepdims=True)) / np.maximum(np.ptp(X, axis=1, keepdims=True), 1e-300) * 1024,
                   1023).astype(np.int64)
    code = np.zeros(n, dtype=np.int64)
    for bit in range(10):
        for d in range(3):
            code |= ((q[d] >> bit) & 1) << (3 * bit + d)
    rank = np.empty(n, dtype=np.int64)
    rank[np.argsort(code, kind="stable")] = np.arange(n)
    for x in range(8):
        r = rank[order[x::8]]
        assert np.all(np.diff(r) == 1)
