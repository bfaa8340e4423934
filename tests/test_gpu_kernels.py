"""Parity of the HIP kernels (called through the C-ABI) with the CPU oracle and the reference's
golden vectors.  Tolerances: FP64 stress/tangent 1e-10 relative, assembly values
1e-12 relative to the largest entry, integer structure bit-exact."""

from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from slope_stability_petsc4py_amd import _lib, fem, mesh, pattern
from slope_stability_petsc4py_amd import workloads as W
from slope_stability_petsc4py_amd.constitutive import (ConstitutiveOperator, constitutive_problem_2d,
                                                       constitutive_problem_3d, reduction)
from slope_stability_petsc4py_amd.device import DevicePattern

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
RTOL = 1e-10


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _rowrel(a, b):
    """Max over points of |a-b|_inf / |b|_inf of that point."""
    num = np.abs(a - b).max(axis=1)
    den = np.maximum(np.abs(b).max(axis=1), 1e-300)
    return num / den


def test_constitutive_3d_golden(gpu):
    print("case: constitutive_3d golden", flush=True)
    d = np.load(GOLD / "constitutive_3d.npz")
    S, DS = constitutive_problem_3d(d["E"], d["c_bar"], d["sin_phi"], d["shear"], d["lame"], return_tangent=True)
    assert _rel(S.cpu().numpy(), d["S"]) < RTOL
    assert _rel(DS.cpu().numpy(), d["DS"]) < RTOL


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200_003])
def test_constitutive_3d_stratified_vs_oracle(gpu, n):
    print(f"case: constitutive_3d stratified n={n}", flush=True)
    E, b = W.stratified_strains(n, seed=n)
    mat = (np.full(n, 4 * np.sqrt(3)), np.full(n, 1 / np.sqrt(3)), np.full(n, W.HOMO.shear), np.full(n, W.HOMO.lame))
    S, DS, DS32, br = constitutive_problem_3d(E, *mat, return_tangent=True, fp32_tangent=True, return_branch=True)
    So, DSo, bro = oracle.constitutive_3d(E, *mat)
    assert np.array_equal(br.cpu().numpy().astype(np.int32), bro)
    assert _rel(S.cpu().numpy(), So) < RTOL
    assert _rel(DS.cpu().numpy(), DSo) < RTOL
    # per-point relative error, except near eigenvalue ties where the formula itself is ill-conditioned
    err = _rowrel(DS.cpu().numpy(), DSo)
    assert np.quantile(err, 0.999) < 1e-8
    np.testing.assert_array_equal(DS32.cpu().numpy(), DS.cpu().numpy().astype(np.float32))


def test_constitutive_3d_packed_symmetric_tangent(gpu):
    """The packed return mapping writes the full layout's upper entries (s <= m), with the same
    stresses and branch ids bit for bit; its tangent entries agree with the full kernel's to 1e-14
    of the point's largest entry (the packed kernel forms only the 21 upper entries, so a product
    that the full kernel shares between (i, j) and (j, i) may fuse into an FMA there: last-bit
    differences), and the unpacked tangent equals the full one (the tangent is symmetric) to the
    same bound.  Both are checked against the oracle to 1e-10 elsewhere."""
    from slope_stability_petsc4py_amd.device import ptr
    print("case: constitutive_3d packed symmetric DS", flush=True)
    n = 50_001
    E, _b = W.stratified_strains(n, seed=4)
    mat = [np.full(n, 4 * np.sqrt(3)), np.full(n, 1 / np.sqrt(3)), np.full(n, W.HOMO.shear), np.full(n, W.HOMO.lame)]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")  # noqa: E731
    Ed, md = t(E), [t(a) for a in mat]
    S1 = torch.empty((n, 6), dtype=torch.float64, device="cuda")
    S2 = torch.empty_like(S1)
    DS = torch.empty((n, 36), dtype=torch.float64, device="cuda")
    DSp = torch.empty((n, 21), dtype=torch.float64, device="cuda")
    b1 = torch.empty(n, dtype=torch.int8, device="cuda")
    b2 = torch.empty_like(b1)
    _lib.call("ssx_constitutive_3d", ptr(Ed), *(ptr(a) for a in md), n, ptr(S1), ptr(DS), None, ptr(b1), 0)
    _lib.call("ssx_constitutive_3d_sym", ptr(Ed), *(ptr(a) for a in md), n, ptr(S2), ptr(DSp), ptr(b2), 0)
    full = [s_ + 6 * m_ for s_ in range(6) for m_ in range(s_, 6)]
    assert torch.equal(S1, S2) and torch.equal(b1, b2)
    DSh = DS.cpu().numpy()
    errp = _rowrel(DSp.cpu().numpy(), DSh[:, full])
    print(f"  packed vs full upper entries: max per-point rel {errp.max():.2e}, "
          f"bitwise rows {(errp == 0).mean():.4f}", flush=True)
    assert errp.max() < 1e-14, errp.max()
    DSu = torch.full_like(DS, np.nan)
    _lib.call("ssx_ds_unpack_sym", n, ptr(DSp), ptr(DSu), 0)
    assert torch.equal(DSu[:, full], DSp)
    err = _rowrel(DSu.cpu().numpy(), DSh)
    assert err.max() < 1e-14, err.max()


def test_constitutive_3d_s_only_matches_sds(gpu):
    E, _ = W.stratified_strains(5000, seed=2)
    mat = (np.full(5000, 4 * np.sqrt(3)), np.full(5000, 1 / np.sqrt(3)), np.full(5000, W.HOMO.shear),
           np.full(5000, W.HOMO.lame))
    S1 = constitutive_problem_3d(E, *mat)
    S2, _ = constitutive_problem_3d(E, *mat, return_tangent=True)
    assert torch.equal(S1, S2)


def test_constitutive_2d_vs_oracle(gpu):
    print("case: constitutive_2d", flush=True)
    rng = np.random.default_rng(9)
    n = 3000
    E = rng.normal(0, 1e-3, (n, 3))
    E[:100] = 0.0                 # I2 == 0 tie branch
    E[100:200, 2] = 0.0
    E[100:200, 1] = E[100:200, 0]  # equal in-plane eigenvalues
    cb, sp = 4 * np.sqrt(3), 1 / np.sqrt(3)
    mat = (np.full(n, cb), np.full(n, sp), np.full(n, W.HOMO.shear), np.full(n, W.HOMO.bulk), np.full(n, W.HOMO.lame))
    S, DS, br = constitutive_problem_2d(E, *mat, return_tangent=True, return_branch=True)
    So, DSo, bro = oracle.mc2d(E, mat[0], mat[1], mat[2], mat[4])
    assert np.array_equal(br.cpu().numpy().astype(np.int32), bro)
    assert len(set(bro.tolist())) >= 4
    assert _rel(S.cpu().numpy(), So) < RTOL
    assert _rel(DS.cpu().numpy(), DSo) < RTOL


def test_reduction_vs_numpy(gpu):
    rng = np.random.default_rng(1)
    n = 1000
    c0 = rng.uniform(5, 20, n)
    phi = np.deg2rad(rng.uniform(20, 45, n))
    psi = np.deg2rad(rng.uniform(0, 10, n))
    for typ in "ABC":
        for lam in (0.7, 1.0, 1.66):
            cb, sp = reduction(c0, phi, psi, lam, typ)
            cbo, spo = oracle.davis_reduction(c0, phi, psi, lam, typ)
            assert _rel(cb.cpu().numpy(), cbo) < 1e-13
            assert _rel(sp.cpu().numpy(), spo) < 1e-13
    with pytest.raises(ValueError):
        reduction(c0, phi, psi, 0.0, "B")
    with pytest.raises(ValueError):
        reduction(c0, phi, psi, 1.0, "D")


@pytest.fixture(scope="module")
def asm3d():
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=4)
    geom = fem.geometry(m.coord, m.elem, m.elem_type)
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    return m, geom, pat


def test_assembly_3d_golden(gpu):
    print("case: assembly golden", flush=True)
    d = np.load(GOLD / "assembly_3d.npz")
    m = mesh.slope_3d(h=5.0, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=2)
    geom = fem.geometry(m.coord, m.elem, m.elem_type)
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    geom.dphi, geom.weight = d["dphi"], d["weight"]
    dp = DevicePattern(pat, geom)
    dev = dp.indptr.device
    E = torch.empty((geom.n_int, 6), dtype=torch.float64, device=dev)
    dp.strain(torch.from_numpy(d["u"]).to(dev), E)
    assert _rel(E.cpu().numpy(), d["E"]) < 1e-12
    fe = torch.empty(geom.n_elem * 30, dtype=torch.float64, device=dev)
    F = torch.empty(pat.n_rows, dtype=torch.float64, device=dev)
    dp.force(torch.from_numpy(d["S"]).to(dev), fe, F)
    assert _rel(F.cpu().numpy(), d["F"]) < 1e-12
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    dp.tangent(torch.from_numpy(d["DS"]).to(dev), K)
    assert _rel(K.cpu().numpy(), d["K"]) < 1e-12


def test_assembly_3d_vs_oracle_and_deterministic(gpu, asm3d):
    print("case: assembly 3D vs oracle", flush=True)
    m, geom, pat = asm3d
    dp = DevicePattern(pat, geom)
    dev = dp.indptr.device
    rng = np.random.default_rng(21)
    u = rng.normal(0, 1e-3, 3 * m.n_nodes)
    E = torch.empty((geom.n_int, 6), dtype=torch.float64, device=dev)
    dp.strain(torch.from_numpy(u).to(dev), E)
    Eo = oracle.strain_3d(geom.dphi, u, pat.elem_dofs, geom.n_q)
    assert _rel(E.cpu().numpy(), Eo) < 1e-12
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    S = rng.normal(size=(geom.n_int, 6))
    fe = torch.empty(geom.n_elem * 30, dtype=torch.float64, device=dev)
    F = torch.empty(pat.n_rows, dtype=torch.float64, device=dev)
    dp.force(torch.from_numpy(S).to(dev), fe, F)
    assert _rel(F.cpu().numpy(), oracle.force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q)) < 1e-12
    # the node-slot gather (default) and the row-slot gather sum in the same order: bit for bit
    assert dp.force_gather == "nodes"
    F_rows = torch.empty_like(F)
    dp.force_gather = "rows"
    dp.force(torch.from_numpy(S).to(dev), fe, F_rows)
    dp.force_gather = "nodes"
    assert torch.equal(F, F_rows)
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    K32 = torch.empty(pat.nnz, dtype=torch.float32, device=dev)
    DSd = torch.from_numpy(DS).to(dev)
    dp.tangent(DSd, K, K32)
    Ko = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    assert _rel(K.cpu().numpy(), Ko) < 1e-12
    np.testing.assert_array_equal(K32.cpu().numpy(), K.cpu().numpy().astype(np.float32))
    # constrained diagonal = 1
    assert np.all(K.cpu().numpy()[pat.constrained_diag_positions] == 1.0)
    # run twice: bitwise identical (no atomics, fixed summation order)
    K2 = torch.empty_like(K)
    dp.tangent(DSd, K2)
    assert torch.equal(K, K2)
    # regularised operator K_r = (1-r) K + r K_el
    el = torch.from_numpy(rng.normal(size=pat.nnz)).to(dev)
    el[torch.from_numpy(pat.constrained_diag_positions.astype(np.int64)).to(dev)] = 1.0
    Kr = torch.empty_like(K)
    dp.tangent(DSd, Kr, r=0.3, elastic=el)
    np.testing.assert_allclose(Kr.cpu().numpy(), 0.7 * Ko + 0.3 * el.cpu().numpy(), rtol=1e-12,
                               atol=1e-12 * np.abs(Ko).max())


def test_tangent_global_fallback_path(gpu, asm3d):
    """The LDS-tiled and the global-memory variants of the tangent kernel agree bitwise."""
    m, geom, pat = asm3d
    dp = DevicePattern(pat, geom)
    dp.tangent_method = "rows"
    dev = dp.indptr.device
    DS = torch.from_numpy(np.random.default_rng(2).normal(size=(geom.n_int, 36))).to(dev)
    K1 = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    K2 = torch.empty_like(K1)
    dp.tangent(DS, K1)
    Ko = oracle.tangent_3d_rows(geom.dphi, DS.cpu().numpy(), geom.weight, pat, geom.n_q)
    assert _rel(K1.cpu().numpy(), Ko) < 1e-12
    dp.max_block_nnz = -1  # forces the global-memory path
    dp.tangent(DS, K2)
    assert torch.equal(K1, K2)


def test_spmv_vs_oracle(gpu, asm3d):
    m, geom, pat = asm3d
    dp = DevicePattern(pat, geom)
    dev = dp.indptr.device
    rng = np.random.default_rng(5)
    val = rng.normal(size=pat.nnz)
    x = rng.normal(size=pat.n_rows)
    K = dp.new_csr()
    K.values.copy_(torch.from_numpy(val))
    K.values32.copy_(torch.from_numpy(val.astype(np.float32)))
    y = K.matvec(torch.from_numpy(x).to(dev))
    yo = oracle.spmv(pat.indptr, pat.indices, val, x)
    assert _rel(y.cpu().numpy(), yo) < 1e-13
    y32 = K.matvec32(torch.from_numpy(x.astype(np.float32)).to(dev))
    assert _rel(y32.cpu().numpy().astype(np.float64), yo) < 1e-5
    # both row forms: aligned quadruples with head/tail entries by lanes 0..2 (16-byte aligned
    # arrays, >= 4 lanes), and one entry per lane (fewer lanes, or arrays off the 16-byte grid)
    from gpu_helpers import misaligned
    Km = misaligned(K)
    for KK in (K, Km):
        for lanes in (1, 2, 4, 8, 16, 32, 64):
            KK.lanes = lanes
            assert _rel(KK.matvec(torch.from_numpy(x).to(dev)).cpu().numpy(), yo) < 1e-13
            y32 = KK.matvec32(torch.from_numpy(x.astype(np.float32)).to(dev))
            assert _rel(y32.cpu().numpy().astype(np.float64), yo) < 1e-5


@pytest.mark.parametrize("k", [1, 8, 13])
def test_spmm_vs_oracle(gpu, asm3d, k):
    """Multi-vector SpMV (k vectors as rows, padded row stride, passes of 8) == k oracle SpMVs."""
    from slope_stability_petsc4py_amd.device import ptr
    print(f"case: spmm k={k}", flush=True)
    m, geom, pat = asm3d
    dp = DevicePattern(pat, geom)
    dev = dp.indptr.device
    rng = np.random.default_rng(15 + k)
    val = rng.normal(size=pat.nnz)
    K = dp.new_csr()
    K.values.copy_(torch.from_numpy(val))
    n = pat.n_rows
    ld = n + 5
    X = rng.normal(size=(k, ld))
    Xd = torch.from_numpy(X).to(dev)
    Yd = torch.full((k, ld), 7.0, dtype=torch.float64, device=dev)
    for lanes in (4, 16):
        _lib.call("ssx_spmm_f64", n, ptr(K.indptr), ptr(K.indices), ptr(K.values), k, ptr(Xd), ld, ptr(Yd), ld,
                  lanes, 0)
        Y = Yd.cpu().numpy()
        for j in range(k):
            assert _rel(Y[j, :n], oracle.spmv(pat.indptr, pat.indices, val, X[j, :n])) < 1e-13
        assert np.all(Y[:, n:] == 7.0)


def test_vector_kernels(gpu):
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(8)
    n, k = 100_003, 7
    V = torch.from_numpy(rng.normal(size=(k, n))).to(dev)
    x = torch.from_numpy(rng.normal(size=n)).to(dev)
    out = torch.empty(k, dtype=torch.float64, device=dev)
    cols = torch.empty(64 * k, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    _lib.call("ssx_dot_cols", n, k, V.data_ptr(), n, x.data_ptr(), cols.data_ptr(), out.data_ptr(), st)
    np.testing.assert_allclose(out.cpu().numpy(), V.cpu().numpy() @ x.cpu().numpy(), rtol=1e-12)
    # deterministic: same bits on a second call; small n (one chunk) and one column
    out2 = torch.empty_like(out)
    _lib.call("ssx_dot_cols", n, k, V.data_ptr(), n, x.data_ptr(), cols.data_ptr(), out2.data_ptr(), st)
    assert torch.equal(out, out2)
    _lib.call("ssx_dot_cols", 5, 1, V.data_ptr(), n, x.data_ptr(), cols.data_ptr(), out2.data_ptr(), st)
    np.testing.assert_allclose(out2[0].item(), float(V[0, :5].cpu().numpy() @ x[:5].cpu().numpy()), rtol=1e-14)
    part = torch.empty(1024, dtype=torch.float64, device=dev)
    d = torch.empty(1, dtype=torch.float64, device=dev)
    _lib.call("ssx_dot", n, x.data_ptr(), x.data_ptr(), part.data_ptr(), d.data_ptr(), torch.cuda.current_stream().cuda_stream)
    np.testing.assert_allclose(d.item(), float(x.cpu().numpy() @ x.cpu().numpy()), rtol=1e-12)
    h = torch.from_numpy(rng.normal(size=k)).to(dev)
    y = x.clone()
    _lib.call("ssx_gemv_cols", n, k, V.data_ptr(), n, h.data_ptr(), -1.0, y.data_ptr(), torch.cuda.current_stream().cuda_stream)
    np.testing.assert_allclose(y.cpu().numpy(), x.cpu().numpy() - h.cpu().numpy() @ V.cpu().numpy(), rtol=1e-12,
                               atol=1e-12)


def test_operator_api_3d(gpu):
    """ConstitutiveOperator methods on a small 3D slope vs the oracle pipeline."""
    print("case: operator API 3D", flush=True)
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=4)
    op = ConstitutiveOperator.from_mesh(m, [W.HOMO], Davis_type="B")
    geom, pat = op.geom, op.pattern.host
    rng = np.random.default_rng(3)
    U = rng.normal(0, 2e-3, (3, m.n_nodes)) * m.q_mask
    op.reduction(1.2)
    F, K = op.build_F_K_tangent_reduced(U)
    c0 = np.full(geom.n_int, 6.0)
    cb, sp = oracle.davis_reduction(c0, np.full(geom.n_int, np.pi / 4), np.zeros(geom.n_int), 1.2, "B")
    E = oracle.strain_3d(geom.dphi, U.reshape(-1, order="F"), pat.elem_dofs, geom.n_q)
    S, DS, br = oracle.constitutive_3d(E, cb, sp, np.full(geom.n_int, W.HOMO.shear), np.full(geom.n_int, W.HOMO.lame))
    assert len(np.unique(br)) >= 3
    Fo = oracle.force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q)
    assert isinstance(F, np.ndarray) and F.shape == (3, m.n_nodes)
    assert _rel(F.reshape(-1, order="F"), Fo) < 1e-10
    Ko = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    assert _rel(K.values.cpu().numpy(), Ko) < 1e-10
    # P2 slope: straight-sided tetrahedra, the tangent kept packed; DS is unpacked on request
    assert op.pattern.affine and op._sym
    assert _rel(op.DS.t().cpu().numpy(), DS) < 1e-10 and op.DS.shape == (36, geom.n_int)
    Ff = op.build_F_free()
    np.testing.assert_array_equal(Ff, F.reshape(-1, order="F")[pat.free_idx])
    Kr = op.build_K_regularized(0.25)
    Kel = oracle.tangent_3d_rows(geom.dphi, np.tile(
        (np.diag([2 * W.HOMO.shear] * 3 + [W.HOMO.shear] * 3) + np.pad(np.full((3, 3), W.HOMO.lame), ((0, 3), (0, 3))))
        .reshape(-1, order="F"), (geom.n_int, 1)), geom.weight, pat, geom.n_q)
    assert _rel(op.elastic_values.cpu().numpy(), Kel) < 1e-12
    assert _rel(Kr.values.cpu().numpy(), 0.75 * Ko + 0.25 * Kel) < 1e-10
    t = op.get_total_time()
    assert set(t) >= {"reduction", "stress_tangent", "build_F", "build_tangent_local"}
    with pytest.raises(ValueError):
        ConstitutiveOperator.from_mesh(m, [W.HOMO]).constitutive_problem_stress(U)


def test_operator_api_2d(gpu):
    print("case: operator API 2D", flush=True)
    m = mesh.slope_2d_p2(h=2.0, x1=6.0, x2=4.0, x3=6.0, y1=4.0, y2=4.0)
    op = ConstitutiveOperator.from_mesh(m, [W.HOMO], Davis_type="B")
    geom, pat = op.geom, op.pattern.host
    B = oracle.strain_operator(geom.dphi, m.elem, m.n_nodes, geom.n_q)
    rng = np.random.default_rng(4)
    U = rng.normal(0, 2e-3, (2, m.n_nodes)) * m.q_mask
    op.reduction(1.1)
    F, K = op.build_F_K_tangent_reduced(U)
    u = U.reshape(-1, order="F")
    E = (B @ u).reshape(-1, 3)
    cb, sp = oracle.davis_reduction(np.full(geom.n_int, 6.0), np.full(geom.n_int, np.pi / 4), np.zeros(geom.n_int),
                                    1.1, "B")
    S, DS, br = oracle.mc2d(E, cb, sp, np.full(geom.n_int, W.HOMO.shear), np.full(geom.n_int, W.HOMO.lame))
    Fo = B.T @ (geom.weight[:, None] * S[:, :3]).reshape(-1)
    Fo[pat.row_free == 0] = 0
    assert _rel(F.reshape(-1, order="F"), Fo) < 1e-10
    import scipy.sparse as sp_
    blocks = [sp_.csr_matrix(geom.weight[g] * DS[g].reshape(3, 3, order="F")) for g in range(geom.n_int)]
    Kg = (B.T @ sp_.block_diag(blocks, format="csr") @ B).tocsr()
    rows = np.repeat(np.arange(pat.n_rows), np.diff(pat.indptr))
    Ko = np.asarray(Kg[rows, pat.indices]).ravel()
    Ko[pat.constrained_diag_positions] = 1.0
    assert _rel(K.values.cpu().numpy(), Ko) < 1e-10


def test_p4_element_matrices_mfma_vs_oracle(gpu):
    """P4 element tangents on the FP64 matrix cores vs the oracle's dense B^T D B, then the row
    gather into the pattern vs the oracle rows kernel."""
    print("case: P4 MFMA element matrices", flush=True)
    m = mesh.slope_3d(h=5.0, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=1, elem_type="P4")
    geom = fem.geometry(m.coord, m.elem, "P4")
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    dp = DevicePattern(pat, geom)
    dev = dp.indptr.device
    rng = np.random.default_rng(31)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    DSd = torch.from_numpy(DS).to(dev)
    ke = dp.element_matrices_p4(DSd).cpu().numpy()
    keo = oracle.element_matrices_3d(geom.dphi, DS, geom.weight, geom.n_elem, geom.n_q)
    assert _rel(ke, keo) < 1e-12
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    dp.tangent(DSd, K)
    Ko = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    assert _rel(K.cpu().numpy(), Ko) < 1e-12


def test_p2_node_gather_elastic_blend_vs_oracle(gpu, asm3d):
    """Element matrices + node gather with the assembled-elastic K_r blend (1-r) K + r K_el and the
    FP32 copy == the oracle rows kernel blended the same way (constrained rows blended too)."""
    m, geom, pat = asm3d
    dp = DevicePattern(pat, geom)
    dev = dp.indptr.device
    rng = np.random.default_rng(33)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    DSd = torch.from_numpy(DS).to(dev)
    ke = dp.element_matrices_3d(DSd).cpu().numpy()
    keo = oracle.element_matrices_3d(geom.dphi, DS, geom.weight, geom.n_elem, geom.n_q)
    assert _rel(ke, keo) < 1e-12
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    K32 = torch.empty(pat.nnz, dtype=torch.float32, device=dev)
    el = rng.normal(size=pat.nnz)
    el[pat.constrained_diag_positions] = 1.0    # K_el, like K, has unit constrained rows
    dp.tangent(DSd, K, K32, r=0.2, elastic=torch.from_numpy(el).to(dev))
    Ko = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    exp = 0.8 * Ko + 0.2 * el
    assert _rel(K.cpu().numpy(), exp) < 1e-12
    assert np.all(K.cpu().numpy()[pat.constrained_diag_positions] == 1.0)
    np.testing.assert_array_equal(K32.cpu().numpy(), K.cpu().numpy().astype(np.float32))


@pytest.mark.parametrize("elem_type", ["P1", "P2", "P4"])
def test_node_gather_kr_fold_vs_oracle(gpu, elem_type):
    """Element matrices with the regularisation folded into the point tangents, gathered per node:
    K_r == (1-r) K(DS) + r K(D_el) of the oracle rows kernel, FP32 copy exact, bitwise repeatable."""
    print(f"case: node gather K_r fold {elem_type}", flush=True)
    nz = 1 if elem_type == "P4" else 3
    m = mesh.slope_3d(h=5.0 if elem_type == "P4" else 2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0,
                      z_len=10.0, nz=nz, elem_type=elem_type)
    geom = fem.geometry(m.coord, m.elem, elem_type)
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    dp = DevicePattern(pat, geom)
    assert dp.tangent_method == "nodes" and dp.ns is not None
    dev = dp.indptr.device
    rng = np.random.default_rng(41)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    G = rng.uniform(1e3, 2e4, geom.n_int)
    lam = rng.uniform(1e3, 3e4, geom.n_int)
    vol = np.array([1, 1, 1, 0, 0, 0.0])
    idn = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
    Del = (lam[:, None, None] * np.outer(vol, vol)[None] + 2 * G[:, None, None] * idn[None]).reshape(-1, 36)
    r = 0.3
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    K32 = torch.empty(pat.nnz, dtype=torch.float32, device=dev)
    dp.tangent(t(DS), K, K32, r=r, mat=(t(G), t(lam)))
    Kt = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    Kel = oracle.tangent_3d_rows(geom.dphi, Del, geom.weight, pat, geom.n_q)
    exp = (1 - r) * Kt + r * Kel
    exp[pat.constrained_diag_positions] = 1.0
    assert _rel(K.cpu().numpy(), exp) < 1e-12
    np.testing.assert_array_equal(K32.cpu().numpy(), K.cpu().numpy().astype(np.float32))
    K2 = torch.empty_like(K)
    dp.tangent(t(DS), K2, r=r, mat=(t(G), t(lam)))
    assert torch.equal(K, K2)
    # unregularised node gather == oracle rows kernel
    dp.tangent(t(DS), K2)
    exp0 = Kt.copy()
    exp0[pat.constrained_diag_positions] = 1.0
    assert _rel(K2.cpu().numpy(), exp0) < 1e-12


@pytest.mark.parametrize("rule", ["P2", "P4"])
def test_p2_element_kernels_vs_oracle(gpu, rule):
    """Full P2 element matrices with K_r folded, on an odd element count (one wavefront per element;
    the 11-point rule, and the 24 points of the P4 rule as the P4 multigrid's P2 level) give the
    oracle's regularised element matrices."""
    print(f"case: P2 element kernel, {rule} points", flush=True)
    from slope_stability_petsc4py_amd import _lib
    from slope_stability_petsc4py_amd.fem import quadrature_3d
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=7.5, nz=3)
    geom = fem.geometry(m.coord, m.elem, "P2", quad=quadrature_3d(rule))
    n_e = geom.n_elem - 1 if geom.n_elem % 2 == 0 else geom.n_elem   # odd count
    dev = gpu
    rng = np.random.default_rng(43)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    G = rng.uniform(1e3, 2e4, geom.n_int)
    lam = rng.uniform(1e3, 3e4, geom.n_int)
    vol = np.array([1, 1, 1, 0, 0, 0.0])
    idn = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
    Del = (lam[:, None, None] * np.outer(vol, vol)[None] + 2 * G[:, None, None] * idn[None]).reshape(-1, 36)
    r = 0.25
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    n_int = n_e * geom.n_q
    dphi = np.ascontiguousarray(geom.dphi[:, :n_int, :])     # the first n_e elements only
    wt, DS, G, lam, Del = geom.weight[:n_int], DS[:n_int], G[:n_int], lam[:n_int], Del[:n_int]
    ke = torch.full((n_e, 30, 30), np.nan, dtype=torch.float64, device=dev)
    args = [t(dphi), t(DS), t(wt), t(G), t(lam)]
    _lib.call("ssx_element_matrices_3d", 10, int(args[0].data_ptr()), int(args[1].data_ptr()),
              int(args[2].data_ptr()), n_e, geom.n_q, r, int(args[3].data_ptr()), int(args[4].data_ptr()),
              int(ke.data_ptr()), 0)
    torch.cuda.synchronize()
    exp = oracle.element_matrices_3d(dphi, (1 - r) * DS + r * Del, wt, n_e, geom.n_q)
    got = ke.cpu().numpy()
    assert np.isfinite(got).all()
    assert _rel(got, exp) < 1e-12


def _upper_blocks(ke):
    """(n_e, 55, 3, 3) upper node-pair blocks (a <= b, row-major) of (n_e, 30, 30) element matrices."""
    kef = ke.reshape(ke.shape[0], 10, 3, 10, 3)
    a, b = np.triu_indices(10)
    return np.ascontiguousarray(kef[:, a, :, b, :].transpose(1, 0, 2, 3))


@pytest.mark.parametrize("fold", [False, True])
def test_p2_element_blocks_vs_oracle(gpu, fold):
    """Upper node-pair blocks of the P2 element matrices (persistent kernel, odd element count so
    the last sweep of the grid is partial) == the upper blocks of the oracle's element matrices,
    with and without the K_r fold; repeatable bit for bit."""
    print(f"case: P2 element blocks, fold={fold}", flush=True)
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=7.5, nz=3)
    geom = fem.geometry(m.coord, m.elem, "P2")
    n_e = geom.n_elem - 1 if geom.n_elem % 2 == 0 else geom.n_elem
    dev = gpu
    rng = np.random.default_rng(47)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    G = rng.uniform(1e3, 2e4, geom.n_int)
    lam = rng.uniform(1e3, 3e4, geom.n_int)
    vol = np.array([1, 1, 1, 0, 0, 0.0])
    idn = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
    Del = (lam[:, None, None] * np.outer(vol, vol)[None] + 2 * G[:, None, None] * idn[None]).reshape(-1, 36)
    r = 0.25 if fold else 0.0
    n_int = n_e * geom.n_q
    dphi = np.ascontiguousarray(geom.dphi[:, :n_int, :])
    wt, DS, G, lam, Del = geom.weight[:n_int], DS[:n_int], G[:n_int], lam[:n_int], Del[:n_int]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    args = [t(dphi), t(DS), t(wt), t(G), t(lam)]
    kb = [torch.full((n_e, 55, 3, 3), np.nan, dtype=torch.float64, device=dev) for _ in range(2)]
    for out in kb:
        _lib.call("ssx_element_blocks_3d", 10, int(args[0].data_ptr()), int(args[1].data_ptr()),
                  int(args[2].data_ptr()), n_e, geom.n_q, r, int(args[3].data_ptr()) if fold else None,
                  int(args[4].data_ptr()) if fold else None, None, int(out.data_ptr()), 0)
    torch.cuda.synchronize()
    exp = _upper_blocks(oracle.element_matrices_3d(dphi, (1 - r) * DS + r * Del, wt, n_e, geom.n_q))
    got = kb[0].cpu().numpy()
    assert np.isfinite(got).all()
    assert _rel(got, exp) < 1e-12
    assert torch.equal(kb[0], kb[1])


def _fan_mesh_p2(n_sphere=40, seed=0):
    """P2 tetrahedra fanned around one vertex: the origin joined to the hull triangles of points on
    the unit sphere (76 elements share the centre for 40 points), outer vertices 0..9 fixed."""
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(n_sphere, 3))
    P /= np.linalg.norm(P, axis=1)[:, None]
    tri = ConvexHull(P).simplices
    coord4 = np.vstack((np.zeros((1, 3)), P)).T
    elem4 = np.vstack((np.zeros(len(tri), dtype=np.int64), (tri + 1).T))
    x = coord4[:, elem4]
    det = np.einsum("ij,ij->j", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], axis=0), x[:, 3] - x[:, 0])
    neg = det < 0
    elem4[1, neg], elem4[2, neg] = elem4[2, neg].copy(), elem4[1, neg].copy()
    coord, elem = mesh.elevate_tet4_to_tet10(coord4, elem4)
    q = np.ones((3, coord.shape[1]), dtype=bool)
    q[:, 1:11] = False
    return coord, elem, q


def test_p2_tangent_node_with_more_than_64_elements(gpu):
    """A node with more element slots than a wavefront has lanes (the block gather reads a node's
    slot list 64 slots at a time): K_r on a P2 fan mesh (76 elements around the centre) equals the
    oracle rows kernel; the force gather likewise."""
    print("case: P2 fan mesh, node with 76 elements", flush=True)
    coord, elem, q = _fan_mesh_p2()
    geom = fem.geometry(coord, elem, "P2")
    pat = pattern.build_owned_pattern(elem, q)
    dp = DevicePattern(pat, geom, coord=coord)
    assert int(np.diff(dp.ns.ptr).max()) > 64
    dev = dp.indptr.device
    rng = np.random.default_rng(59)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    G = rng.uniform(1e3, 2e4, geom.n_int)
    lam = rng.uniform(1e3, 3e4, geom.n_int)
    vol = np.array([1, 1, 1, 0, 0, 0.0])
    idn = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
    Del = (lam[:, None, None] * np.outer(vol, vol)[None] + 2 * G[:, None, None] * idn[None]).reshape(-1, 36)
    r = 0.1
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    dp.tangent(t(DS), K, r=r, mat=(t(G), t(lam)))
    exp = (1 - r) * oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q) + \
        r * oracle.tangent_3d_rows(geom.dphi, Del, geom.weight, pat, geom.n_q)
    exp[pat.constrained_diag_positions] = 1.0
    assert _rel(K.cpu().numpy(), exp) < 1e-12
    S = rng.normal(size=(geom.n_int, 6))
    fe = torch.empty(geom.n_elem * 30, dtype=torch.float64, device=dev)
    F = torch.empty(pat.n_rows, dtype=torch.float64, device=dev)
    dp.force(t(S), fe, F)
    Fo = oracle.force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q)
    assert _rel(F.cpu().numpy(), Fo) < 1e-12


def test_p2_tangent_with_branch_ids_vs_oracle(gpu):
    """Tangent from the return mapping's own DS and branch ids (elastic points in closed form) on a
    stratified strain field covering every branch: K_r and K equal the oracle rows kernel on the
    same DS to 1e-12, agree with the path without branch ids, and repeat bit for bit."""
    print("case: P2 tangent with branch ids", flush=True)
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=3)
    geom = fem.geometry(m.coord, m.elem, "P2")
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    dp = DevicePattern(pat, geom)
    dev = dp.indptr.device
    n = geom.n_int
    E, _b = W.stratified_strains(n, seed=5)
    cb, sp = np.full(n, 4 * np.sqrt(3)), np.full(n, 1 / np.sqrt(3))
    G, lam = np.full(n, W.HOMO.shear), np.full(n, W.HOMO.lame)
    S, DSd, br = constitutive_problem_3d(E, cb, sp, G, lam, return_tangent=True, return_branch=True)
    hist = np.bincount(br.cpu().numpy().astype(np.int64), minlength=5)
    assert (hist > 0).all(), hist
    DS = DSd.cpu().numpy()
    Gd, lamd = (torch.from_numpy(a).to(dev) for a in (G, lam))
    vol = np.array([1, 1, 1, 0, 0, 0.0])
    idn = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
    Del = (lam[:, None, None] * np.outer(vol, vol)[None] + 2 * G[:, None, None] * idn[None]).reshape(-1, 36)
    Kt = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    Kel = oracle.tangent_3d_rows(geom.dphi, Del, geom.weight, pat, geom.n_q)
    for r in (0.0, 1e-4, 0.5):
        exp = (1 - r) * Kt + r * Kel
        exp[pat.constrained_diag_positions] = 1.0
        K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
        K32 = torch.empty(pat.nnz, dtype=torch.float32, device=dev)
        dp.tangent(DSd, K, K32, r=r, mat=(Gd, lamd), branch=br)
        assert _rel(K.cpu().numpy(), exp) < 1e-12, r
        np.testing.assert_array_equal(K32.cpu().numpy(), K.cpu().numpy().astype(np.float32))
        K2 = torch.empty_like(K)
        dp.tangent(DSd, K2, r=r, mat=(Gd, lamd), branch=br)
        assert torch.equal(K, K2)
        dp.tangent(DSd, K2, r=r, mat=(Gd, lamd))
        assert _rel(K.cpu().numpy(), K2.cpu().numpy()) < 1e-13
    with pytest.raises(ValueError):
        dp.tangent(DSd, K2, r=0.1, branch=br)


@pytest.mark.parametrize("kind", ["slope", "fan"])
def test_p2_affine_geometry_kernels_vs_oracle(gpu, kind):
    """Strain, force and tangent of straight-sided P2 tetrahedra with the gradients formed in the
    kernels from the per-element factors: equal to the oracle on the per-point dphi and w to 1e-12,
    to the per-point kernels to 1e-13, and the tangent repeats bit for bit (a slope mesh whose
    point count is not a multiple of the strain block, and the fan mesh of 76 elements)."""
    print(f"case: P2 affine geometry kernels, {kind}", flush=True)
    if kind == "slope":
        m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=3)
        coord, elem, q = m.coord, m.elem, m.q_mask
    else:
        coord, elem, q = _fan_mesh_p2()
    geom = fem.geometry(coord, elem, "P2")
    pat = pattern.build_owned_pattern(elem, q)
    aff = fem.affine_geometry(coord, elem, "P2")
    assert aff is not None
    dp = DevicePattern(pat, geom, coord=coord, affine=aff)
    pts = DevicePattern(pat, geom, coord=coord)
    assert dp.affine and not pts.affine
    dev = dp.indptr.device
    n = geom.n_int
    if kind == "slope":
        assert n % 256 != 0
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    rng = np.random.default_rng(61)
    u = rng.normal(0, 1e-3, 3 * coord.shape[1])
    E = torch.empty((n, 6), dtype=torch.float64, device=dev)
    E2 = torch.empty_like(E)
    dp.strain(t(u), E)
    pts.strain(t(u), E2)
    assert _rel(E.cpu().numpy(), oracle.strain_3d(geom.dphi, u, pat.elem_dofs, geom.n_q)) < 1e-12
    assert _rel(E.cpu().numpy(), E2.cpu().numpy()) < 1e-13
    S = rng.normal(size=(n, 6))
    fe = torch.empty(geom.n_elem * 30, dtype=torch.float64, device=dev)
    F = torch.empty(pat.n_rows, dtype=torch.float64, device=dev)
    F2 = torch.empty_like(F)
    dp.force(t(S), fe, F)
    pts.force(t(S), fe, F2)
    assert _rel(F.cpu().numpy(), oracle.force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q)) < 1e-12
    assert _rel(F.cpu().numpy(), F2.cpu().numpy()) < 1e-13
    # tangent from the return mapping's DS and branch ids, K_r folded
    Es, _b = W.stratified_strains(n, seed=8)
    cb, sp = np.full(n, 4 * np.sqrt(3)), np.full(n, 1 / np.sqrt(3))
    G, lam = np.full(n, W.HOMO.shear), np.full(n, W.HOMO.lame)
    _S, DSd, br = constitutive_problem_3d(Es, cb, sp, G, lam, return_tangent=True, return_branch=True)
    DS = DSd.cpu().numpy()
    vol = np.array([1, 1, 1, 0, 0, 0.0])
    idn = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
    Del = (lam[:, None, None] * np.outer(vol, vol)[None] + 2 * G[:, None, None] * idn[None]).reshape(-1, 36)
    Kt = oracle.tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    Kel = oracle.tangent_3d_rows(geom.dphi, Del, geom.weight, pat, geom.n_q)
    for r in (0.0, 0.25):
        exp = (1 - r) * Kt + r * Kel
        exp[pat.constrained_diag_positions] = 1.0
        K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
        K32 = torch.empty(pat.nnz, dtype=torch.float32, device=dev)
        dp.tangent(DSd, K, K32, r=r, mat=(t(G), t(lam)), branch=br)
        assert _rel(K.cpu().numpy(), exp) < 1e-12, r
        np.testing.assert_array_equal(K32.cpu().numpy(), K.cpu().numpy().astype(np.float32))
        K2 = torch.empty_like(K)
        dp.tangent(DSd, K2, r=r, mat=(t(G), t(lam)), branch=br)
        assert torch.equal(K, K2)
        pts.tangent(DSd, K2, r=r, mat=(t(G), t(lam)), branch=br)
        assert _rel(K.cpu().numpy(), K2.cpu().numpy()) < 1e-13
        # without branch ids and without the fold
        dp.tangent(DSd, K2, r=r, mat=(t(G), t(lam)))
        assert _rel(K.cpu().numpy(), K2.cpu().numpy()) < 1e-13
        # the packed symmetric tangent (21 values per point)
        DSp = DSd[:, [s_ + 6 * m_ for s_ in range(6) for m_ in range(s_, 6)]].contiguous()
        dp.tangent(DSp, K2, K32, r=r, mat=(t(G), t(lam)), branch=br, sym=True)
        assert _rel(K2.cpu().numpy(), exp) < 1e-12
        assert _rel(K.cpu().numpy(), K2.cpu().numpy()) < 1e-13
        np.testing.assert_array_equal(K32.cpu().numpy(), K2.cpu().numpy().astype(np.float32))
        with pytest.raises(ValueError):
            pts.tangent(DSp, K2, r=r, mat=(t(G), t(lam)), branch=br, sym=True)
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    dp.tangent(DSd, K)
    exp = Kt.copy()
    exp[pat.constrained_diag_positions] = 1.0
    assert _rel(K.cpu().numpy(), exp) < 1e-12


@pytest.mark.parametrize("n", [1, 64, 200, 1000])
@pytest.mark.parametrize("f64", [False, True])
def test_sym_tiles_gemv_vs_numpy(gpu, n, f64):
    """Coarse solve with the symmetric tile store == the dense product with the full matrix."""
    from slope_stability_petsc4py_amd.device import ptr
    from slope_stability_petsc4py_amd.linear import sym_tile_sizes
    print(f"case: sym tiles gemv n={n} f64={f64}", flush=True)
    rng = np.random.default_rng(n)
    M = rng.normal(size=(n, n))
    A = M + M.T
    x = rng.normal(size=n)
    dt = torch.float64 if f64 else torch.float32
    te, pe = sym_tile_sizes(n)
    nt = (n + 63) // 64
    assert te == nt * (nt + 1) // 2 * 4096 and pe == nt * nt * 64
    Ad = torch.from_numpy(A).to(gpu)              # FP64 source, packed into the level precision
    tiles = torch.full((te,), np.nan, dtype=dt, device=gpu)
    part = torch.empty(pe, dtype=dt, device=gpu)
    y = torch.full((n,), np.nan, dtype=dt, device=gpu)
    _lib.call("ssx_sym_tiles_pack", n, ptr(Ad), 1, ptr(tiles), int(f64), 0)
    xd = torch.from_numpy(x).to(dt).to(gpu)
    _lib.call("ssx_sym_tiles_gemv_f64" if f64 else "ssx_sym_tiles_gemv_f32", n, ptr(tiles), ptr(xd), ptr(part),
              ptr(y), 0)
    got = y.cpu().numpy().astype(np.float64)
    exp = A @ x
    assert _rel(got, exp) < (1e-13 if f64 else 2e-6)
    # FP32 source packed into FP32 tiles gives the same product
    if not f64:
        A32 = torch.from_numpy(A.astype(np.float32)).to(gpu)
        _lib.call("ssx_sym_tiles_pack", n, ptr(A32), 0, ptr(tiles), 0, 0)
        _lib.call("ssx_sym_tiles_gemv_f32", n, ptr(tiles), ptr(xd), ptr(part), ptr(y), 0)
        assert _rel(y.cpu().numpy().astype(np.float64), exp) < 2e-6


@pytest.mark.parametrize("n", [1, 63, 64, 65, 300, 1000])
@pytest.mark.parametrize("f64", [False, True])
def test_spd_inverse_tiles_vs_numpy(gpu, n, f64):
    """Explicit inverse of an SPD matrix (blocked Cholesky, triangular inverse, W^T W on 64 x 64
    tiles) into the symmetric tile store: the coarse solve with it equals numpy.linalg.solve; the
    computation and the stored tiles in FP64 or FP32; repeated calls give the same bits."""
    from slope_stability_petsc4py_amd.device import ptr
    from slope_stability_petsc4py_amd.linear import spd_inverse_work, sym_tile_sizes
    print(f"case: spd inverse n={n} f64={f64}", flush=True)
    rng = np.random.default_rng(7 + n)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = (Q * np.logspace(0, 4, n)) @ Q.T           # condition number 1e4
    A = 0.5 * (A + A.T)
    x = rng.normal(size=n)
    dt = torch.float64 if f64 else torch.float32
    te, pe = sym_tile_sizes(n)
    work = torch.empty(spd_inverse_work(n), dtype=dt, device=gpu)
    info = torch.full((1,), -7, dtype=torch.int32, device=gpu)
    tiles = torch.full((te,), np.nan, dtype=dt, device=gpu)
    Ad = torch.from_numpy(A).to(gpu)
    _lib.call("ssx_spd_inverse_tiles", n, ptr(Ad), int(f64), ptr(work), ptr(tiles), int(f64), ptr(info), 0)
    assert int(info.item()) == 0
    part = torch.empty(pe, dtype=dt, device=gpu)
    y = torch.empty(n, dtype=dt, device=gpu)
    _lib.call("ssx_sym_tiles_gemv_f64" if f64 else "ssx_sym_tiles_gemv_f32", n, ptr(tiles),
              ptr(torch.from_numpy(x).to(dt).to(gpu)), ptr(part), ptr(y), 0)
    exp = np.linalg.solve(A, x)
    assert _rel(y.cpu().numpy().astype(np.float64), exp) < (1e-10 if f64 else 2e-3)
    tiles2 = torch.empty_like(tiles)
    _lib.call("ssx_spd_inverse_tiles", n, ptr(Ad), int(f64), ptr(work), ptr(tiles2), int(f64), ptr(info), 0)
    assert torch.equal(tiles, tiles2)
    if n >= 64:
        # an indefinite matrix reports its first non-positive pivot
        B = A.copy()
        B[n // 2, n // 2] = -1e6
        _lib.call("ssx_spd_inverse_tiles", n, ptr(torch.from_numpy(B).to(gpu)), int(f64), ptr(work), ptr(tiles2),
                  int(f64), ptr(info), 0)
        assert 0 < int(info.item()) <= n // 2 + 1


@pytest.mark.parametrize("f64", [False, True])
def test_inv_diag_vs_scipy(gpu, asm3d, f64):
    """Inverse diagonal of the CSR operator (Jacobi/Chebyshev smoother) == 1 / diag, 0 on a zero
    diagonal."""
    from slope_stability_petsc4py_amd.device import ptr
    m, geom, pat = asm3d
    dp = DevicePattern(pat, geom)
    rng = np.random.default_rng(31)
    val = rng.uniform(0.5, 2.0, pat.nnz)
    diag_pos = [pat.indptr[i] + int(np.flatnonzero(pat.indices[pat.indptr[i]:pat.indptr[i + 1]] == i)[0])
                for i in (0, 7, pat.n_rows - 1)]
    val[diag_pos[1]] = 0.0
    K = dp.new_csr()
    K.values.copy_(torch.from_numpy(val))
    K.values32.copy_(torch.from_numpy(val.astype(np.float32)))
    dt = torch.float64 if f64 else torch.float32
    dinv = torch.full((pat.n_rows,), np.nan, dtype=dt, device=gpu)
    _lib.call("ssx_inv_diag_f64" if f64 else "ssx_inv_diag_f32", pat.n_rows, 0, ptr(K.indptr), ptr(K.indices),
              ptr(K.values if f64 else K.values32), ptr(dinv), 0)
    import scipy.sparse as sp
    A = sp.csr_matrix((val, pat.indices, pat.indptr), shape=(pat.n_rows, pat.n_rows))
    d = A.diagonal()
    exp = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    got = dinv.cpu().numpy().astype(np.float64)
    assert got[7] == 0.0
    np.testing.assert_allclose(got, exp, rtol=1e-15 if f64 else 1e-6, atol=0)
