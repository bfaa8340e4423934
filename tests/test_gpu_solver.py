"""Device linear solvers and the end-to-end SSR against the reference's committed results."""

from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from slope_stability_petsc4py_amd import cases, mesh
from slope_stability_petsc4py_amd.constitutive import ConstitutiveOperator
from slope_stability_petsc4py_amd.linear import ChebyshevJacobi, DeflatedFGMRES, PMultigrid
from slope_stability_petsc4py_amd.workloads import HOMO

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def elastic3d():
    m = mesh.slope_3d(h=2.5, nz=4, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0)
    op = ConstitutiveOperator.from_mesh(m, [HOMO])
    K = op.elastic_matrix()
    f = op.gravity_load(m).T.reshape(-1).copy()
    return m, op, K, f


@pytest.mark.parametrize("use32", [False, True])
@pytest.mark.parametrize("pc", ["pmg", "cheb"])
def test_fgmres_solves_elastic_system(gpu, elastic3d, use32, pc):
    print(f"case: fgmres pc={pc} fp32={use32}", flush=True)
    m, op, K, f = elastic3d
    prec = PMultigrid(m.elem, m.q_mask, use32=use32) if pc == "pmg" else ChebyshevJacobi(use32=use32)
    solver = DeflatedFGMRES(prec, tolerance=1e-8, max_iterations=300)
    solver.setup_preconditioner(K)
    solver.A_orthogonalize(K)
    b = torch.from_numpy(f).to(gpu)
    x = solver.solve(K, b).cpu().numpy()
    Ks = K.to_scipy()
    x_ref = spla.spsolve(Ks.tocsc(), f)
    assert np.linalg.norm(Ks @ x - f) <= 1.1e-8 * np.linalg.norm(f)
    assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref)
    its = solver.iteration_collector.iterations[-1]
    if pc == "pmg":
        assert its < 60, its


def test_deflation_basis_reuses_previous_solutions(gpu, elastic3d):
    m, op, K, f = elastic3d
    solver = DeflatedFGMRES(PMultigrid(m.elem, m.q_mask, use32=True), tolerance=1e-6, max_iterations=200)
    solver.setup_preconditioner(K)
    b = torch.from_numpy(f).to(gpu)
    solver.A_orthogonalize(K)
    x = solver.solve(K, b)
    n0 = solver.iteration_collector.iterations[-1]
    solver.expand_deflation_basis(x)
    solver.A_orthogonalize(K)
    solver.solve(K, b)
    # the solution is in the A-orthonormal basis: the coarse guess solves the system outright
    assert solver.iteration_collector.iterations[-1] == 0 < n0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ssr_homo2d_matches_reference_fos(gpu, precision):
    """2D homogeneous SSR (benchmarks/run_2D_homo_SSR_capture): FoS within 1e-3 of the reference's
    PETSc run (1.2113326) and lambda history close to its history."""
    print(f"case: SSR homo2d {precision}", flush=True)
    h = np.load(GOLD / "ssr_histories.npz")
    ref = h["homo2d_lambda_hist"]
    s = cases.make_solver(cases.homo2d(), precision=precision)
    res = cases.run_ssr(s)
    print("fos", res.fos, "ref", ref[-1], "steps", len(res.lambda_hist), "newton", res.newton_iterations,
          "wall", res.wall_time, flush=True)
    assert abs(res.fos - ref[-1]) < 1e-3
    assert abs(res.lambda_hist[0] - ref[0]) < 1e-12 and abs(res.lambda_hist[1] - ref[1]) < 1e-12


def test_a_orthogonalize_matches_oracle_and_drops_dependent_columns(gpu, elastic3d):
    """Device A-orthonormalisation (no per-column host reads) == the sequential drop rule of
    linear/orthogonalize.py; an exactly dependent column is dropped; W^T A W = I on the kept set."""
    import oracle
    m, op, K, f = elastic3d
    Ks = K.to_scipy()
    rng = np.random.default_rng(7)
    n = K.n_rows
    free = m.q_mask.T.reshape(-1).astype(np.float64)
    B = rng.normal(size=(5, n)) * free
    B = np.vstack([B, B[1] + 2.0 * B[3]])       # newest column depends on two older ones
    solver = DeflatedFGMRES(PMultigrid(m.elem, m.q_mask, use32=False))
    solver.setup_preconditioner(K)
    for row in B:
        solver.expand_deflation_basis(torch.from_numpy(row).to(gpu))
    solver.A_orthogonalize(K)
    W = solver._W.cpu().numpy()
    Wo, So = oracle.a_orthogonalize(B, Ks)
    assert W.shape == Wo.shape == (5, n)
    assert np.abs(W - Wo).max() <= 1e-8 * np.abs(Wo).max()
    G = W @ (Ks @ W.T)
    assert np.abs(G - np.eye(5)).max() < 1e-8
    np.testing.assert_allclose(solver._AW.cpu().numpy(), (Ks @ W.T).T, rtol=0, atol=1e-9 * np.abs(G).max() * 1e3)


@pytest.mark.parametrize("case_name", ["homo2d_ll", "homo3d_ll_L1"])
def test_ll_matches_reference_limit_load(gpu, case_name):
    """Limit-load continuation (continuation/limit_load.py) on the reference's LL benchmarks
    (benchmarks/slope_stability_2D_homo_LL, slope_stability_3D_homo_LL on LL_homo_ada_L1): the load
    factor reached at the end of the branch within 1e-3 (relative) of the reference's last entry,
    and the elastic start on the reference's line omega = c t (c = f.K^-1 f; the reference's first
    omega target comes from an elastic solve at linear tolerance 1e-1, so only the ratio is fixed)."""
    print(f"case: LL {case_name}", flush=True)
    h = np.load(GOLD / "ll_histories.npz")
    key = "homo2d" if case_name == "homo2d_ll" else "homo3d"
    ref_t, ref_om = h[key + "_lambda_hist"], h[key + "_omega_hist"]
    case = cases.homo2d_ll() if case_name == "homo2d_ll" else cases.homo3d_ll_L1()
    s = cases.make_solver(case, precision="fp64" if case_name == "homo2d_ll" else "fp32")
    res = cases.run_ll(s)
    t = np.asarray(res.t_hist)
    om = np.asarray(res.omega_hist)
    print("t", t[-1], "ref", ref_t[-1], "steps", len(t), "wall", res.wall_time, flush=True)
    assert abs(t[-1] - ref_t[-1]) <= 1e-3 * ref_t[-1]
    assert abs(om[1] / t[1] - ref_om[1] / ref_t[1]) <= 1e-3 * ref_om[1] / ref_t[1]
    assert np.all(np.diff(t[:-1]) > 0)
