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
@pytest.mark.parametrize("pc", ["pmg", "pmg_agg", "cheb"])
def test_fgmres_solves_elastic_system(gpu, elastic3d, use32, pc):
    print(f"case: fgmres pc={pc} fp32={use32}", flush=True)
    m, op, K, f = elastic3d
    if pc == "pmg":
        prec = PMultigrid(m.elem, m.q_mask, use32=use32)
        assert prec.coarse == "dense"
    elif pc == "pmg_agg":
        # sparse P1 level + dense rigid-body-mode aggregate level
        prec = PMultigrid(m.elem, m.q_mask, use32=use32, coarse="aggregate", coord=m.coord)
    else:
        prec = ChebyshevJacobi(use32=use32)
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
    if pc == "pmg_agg":
        assert its < 120, its


def test_fgmres_mgs_order_matches_cgs2(gpu, elastic3d):
    """The reference's modified Gram-Schmidt order (orthogonalization="mgs") and the default
    CGS + one re-orthogonalisation pass build the same Krylov basis in exact arithmetic: the same
    solution, iteration count and residual history to rounding."""
    m, op, K, f = elastic3d
    b = torch.from_numpy(f).to(gpu)
    out = {}
    for orth in ("cgs2", "mgs"):
        prec = PMultigrid(m.elem, m.q_mask, use32=False)
        solver = DeflatedFGMRES(prec, tolerance=1e-8, max_iterations=300, orthogonalization=orth)
        solver.setup_preconditioner(K)
        solver.A_orthogonalize(K)
        x = solver.solve(K, b).cpu().numpy()
        out[orth] = (x, solver.last_residual_history)
    (x1, h1), (x2, h2) = out["cgs2"], out["mgs"]
    assert abs(len(h1) - len(h2)) <= 1
    k = min(len(h1), len(h2))
    np.testing.assert_allclose(h2[:k], h1[:k], rtol=1e-6)
    assert np.linalg.norm(x2 - x1) <= 1e-7 * np.linalg.norm(x1)
    with pytest.raises(ValueError):
        DeflatedFGMRES(None, orthogonalization="householder")


def test_sparse_galerkin_matches_dense_on_the_p1_pattern(gpu, elastic3d):
    """P^T A P on the P1 pattern (ssx_galerkin_csr) against the dense product (ssx_galerkin_dense):
    equal at the pattern's positions, and the dense product vanishes (to rounding) outside them."""
    m, op, K, f = elastic3d
    prec = PMultigrid(m.elem, m.q_mask, use32=False, coarse="aggregate", coord=m.coord)
    prec.setup(K)
    Kc = prec.Kc.to_scipy().toarray()
    from slope_stability_petsc4py_amd import _lib
    from slope_stability_petsc4py_amd.device import ptr, stream_ptr
    n_c = prec.n_c
    Ad = torch.zeros((n_c, n_c), dtype=torch.float64, device=gpu)
    _lib.call("ssx_galerkin_dense", K.n_rows, ptr(K.indptr), ptr(K.indices), ptr(K.values), K.nnz, ptr(prec.p_ptr),
              ptr(prec.p_idx), ptr(prec.p_w), n_c, prec._w_bound(K, "p"), ptr(Ad), ptr(prec._gwork), stream_ptr())
    D = Ad.cpu().numpy().T          # column-major fill
    fix = np.flatnonzero(~prec.c_free)
    D[fix, fix] = 1.0
    mask = prec.Kc.to_scipy().toarray() != 0
    pat = np.zeros_like(mask)
    cp = prec.c_pat
    for i in range(n_c):
        pat[i, cp.indices[cp.indptr[i]:cp.indptr[i + 1]]] = True
    scale = np.abs(D).max()
    assert np.abs(Kc - np.where(pat, D, 0.0)).max() <= 1e-12 * scale
    assert np.abs(D[~pat]).max() <= 1e-12 * scale
    assert mask.sum() > 0 and np.allclose(Kc, Kc.T, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("kind", ["dense", "csr"])
def test_galerkin_bitwise_repeatable_and_matches_scipy(gpu, elastic3d, kind):
    """The fixed-point Galerkin sums give the same bits on every call, and equal the FP64 product
    P^T K P of scipy to 1e-13 of its largest entry (the fixed-point resolution is ~1e-16 of it)."""
    import scipy.sparse as sps
    from slope_stability_petsc4py_amd import _lib
    from slope_stability_petsc4py_amd.device import ptr, stream_ptr
    m, op, K, f = elastic3d
    prec = PMultigrid(m.elem, m.q_mask, use32=False, coarse="aggregate", coord=m.coord)
    prec.setup(K)
    p_ptr, p_idx, p_w = prec._p_host
    P = sps.csr_matrix((p_w, p_idx, p_ptr), shape=(K.n_rows, prec.n_c))
    exp = (P.T @ K.to_scipy() @ P).toarray()
    wb = prec._w_bound(K, "p")
    runs = []
    for _ in range(2):
        if kind == "dense":
            out = torch.zeros((prec.n_c, prec.n_c), dtype=torch.float64, device=gpu)
            _lib.call("ssx_galerkin_dense", K.n_rows, ptr(K.indptr), ptr(K.indices), ptr(K.values), K.nnz,
                      ptr(prec.p_ptr), ptr(prec.p_idx), ptr(prec.p_w), prec.n_c, wb, ptr(out), ptr(prec._gwork),
                      stream_ptr())
            got = out.cpu().numpy().T
        else:
            out = torch.full((prec.Kc.nnz,), np.nan, dtype=torch.float64, device=gpu)
            _lib.call("ssx_galerkin_csr", K.n_rows, ptr(K.indptr), ptr(K.indices), ptr(K.values), K.nnz,
                      ptr(prec.p_ptr), ptr(prec.p_idx), ptr(prec.p_w), ptr(prec.PT.indptr), ptr(prec.PT.indices),
                      ptr(prec.PT.values), prec.n_c, ptr(prec.Kc.indptr), ptr(prec.Kc.indices), prec.Kc.nnz,
                      prec.kc_max_row, wb, ptr(out), ptr(prec._gwork), stream_ptr())
            cp = prec.c_pat
            got = sps.csr_matrix((out.cpu().numpy(), cp.indices, cp.indptr), shape=exp.shape).toarray()
            on = sps.csr_matrix((np.ones(cp.indices.size), cp.indices, cp.indptr), shape=exp.shape).toarray() != 0
            exp = np.where(on, exp, 0.0)
        runs.append(out)
        scale = np.abs(exp).max()
        assert np.abs(got - exp).max() <= 1e-13 * scale
    assert torch.equal(runs[0], runs[1])
    # through the preconditioner: the sparse P1 level is the same bits on a second setup
    kc0 = prec.Kc.values.clone()
    prec.setup(K)
    assert torch.equal(kc0, prec.Kc.values)


@pytest.mark.parametrize("coarse", ["dense", "aggregate"])
def test_p4_multigrid_coarse_levels_solve(gpu, coarse):
    """P4 -> P2 -> P1 V-cycle with the dense P1 level and with the sparse P1 level above the
    aggregate level (which needs the P2 level's coordinates): FGMRES solves the elastic P4 system."""
    print(f"case: P4 multigrid coarse={coarse}", flush=True)
    from slope_stability_petsc4py_amd.linear import PMultigridP4
    m = mesh.slope_3d(h=5.0, nz=1, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, elem_type="P4")
    op = ConstitutiveOperator.from_mesh(m, [HOMO])
    op.reduction(1.0)
    op.constitutive_problem_stress_tangent(np.zeros((3, m.n_nodes)))
    K = op.build_K_regularized(1e-4)
    f = op.gravity_load(m).T.reshape(-1).copy()
    prec = PMultigridP4(m, op, use32=True, coarse=coarse)
    assert prec.coarse.coarse == coarse
    solver = DeflatedFGMRES(prec, tolerance=1e-8, max_iterations=300)
    solver.setup_preconditioner(K)
    solver.A_orthogonalize(K)
    x = solver.solve(K, torch.from_numpy(f).to(gpu)).cpu().numpy()
    Ks = K.to_scipy()
    assert np.linalg.norm(Ks @ x - f) <= 1.1e-8 * np.linalg.norm(f)


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


@pytest.mark.parametrize("use32", [False, True])
@pytest.mark.parametrize("degree", [2, 3])
def test_fused_chebyshev_step_matches_unfused(gpu, elastic3d, use32, degree):
    """One launch per Chebyshev(Jacobi) step (residual + smoother + update fused, iterate
    ping-ponged) == the residual / update sequence it replaces, from a zero and a given guess."""
    from slope_stability_petsc4py_amd.linear import _Level
    print(f"case: fused chebyshev fp32={use32} degree={degree}", flush=True)
    m, op, K, f = elastic3d
    dt = torch.float32 if use32 else torch.float64
    rng = np.random.default_rng(degree)
    b = torch.from_numpy(rng.normal(size=K.n_rows)).to(gpu).to(dt)
    x0 = torch.from_numpy(rng.normal(size=K.n_rows)).to(gpu).to(dt)
    from slope_stability_petsc4py_amd import _lib
    out = {}
    from gpu_helpers import misaligned
    for fused in (False, True, "vec4"):
        # fused=True: one entry per lane (arrays off the 16-byte grid); "vec4": aligned quadruples
        lv = _Level(K if fused == "vec4" else misaligned(K), use32)
        lv.refresh()
        lv.fused = bool(fused)
        xa = torch.empty_like(b)
        lv.cheb(b, xa, degree, 0.1, zero_guess=True)
        xb = x0.clone()
        lv.cheb(b, xb, degree, 0.1, zero_guess=False)
        out[fused] = (xa.double().cpu().numpy(), xb.double().cpu().numpy())
    tol = 1e-5 if use32 else 1e-12
    for i in range(2):
        ref = out[False][i]
        assert np.abs(out[True][i] - ref).max() <= tol * np.abs(ref).max()
        assert np.abs(out["vec4"][i] - ref).max() <= tol * np.abs(ref).max()


def test_footing_limit_load_converges_to_prandtl(gpu):
    """Strip footing on a weightless associated Mohr-Coulomb layer (phi = 10 deg, c = 10 kPa,
    pressure q = c scaled by t): the limit-load continuation's last t approaches Prandtl's N_c =
    8.345 from above as the mesh is refined (measured 9.08 at h = 0.5 and 8.67 at h = 0.25), with
    the error at least halving per refinement; this pins the surface load and the limit-load
    continuation against a closed form."""
    out = []
    for h in (0.5, 0.25):
        print(f"case: footing2d_prandtl h={h}", flush=True)
        case = cases.footing2d_prandtl(h=h)
        res = cases.run_ll(cases.make_solver(case, precision="fp64"))
        out.append(res.limit_load)
        print("t_limit", res.limit_load, "Nc", case.extra["Nc"], "steps", len(res.t_hist), "wall", res.wall_time,
              flush=True)
    nc = cases.prandtl_nc(10.0)
    assert nc < out[1] < out[0]
    assert out[1] - nc < 0.55 * (out[0] - nc)
    assert out[1] < 1.05 * nc and out[0] < 1.10 * nc


def test_footing_under_gravity_fp32_matches_fp64(gpu):
    """Limit load of a strip footing on the layered slope under its own weight (constant gravity
    load, F(U) = f_g + t f): the FP32-preconditioner form follows the plain FP64 form's load path
    -- t at equal omega within 1e-3 -- and the branch starts from the gravity equilibrium
    (omega_0 = f^T U_g > 0, t = 0).  The FP32 form with the sparse P1 level above the dense
    aggregate level (the coarse form of large meshes) follows the same path."""
    print("case: footing3d_hetero h=2", flush=True)
    out = {}
    for prec in ("fp64", "fp32", "fp32_agg"):
        coarse = "aggregate" if prec.endswith("_agg") else "dense"
        res = cases.run_ll(cases.make_solver(cases.footing3d_hetero(h=2.0), precision=prec[:4], coarse=coarse))
        out[prec] = (np.asarray(res.omega_hist), np.asarray(res.t_hist))
        print(prec, "t", res.t_hist[-1], "omega", res.omega_hist[-1], "steps", len(res.t_hist), "newton",
              res.newton_iterations, "wall", res.wall_time, flush=True)
        assert res.t_hist[0] == 0.0 and res.omega_hist[0] > 0.0
        assert np.all(np.diff(res.omega_hist) > 0)
    w = min(out["fp64"][0][-1], out["fp32"][0][-1])
    t64 = np.interp(w, *out["fp64"])
    t32 = np.interp(w, *out["fp32"])
    assert abs(t64 - t32) <= 1e-3 * t64
    assert t64 > 1.0
    w = min(out["fp64"][0][-1], out["fp32_agg"][0][-1])
    assert abs(np.interp(w, *out["fp64"]) - np.interp(w, *out["fp32_agg"])) <= 1e-3 * t64


def test_p4_ssr_matches_reference_history():
    """The reference's P4 problem (SSR_hetero_ada_L1 elevated to P4, 616,322 free dofs, p-multigrid
    P4 -> P2 -> P1, FP32 preconditioner) to omega = 6.7e6: lambda lies between the reference's
    lambda at omega = 6.57e6 (1.566538) and at 7.0e6 (1.569076) and within 1e-3 of their linear
    interpolation (benchmarks/slope_stability_3D_hetero_SSR_default/archive/
    report_pmg_rank8_p2_levels_p4_omega7e6.md:72-80); the branch starts at lambda 1.0, 1.1 as the
    reference's (lambda_init 1.0, d_lambda_init 0.1)."""
    print("case: SSR hetero3d_L1 P4 fp32", flush=True)
    case = cases.hetero3d_L1("P4")
    assert case.mesh.elem.shape[0] == 35
    case.params.omega_max_stop = 6.7e6
    s = cases.make_solver(case, precision="fp32")
    assert int(s.free_mask.sum().item()) == 616322
    res = cases.run_ssr(s)
    lam = res.fos
    print("lambda", lam, "omega", res.omega_hist[-1], "steps", len(res.lambda_hist), "newton", res.newton_iterations,
          "wall", res.wall_time, flush=True)
    assert abs(res.omega_hist[-1] - 6.7e6) < 1.0
    lo, hi = 1.566538452, 1.569076398
    interp = lo + (6.7e6 - 6570273.222) / (7.0e6 - 6570273.222) * (hi - lo)
    assert lo < lam < hi
    assert abs(lam - interp) < 1e-3
    assert abs(res.lambda_hist[0] - 1.0) < 1e-12 and abs(res.lambda_hist[1] - 1.1) < 1e-12


def test_footing_config_size_forms_agree(gpu):
    """The strip-footing limit load at the size of the large LL configuration (528,504 P2 tets,
    2,194,269 dofs; sparse P1 level above the aggregate level): plain FP64, FP32 and learned + FP32
    forms followed to omega = 1000 give t at equal omega within 1e-3 of each other."""
    import time
    from pathlib import Path as _P
    from slope_stability_petsc4py_amd.learn import LearnedPredictor, NodeMLP, node_info
    weights = _P(cases.__file__).resolve().parent / "data" / "predictor_hetero3d_L1.npz"
    case = cases.footing3d_hetero(h=0.53)
    case.params.omega_max_stop = 1000.0
    assert case.mesh.n_elem > 500_000
    out = {}
    for form in ("plain_fp64", "fp32", "learned_fp32"):
        t0 = time.perf_counter()
        s = cases.make_solver(case, precision="fp64" if form == "plain_fp64" else "fp32")
        assert s.op.pattern.n_rows > 2_100_000 and s.linear.pc.coarse == "aggregate"
        pred = None
        if form == "learned_fp32":
            pred = LearnedPredictor(NodeMLP.load(weights), node_info(case.mesh, case.materials))
        res = cases.run_ll(s, predictor=pred)
        out[form] = (np.asarray(res.omega_hist), np.asarray(res.t_hist))
        print(f"case: footing config size {form}: t {res.t_hist[-1]:.6f} omega {res.omega_hist[-1]:.1f} "
              f"steps {len(res.t_hist)} newton {res.newton_iterations} wall {res.wall_time:.1f}s "
              f"total {time.perf_counter() - t0:.1f}s", flush=True)
        del s, res
        torch.cuda.empty_cache()
    ref = out["plain_fp64"]
    for form in ("fp32", "learned_fp32"):
        w = min(ref[0][-1], out[form][0][-1])
        t_ref = np.interp(w, *ref)
        assert abs(np.interp(w, *out[form]) - t_ref) <= 1e-3 * t_ref, form


def test_p4_layered_slope_fos_forms_agree(gpu):
    """The heterogeneous P4 configuration to its factor of safety: the reference's L2 slope mesh
    elevated to P4 (33,935 tets, 1.14M dofs) with the four layer materials, p-multigrid
    P4 -> P2 -> P1.  The FP32-preconditioner form and the plain FP64 form reach the same FoS within
    1e-3; at the last common omega the FP32 form's lambda is within 1e-3 of the FP64 form's."""
    import time
    case = cases.hetero3d_L2_layers()
    assert case.mesh.elem_type == "P4" and case.mesh.n_elem == 33_935
    out = {}
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        s = cases.make_solver(case, precision=prec)
        assert s.op.pattern.n_rows > 1_100_000
        res = cases.run_ssr(s)
        out[prec] = res
        print(f"case: P4 layered slope {prec}: FoS {res.fos:.6f} stop {res.stop_reason} steps "
              f"{len(res.lambda_hist)} newton {res.newton_iterations} fgmres {res.linear_iterations} "
              f"wall {res.wall_time:.1f}s total {time.perf_counter() - t0:.1f}s", flush=True)
        del s
        torch.cuda.empty_cache()
    a, b = out["fp64"], out["fp32"]
    assert a.stop_reason == "d_lambda_diff_scaled_min" and b.stop_reason == "d_lambda_diff_scaled_min"
    assert abs(a.fos - b.fos) < 1e-3
    w = min(a.omega_hist[-1], b.omega_hist[-1])
    assert abs(np.interp(w, a.omega_hist, a.lambda_hist) - np.interp(w, b.omega_hist, b.lambda_hist)) < 1e-3
