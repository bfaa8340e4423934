"""The CPU restatement of the plain 2D SSR (oracle/ssr2d.py), pinned three ways:

* its vectorised return mapping equals the loop form `oracle.mc2d` (itself pinned by the
  closest-point projection) entry for entry, on strains spread over all five branches;
* its sparse tangent K = B^T diag(w DS) B is the derivative of its internal force F = B^T (w S)
  (central differences on a small slope), and K_el is the tangent of the elastic branch;
* its SSR run on the reference's 2D homogeneous slope (h = 1: 1300 P2 triangles, 7-point rule)
  reproduces the reference PETSc run's lambda history (tests/golden/ssr_histories.npz), FoS within
  1e-3.
"""

from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import ssr2d
from slope_stability_petsc4py_amd import cases, fem
from test_oracle_mc2d import _branch_mix

GOLDEN = Path(__file__).resolve().parent / "golden" / "ssr_histories.npz"


def _problem(h):
    c = cases.homo2d(h=h)
    m = c.mesh
    g = fem.geometry(m.coord, m.elem, m.elem_type)
    P = ssr2d.Problem2D(m, g, c.materials[0])
    f_int = np.zeros((2, g.n_int))
    f_int[1] = -c.materials[0].gamma_sat
    f = fem.volume_load(g, m.elem, m.n_nodes, f_int)
    f[~m.q_mask.astype(bool)] = 0.0
    return c, P, f.T.reshape(-1).copy()


@pytest.mark.parametrize("seed", [3, 11])
def test_vectorised_return_mapping_equals_loop_form(seed):
    E = np.concatenate([_branch_mix(1200, seed), np.zeros((3, 3)), np.tile([2e-3, 2e-3, 0.0], (3, 1))])
    n = E.shape[0]
    rng = np.random.default_rng(seed)
    args = (rng.uniform(5.0, 30.0, n), rng.uniform(0.2, 0.7, n), np.full(n, 15384.6), np.full(n, 23076.9))
    S1, D1, b1 = oracle.mc2d(E, *args)
    S2, D2, b2 = ssr2d.mc2d_vec(E, *args)
    assert np.bincount(b1, minlength=5).min() > 0
    np.testing.assert_array_equal(b1, b2)
    np.testing.assert_allclose(S2, S1, rtol=0, atol=1e-12 * np.abs(S1).max())
    np.testing.assert_allclose(D2, D1, rtol=0, atol=1e-12 * np.abs(D1).max())


def test_sparse_tangent_is_derivative_of_force():
    _c, P, f = _problem(h=5.0)
    P.reduction(1.3)
    rng = np.random.default_rng(0)
    # a displacement that puts points on every branch: a scaled gravity-like field plus noise
    U = np.zeros(P.B.shape[1])
    U[P.free] = rng.normal(0, 2e-3, P.free.size)
    F, DS = P.stress(U, tangent=True)
    br = ssr2d.mc2d_vec((P.B @ U).reshape(-1, 3), P.cb, P.sp, P.G, P.lame)[2]
    assert np.unique(br).size >= 3
    K = P.K_r(DS, 0.0)
    for k in range(3):
        v = np.zeros_like(U)
        v[P.free] = rng.normal(0, 1.0, P.free.size)
        h = 1e-7
        fd = (P.stress(U + h * v) - P.stress(U - h * v)) / (2 * h)
        Kv = K @ v
        Kv[np.setdiff1d(np.arange(U.size), P.free)] = 0.0
        assert np.abs(fd - Kv).max() < 1e-5 * np.abs(Kv).max()
    # elastic tangent: K_r at r = 1, and F(U) = K_el U for a tiny (elastic) U
    np.testing.assert_allclose((P.K_r(DS, 1.0) - P.K_el).data, 0.0, atol=1e-9 * np.abs(P.K_el.data).max())
    u = 1e-9 * U
    Fe = P.K_el @ u
    Fe[np.setdiff1d(np.arange(U.size), P.free)] = 0.0
    np.testing.assert_allclose(P.stress(u), Fe, rtol=0, atol=1e-9 * np.abs(Fe).max())


def test_ssr_2d_homogeneous_matches_reference_history():
    g = np.load(GOLDEN)
    ref = g["homo2d_lambda_hist"]
    c, P, f = _problem(h=1.0)
    p = c.params
    r = ssr2d.ssr(P, f, lambda_init=p.lambda_init, d_lambda_init=p.d_lambda_init, d_lambda_min=p.d_lambda_min,
                  d_lambda_diff_scaled_min=p.d_lambda_diff_scaled_min, step_max=p.step_max,
                  omega_max_stop=p.omega_max_stop, it_newt_max=p.it_newt_max, it_damp_max=p.it_damp_max, tol=p.tol,
                  r_min=p.r_min)
    assert abs(r["fos"] - ref[-1]) < 1e-3
    assert r["stop"] == "d_lambda_diff_scaled_min"
    assert r["lambda_hist"].size == ref.size
    np.testing.assert_allclose(r["lambda_hist"], ref, atol=2e-3)
    np.testing.assert_allclose(r["omega_hist"], g["homo2d_omega_hist"], rtol=2e-2)
