"""Pin the CPU oracle of the plane-strain return mapping (oracle.mc2d, a loop restatement of
constitutive/problem.py:143-475) without importing the reference.

The independent check is the closest-point projection: for associated Mohr-Coulomb with linear
isotropic elasticity the return mapping is the projection of the trial stress onto the yield set
in the metric of the elastic compliance, in principal space with sigma_33 as the third principal
stress (plane strain: eps_33 = 0).  ``_project`` solves that 3-variable quadratic programme by
exact active-set enumeration over the six Mohr-Coulomb planes, with no sorting and no branch
formulas, so it shares nothing with the restatement.  Around it: elastic and apex closed forms,
the yield surface and edge equalities of returned stresses, coaxiality, DS against central
differences of S, the tie cases of the eigenvalue sort (:173-184, :214-241), and planted mistakes
(the oracle's ``_mutation`` hook) that these checks must catch.
"""

from itertools import combinations

import numpy as np
import pytest

import oracle
from slope_stability_petsc4py_amd import workloads as W

CB, SP = 4.0 * np.sqrt(3.0), 1.0 / np.sqrt(3.0)
G, LAM = W.HOMO.shear, W.HOMO.lame


def _mat(n, cb=CB, sp=SP):
    return np.full(n, cb), np.full(n, sp), np.full(n, G), np.full(n, LAM)


def _planes(sp):
    """Rows a with a . sigma <= c_bar: (1 + sin phi) s_i - (1 - sin phi) s_j for every i != j."""
    rows = []
    for i in range(3):
        for j in range(3):
            if i != j:
                a = np.zeros(3)
                a[i] += 1.0 + sp
                a[j] -= 1.0 - sp
                rows.append(a)
    return np.array(rows)


def _project(sig_tr, cb, sp, G_, lam):
    """Closest point of {a_k . s <= cb} to sig_tr in the compliance metric (C = lam 11^T + 2G I):
    minimise (s - sig_tr)^T C^-1 (s - sig_tr); KKT: s = sig_tr - C A^T mu, A s = cb on the active
    set, mu >= 0.  Every active set of at most 3 independent planes is tried; the feasible KKT
    point with the smallest objective is the (unique) projection."""
    A = _planes(sp)
    C = lam * np.ones((3, 3)) + 2.0 * G_ * np.eye(3)
    Ci = np.linalg.inv(C)
    if np.all(A @ sig_tr <= cb * (1 + 1e-13)):
        return sig_tr.copy()
    best, best_obj = None, np.inf
    for k in (1, 2, 3):
        for act in combinations(range(6), k):
            Aa = A[list(act)]
            M = Aa @ C @ Aa.T
            if np.linalg.matrix_rank(M, tol=1e-9 * np.abs(M).max()) < k:
                continue
            mu = np.linalg.solve(M, Aa @ sig_tr - cb)
            if np.any(mu < -1e-10 * np.abs(mu).max()):
                continue
            s = sig_tr - C @ Aa.T @ mu
            if np.any(A @ s > cb + 1e-9 * (abs(cb) + np.abs(s).max())):
                continue
            d = s - sig_tr
            obj = d @ Ci @ d
            if obj < best_obj:
                best, best_obj = s, obj
    assert best is not None
    return best


def _expected_stress(E, cb, sp):
    """S (n, 4) from the projection of the trial principal stresses, mapped back onto the trial
    strain's principal directions (in-plane eigenvectors, e_z)."""
    out = np.zeros((E.shape[0], 4))
    for p, (ex, ey, g) in enumerate(E):
        w, V = np.linalg.eigh(np.array([[ex, g / 2], [g / 2, ey]]))
        e = np.array([w[0], w[1], 0.0])
        s = _project(LAM * e.sum() + 2 * G * e, cb, sp, G, LAM)
        T = s[0] * np.outer(V[:, 0], V[:, 0]) + s[1] * np.outer(V[:, 1], V[:, 1])
        out[p] = [T[0, 0], T[1, 1], T[0, 1], s[2]]
    return out


def _branch_mix(n, seed):
    """Trial strains spread over the five branches: random in-plane strains at several scales and
    orientations, with compressive and tensile volumetric parts."""
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-4.5, -2.0, n)
    th = rng.uniform(0, np.pi, n)
    e1 = mag * rng.uniform(-1, 1, n)
    e2 = e1 - mag * rng.uniform(0, 2, n)
    # a quarter each near the edges: nearly equal tensile in-plane eigenvalues above eps33 = 0
    # (left edge, s1 = s2), smaller in-plane eigenvalue close to eps33 = 0 (right edge, s2 = s3)
    q = n // 4
    e1[:q] = mag[:q] * rng.uniform(0.2, 1.0, q)
    e2[:q] = e1[:q] - mag[:q] * rng.uniform(0.0, 0.1, q)
    e2[q:2 * q] = mag[q:2 * q] * rng.uniform(-0.05, 0.05, q)
    e1[q:2 * q] = mag[q:2 * q] * rng.uniform(0.5, 1.5, q)
    c, s = np.cos(th), np.sin(th)
    ex = c * c * e1 + s * s * e2
    ey = s * s * e1 + c * c * e2
    g = 2 * c * s * (e1 - e2)
    return np.stack([ex, ey, g], axis=1)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_elastic_closed_form_with_sigma33():
    rng = np.random.default_rng(0)
    E = rng.normal(0, 1e-5, (40, 3))
    E[:, :2] -= 1e-4
    S, DS, br = oracle.mc2d(E, *_mat(40))
    assert np.all(br == 0)
    tr = E[:, 0] + E[:, 1]
    S_exp = np.stack([LAM * tr + 2 * G * E[:, 0], LAM * tr + 2 * G * E[:, 1], G * E[:, 2], LAM * tr], axis=1)
    np.testing.assert_allclose(S, S_exp, rtol=1e-13, atol=1e-13 * np.abs(S_exp).max())
    D = np.array([[LAM + 2 * G, LAM, 0], [LAM, LAM + 2 * G, 0], [0, 0, G]])
    np.testing.assert_allclose(DS, np.tile(D.reshape(-1, order="F"), (40, 1)), rtol=1e-15)


def test_apex_closed_form():
    rng = np.random.default_rng(1)
    E = np.tile([4e-3, 4e-3, 0.0], (6, 1)) + rng.normal(0, 1e-8, (6, 3))
    S, DS, br = oracle.mc2d(E, *_mat(6))
    assert np.all(br == 4)
    np.testing.assert_allclose(S, np.tile([CB / (2 * SP), CB / (2 * SP), 0, CB / (2 * SP)], (6, 1)), rtol=1e-15)
    assert np.all(DS == 0)


@pytest.mark.parametrize("sp", [SP, np.sin(np.deg2rad(25.0))])
def test_stress_is_closest_point_projection(sp):
    E = _branch_mix(600, seed=3)
    cb = 2.0 * 6.0 * np.sqrt(1 - sp * sp)
    S, _, br = oracle.mc2d(E, *_mat(E.shape[0], cb, sp))
    assert set(np.unique(br)) == {0, 1, 2, 3, 4}
    S_exp = _expected_stress(E, cb, sp)
    err = np.abs(S - S_exp).max(axis=1) / np.abs(S_exp).max(axis=1).clip(1e-30)
    assert err.max() < 1e-9, (err.max(), br[np.argmax(err)])


def test_returned_stress_on_yield_surface_and_edges():
    E = _branch_mix(800, seed=5)
    S, _, br = oracle.mc2d(E, *_mat(E.shape[0]))
    sig = np.zeros((E.shape[0], 3))
    for p in range(E.shape[0]):
        w = np.linalg.eigvalsh(np.array([[S[p, 0], S[p, 2]], [S[p, 2], S[p, 1]]]))
        sig[p] = np.sort([w[0], w[1], S[p, 3]])[::-1]
    f = (1 + SP) * sig[:, 0] - (1 - SP) * sig[:, 2] - CB
    plastic = (br >= 1) & (br <= 3)
    assert plastic.sum() > 100
    assert np.abs(f[plastic]).max() < 1e-9 * CB
    assert np.all(f[br == 0] <= 1e-9 * CB)
    left, right = br == 2, br == 3
    assert left.sum() > 5 and right.sum() > 5
    assert np.abs(sig[left, 0] - sig[left, 1]).max() < 1e-9 * CB
    assert np.abs(sig[right, 1] - sig[right, 2]).max() < 1e-9 * CB


def test_tangent_matches_central_differences():
    E = _branch_mix(400, seed=7)
    mat = _mat(E.shape[0])
    S0, DS, br = oracle.mc2d(E, *mat)
    h = 1e-10
    fd = np.zeros_like(DS)
    same = np.ones(E.shape[0], dtype=bool)
    for j in range(3):
        Ep, Em = E.copy(), E.copy()
        Ep[:, j] += h
        Em[:, j] -= h
        Sp, _, bp = oracle.mc2d(Ep, *mat)
        Sm, _, bm = oracle.mc2d(Em, *mat)
        fd[:, 3 * j:3 * j + 3] = (Sp[:, :3] - Sm[:, :3]) / (2 * h)
        same &= (bp == br) & (bm == br)
    scale = np.maximum(np.abs(DS).max(axis=1), G)
    err = np.abs(fd - DS).max(axis=1) / scale
    assert same.sum() > 380
    assert err[same].max() < 2e-5, (err[same].max(), br[same][np.argmax(err[same])])
    for b in range(5):
        assert np.sum(same & (br == b)) > 5


def test_tangent_symmetric():
    E = _branch_mix(300, seed=9)
    _, DS, _ = oracle.mc2d(E, *_mat(E.shape[0]))
    D = DS.reshape(-1, 3, 3)
    assert np.abs(D - D.transpose(0, 2, 1)).max() < 1e-9 * np.abs(D).max()


def test_tie_cases_of_the_sort():
    """I2 == 0 (:173-184), eps33 = 0 equal to the larger in-plane eigenvalue (the `>=` of :224),
    eps33 between and above the in-plane eigenvalues (:224-241): stresses equal the projection.
    phi = 15 deg, so that plane-strain compression with eps33 at the tie yields."""
    sp, cb = np.sin(np.deg2rad(15.0)), 12.0 * np.cos(np.deg2rad(15.0))
    cases = np.array([
        [-3e-3, -3e-3, 0.0],      # I2 == 0, elastic (sigma_33 = lambda tr)
        [1e-3, 1e-3, 0.0],        # I2 == 0, plastic (s1 = s2 in plane)
        [5e-3, 5e-3, 0.0],        # I2 == 0, apex
        [0.0, -4e-3, 0.0],        # larger in-plane eigenvalue == eps33 == 0: the tie of the >=
        [0.0, -1.5e-3, 0.0],
        [-4e-3, 0.0, 0.0],
        [1e-3, -3e-3, 0.0],       # eps33 between the in-plane eigenvalues
        [-1e-3, -4e-3, 1e-3],     # eps33 above both
        [2e-3, -2e-3, 0.0],       # I1 == 0
    ])
    n = cases.shape[0]
    S, DS, br = oracle.mc2d(cases, *_mat(n, cb, sp))
    assert np.all(np.isfinite(S)) and np.all(np.isfinite(DS))
    S_exp = _expected_stress(cases, cb, sp)
    np.testing.assert_allclose(S, S_exp, rtol=0, atol=1e-12 * np.abs(S_exp).max())
    assert br[0] == 0 and br[1] != 0 and br[2] == 4
    assert np.all(br[3:6] != 0)          # the ties are plastic, so a wrong sort changes S
    for mutation in ("sort_strict", "no_sort"):
        Sm, _, _ = oracle.mc2d(cases, *_mat(n, cb, sp), _mutation=mutation)
        assert _rel(Sm, S_exp) > 1e-3, mutation


MUTATIONS = ["sort_strict", "no_sort", "denom_lr", "sign_s3", "shear_sign", "no_s33", "mat3_half"]


@pytest.mark.parametrize("mutation", MUTATIONS)
def test_planted_mistakes_are_caught(mutation):
    """Each planted mistake must fail at least one of the checks above (projection over the
    branch mix and the tie cases, the elastic closed form, or DS against central differences)."""
    E = np.concatenate([_branch_mix(300, seed=3), np.random.default_rng(0).normal(0, 1e-5, (10, 3)) - [1e-4, 1e-4, 0]])
    mat = _mat(E.shape[0])
    S, DS, br = oracle.mc2d(E, *mat, _mutation=mutation)
    caught = _rel(S, _expected_stress(E, CB, SP)) > 1e-6
    ties = np.array([[0.0, -4e-3, 0.0], [0.0, -1.5e-3, 0.0], [-4e-3, 0.0, 0.0]])
    sp15, cb15 = np.sin(np.deg2rad(15.0)), 12.0 * np.cos(np.deg2rad(15.0))
    St, _, _ = oracle.mc2d(ties, *_mat(3, cb15, sp15), _mutation=mutation)
    caught |= _rel(St, _expected_stress(ties, cb15, sp15)) > 1e-6
    if not caught:
        h = 1e-10
        S0, DS0, br0 = oracle.mc2d(E, *mat)
        for j in range(3):
            Ep, Em = E.copy(), E.copy()
            Ep[:, j] += h
            Em[:, j] -= h
            Sp, _, bp = oracle.mc2d(Ep, *mat, _mutation=mutation)
            Sm, _, bm = oracle.mc2d(Em, *mat, _mutation=mutation)
            ok = (bp == br) & (bm == br)
            fd = (Sp[:, :3] - Sm[:, :3]) / (2 * h)
            caught |= np.abs(fd - DS[:, 3 * j:3 * j + 3])[ok].max() > 1e-3 * G
    assert caught, mutation


def test_spmv_matches_scipy():
    """oracle.spmv (the C restatement's CSR SpMV) against scipy on a random CSR with empty rows,
    a dense row and unsorted-free random columns."""
    import scipy.sparse as sps
    rng = np.random.default_rng(4)
    n, m = 700, 650
    A = sps.random(n, m, density=0.02, random_state=5, format="lil")
    A[3, :] = 0.0
    A[10, :] = rng.normal(size=m)
    A = A.tocsr()
    A.sort_indices()
    x = rng.normal(size=m)
    y = oracle.spmv(A.indptr, A.indices, A.data, x)
    assert A.getrow(3).nnz == 0 and y[3] == 0.0
    np.testing.assert_allclose(y, A @ x, rtol=1e-13, atol=1e-13 * np.abs(A @ x).max())
