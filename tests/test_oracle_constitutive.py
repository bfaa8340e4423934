"""Pin the CPU oracle of the return mapping: against the reference's own C (golden vectors and,
where built, oracle/_ref), closed forms, yield-surface consistency, finite differences,
symmetry and objectivity.  Each check is chosen so that a dropped term, a wrong sign or a
transposed DS index fails it."""

from pathlib import Path

import numpy as np
import pytest

import oracle
from slope_stability_petsc4py_amd import workloads as W

GOLD = Path(__file__).resolve().parent / "golden"
CB, SP = 4.0 * np.sqrt(3.0), 1.0 / np.sqrt(3.0)
G, L = W.HOMO.shear, W.HOMO.lame


def _mat(n):
    return np.full(n, CB), np.full(n, SP), np.full(n, G), np.full(n, L)


def test_oracle_matches_reference_golden():
    d = np.load(GOLD / "constitutive_3d.npz")
    S, DS, _ = oracle.constitutive_3d(d["E"], d["c_bar"], d["sin_phi"], d["shear"], d["lame"])
    np.testing.assert_allclose(S, d["S"], rtol=0, atol=1e-12 * np.abs(d["S"]).max())
    np.testing.assert_allclose(DS, d["DS"], rtol=0, atol=1e-12 * np.abs(d["DS"]).max())


@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built here")
def test_oracle_matches_compiled_reference_stratified():
    E, _ = W.stratified_strains(4000, seed=3)
    S, DS, _ = oracle.constitutive_3d(E, *_mat(E.shape[0]))
    Sr, DSr = oracle.ref_constitutive_3d(E, *_mat(E.shape[0]))
    assert np.array_equal(S, Sr)
    np.testing.assert_allclose(DS, DSr, rtol=0, atol=1e-13 * np.abs(DSr).max())


def test_all_branches_present_and_classified():
    E, b = W.stratified_strains(2000, seed=5)
    _, _, br = oracle.constitutive_3d(E, *_mat(E.shape[0]))
    assert set(np.unique(br)) == {0, 1, 2, 3, 4}
    assert np.array_equal(br, b)


def test_elastic_closed_form():
    rng = np.random.default_rng(0)
    E = rng.normal(0, 1e-5, (50, 6))
    E[:, :3] -= 2e-3  # compression: elastic
    S, DS, br = oracle.constitutive_3d(E, *_mat(50))
    assert np.all(br == 0)
    tr = E[:, :3].sum(1)
    S_exp = np.concatenate([L * tr[:, None] + 2 * G * E[:, :3], G * E[:, 3:]], axis=1)
    np.testing.assert_allclose(S, S_exp, rtol=1e-13, atol=1e-12)
    D = np.zeros((6, 6))
    D[:3, :3] = L
    D += np.diag([2 * G] * 3 + [G] * 3)
    np.testing.assert_allclose(DS, np.tile(D.reshape(-1, order="F"), (50, 1)), rtol=1e-14)


def test_apex_closed_form():
    E = np.tile([3e-3, 3e-3, 3e-3, 0, 0, 0], (5, 1)) + np.random.default_rng(1).normal(0, 1e-9, (5, 6))
    S, DS, br = oracle.constitutive_3d(E, *_mat(5))
    assert np.all(br == 4)
    np.testing.assert_allclose(S, np.tile([CB / (2 * SP)] * 3 + [0] * 3, (5, 1)), rtol=1e-15)
    assert np.all(DS == 0)


def _principal_stress(S):
    T = np.zeros((S.shape[0], 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = S[:, 0], S[:, 1], S[:, 2]
    T[:, 0, 1] = T[:, 1, 0] = S[:, 3]
    T[:, 1, 2] = T[:, 2, 1] = S[:, 4]
    T[:, 0, 2] = T[:, 2, 0] = S[:, 5]
    return np.linalg.eigvalsh(T)[:, ::-1]


def test_returned_stress_on_yield_surface():
    """Plastic branches return onto (1+sin phi) s1 - (1-sin phi) s3 = c_bar (smooth face and the
    edges), with s1 = s2 on the left edge and s2 = s3 on the right edge."""
    E, b = W.stratified_strains(3000, seed=11)
    S, _, br = oracle.constitutive_3d(E, *_mat(E.shape[0]))
    sig = _principal_stress(S)
    plastic = (br >= 1) & (br <= 3)
    f = (1 + SP) * sig[:, 0] - (1 - SP) * sig[:, 2] - CB
    assert np.abs(f[plastic]).max() < 1e-9 * CB * 1e3
    left, right = br == 2, br == 3
    assert np.abs(sig[left, 0] - sig[left, 1]).max() < 1e-8
    assert np.abs(sig[right, 1] - sig[right, 2]).max() < 1e-8


def test_tangent_matches_central_differences():
    E, b = W.stratified_strains(400, seed=13, mix=(0.2, 0.3, 0.25, 0.25, 0.0))
    S0, DS, br = oracle.constitutive_3d(E, *_mat(E.shape[0]))
    h = 1e-9
    fd = np.zeros_like(DS)
    for j in range(6):
        Ep, Em = E.copy(), E.copy()
        Ep[:, j] += h
        Em[:, j] -= h
        Sp, _, bp = oracle.constitutive_3d(Ep, *_mat(E.shape[0]), tangent=False)
        Sm, _, bm = oracle.constitutive_3d(Em, *_mat(E.shape[0]), tangent=False)
        fd[:, 6 * j:6 * j + 6] = (Sp - Sm) / (2 * h)
    scale = np.abs(DS).max(axis=1, keepdims=True)
    err = np.abs(fd - DS).max(axis=1) / scale[:, 0]
    assert np.median(err) < 1e-5
    assert np.mean(err < 1e-3) > 0.97


def test_tangent_symmetric_associated_flow():
    E, _ = W.stratified_strains(1000, seed=17)
    _, DS, _ = oracle.constitutive_3d(E, *_mat(E.shape[0]))
    D = DS.reshape(-1, 6, 6)
    # engineering-shear convention: DS_ij with the shear halving makes W D W^T symmetric, W = diag(1,1,1,1,1,1)
    asym = np.abs(D - D.transpose(0, 2, 1)).max(axis=(1, 2)) / np.maximum(np.abs(D).max(axis=(1, 2)), 1e-30)
    assert asym.max() < 1e-6


def test_objectivity():
    """S(Q e Q^T) = Q S(e) Q^T for a rotation Q."""
    E, _ = W.stratified_strains(300, seed=19)
    rng = np.random.default_rng(2)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))

    def tens(v, eng):
        f = 0.5 if eng else 1.0
        return np.array([[v[0], f * v[3], f * v[5]], [f * v[3], v[1], f * v[4]], [f * v[5], f * v[4], v[2]]])

    Er = np.array([[t[0, 0], t[1, 1], t[2, 2], 2 * t[0, 1], 2 * t[1, 2], 2 * t[0, 2]]
                   for t in (Q @ tens(e, True) @ Q.T for e in E)])
    S, _, _ = oracle.constitutive_3d(E, *_mat(E.shape[0]), tangent=False)
    Sr, _, _ = oracle.constitutive_3d(Er, *_mat(E.shape[0]), tangent=False)
    Sexp = np.array([[t[0, 0], t[1, 1], t[2, 2], t[0, 1], t[1, 2], t[0, 2]]
                     for t in (Q @ tens(s, False) @ Q.T for s in S)])
    np.testing.assert_allclose(Sr, Sexp, rtol=0, atol=1e-8 * np.abs(S).max())


def test_davis_closed_forms():
    cb, sp = oracle.davis_reduction(np.array([6.0]), np.array([np.pi / 4]), np.array([0.0]), 1.0, "B")
    np.testing.assert_allclose(cb, 4 * np.sqrt(3), rtol=1e-14)
    np.testing.assert_allclose(sp, 1 / np.sqrt(3), rtol=1e-14)
    # psi = 0, general lambda: tan(phi_l) = sin(phi_1), c_bar = 2 cos(phi_1) c0/lambda cos(phi_l)
    lam = 1.37
    phi = np.deg2rad(np.array([25.0, 35.0, 40.0]))
    c0 = np.array([5.0, 10.0, 20.0])
    cb, sp = oracle.davis_reduction(c0, phi, np.zeros(3), lam, "B")
    p1 = np.arctan(np.tan(phi) / lam)
    pl = np.arctan(np.sin(p1))
    np.testing.assert_allclose(sp, np.sin(pl), rtol=1e-14)
    np.testing.assert_allclose(cb, 2 * np.cos(p1) * c0 / lam * np.cos(pl), rtol=1e-13)
    # Davis A with psi = 0 and lambda = 1 reduces to c_bar = 2 c0 cos^2(phi)... (beta = cos phi)
    cbA, spA = oracle.davis_reduction(c0, phi, np.zeros(3), 1.0, "A")
    plA = np.arctan(np.cos(phi) * np.tan(phi))
    np.testing.assert_allclose(spA, np.sin(plA), rtol=1e-14)
    np.testing.assert_allclose(cbA, 2 * np.cos(phi) * c0 * np.cos(plA), rtol=1e-13)
    # Davis C with psi >= phi_1 has beta = 1
    cbC, spC = oracle.davis_reduction(c0, phi, np.full(3, 1.2), 2.0, "C")
    p1 = np.arctan(np.tan(phi) / 2.0)
    np.testing.assert_allclose(spC, np.sin(p1), rtol=1e-14)
    np.testing.assert_allclose(cbC, 2 * c0 / 2.0 * np.cos(p1), rtol=1e-13)
