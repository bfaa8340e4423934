"""Synthetic inputs for the kernel benchmark and the parity tests: branch-stratified strain states
and the problem definitions of the benchmark configurations.

Strains are built from principal values and a random rotation, so every return-mapping branch
(elastic, smooth face, left edge, right edge, apex) is present in chosen proportions; a vectorised
branch classifier (the trial tests of cython/constitutive_3D_kernel.h:114-149,211-217,275-281) keeps
only the samples that land where intended.  This is input generation, not part of the solver.
"""

from __future__ import annotations

import numpy as np

from .materials import MaterialSpec

# Homogeneous material of the 2D/3D homogeneous slopes (c=6 kPa, phi=45, psi=0, E=40 MPa, nu=0.3).
HOMO = MaterialSpec(c0=6.0, phi=45.0, psi=0.0, young=40000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0)

# Four horizontal layers of the heterogeneous 3D slope (top to bottom).
HETERO_LAYERS = [
    MaterialSpec(c0=15.0, phi=30.0, psi=0.0, young=50000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0),
    MaterialSpec(c0=10.0, phi=38.0, psi=0.0, young=30000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0),
    MaterialSpec(c0=18.0, phi=35.0, psi=0.0, young=60000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0),
    MaterialSpec(c0=12.0, phi=32.0, psi=0.0, young=40000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0),
]


def _principal_eigs(E: np.ndarray) -> np.ndarray:
    n = E.shape[0]
    T = np.empty((n, 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = E[:, 0], E[:, 1], E[:, 2]
    T[:, 0, 1] = T[:, 1, 0] = 0.5 * E[:, 3]
    T[:, 1, 2] = T[:, 2, 1] = 0.5 * E[:, 4]
    T[:, 0, 2] = T[:, 2, 0] = 0.5 * E[:, 5]
    return np.linalg.eigvalsh(T)[:, ::-1]


def classify_3d(E, c_bar, sin_phi, shear, lame) -> np.ndarray:
    """Branch id per point (0 elastic, 1 smooth, 2 left edge, 3 right edge, 4 apex)."""
    ev = _principal_eigs(np.asarray(E, dtype=np.float64))
    e1, e2, e3 = ev[:, 0], ev[:, 1], ev[:, 2]
    sp, G, lam, cb = sin_phi, shear, lame, c_bar
    I1 = e1 + e2 + e3
    f = 2 * G * ((1 + sp) * e1 - (1 - sp) * e3) + 2 * lam * sp * I1 - cb
    gsl = (e1 - e2) / (1 + sp)
    gsr = (e2 - e3) / (1 - sp)
    ls = f / (4 * lam * sp * sp + 4 * G * (1 + sp * sp))
    den_l = 4 * lam * sp * sp + G * (1 + sp) ** 2 + 2 * G * (1 - sp) ** 2
    ll = (G * ((1 + sp) * (e1 + e2) - 2 * (1 - sp) * e3) + 2 * lam * sp * I1 - cb) / den_l
    gla = (e1 + e2 - 2 * e3) / (3 - sp)
    den_r = 4 * lam * sp * sp + 2 * G * (1 + sp) ** 2 + G * (1 - sp) ** 2
    lr = (G * (2 * (1 + sp) * e1 - (1 - sp) * (e2 + e3)) + 2 * lam * sp * I1 - cb) / den_r
    gra = (2 * e1 - e2 - e3) / (3 + sp)
    br = np.full(e1.shape, 4, dtype=np.int32)
    br[(gsl > gsr) & (lr >= gsr) & (lr <= gra)] = 3
    br[(gsl < gsr) & (ll >= gsl) & (ll <= gla)] = 2
    br[(ls <= gsl) & (ls <= gsr)] = 1
    br[f <= 0] = 0
    return br


def _rotations(rng, n):
    q, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    return q


def _from_principal(rng, ev: np.ndarray) -> np.ndarray:
    Q = _rotations(rng, ev.shape[0])
    T = np.einsum("nij,nj,nkj->nik", Q, ev, Q)
    return np.stack([T[:, 0, 0], T[:, 1, 1], T[:, 2, 2], 2 * T[:, 0, 1], 2 * T[:, 1, 2], 2 * T[:, 0, 2]], axis=1)


def stratified_strains(n: int, seed: int = 0, c_bar: float | None = None, sin_phi: float | None = None,
                       shear: float | None = None, lame: float | None = None,
                       mix=(0.70, 0.12, 0.06, 0.06, 0.06)) -> tuple[np.ndarray, np.ndarray]:
    """(E (n, 6), branch (n,)) with the requested branch fractions (elastic, smooth, left, right,
    apex) for one material (default: the homogeneous slope material, Davis B at lambda = 1)."""
    rng = np.random.default_rng(seed)
    if c_bar is None:
        c_bar, sin_phi = 4.0 * np.sqrt(3.0), 1.0 / np.sqrt(3.0)
        shear, lame = HOMO.shear, HOMO.lame
    counts = np.floor(np.asarray(mix) * n).astype(int)
    counts[0] += n - counts.sum()
    out_E, out_b = [], []
    for kind, k in enumerate(counts):
        got = []
        have = 0
        while have < k:
            m = max(2 * (k - have), 64)
            a = rng.uniform(0.5e-3, 3e-3, m)
            noise = rng.normal(0.0, 1.0, (m, 3))
            if kind == 0:
                ev = -1e-3 * rng.uniform(0.5, 1.5, m)[:, None] + 1e-5 * noise
            elif kind == 1:
                ev = np.stack([a, 0.2 * a, -a], 1) + 1e-5 * noise
            elif kind == 2:
                ev = np.stack([a, a - 2e-6 * np.abs(noise[:, 0]), -a], 1)
            elif kind == 3:
                ev = np.stack([a, -a + 2e-6 * np.abs(noise[:, 0]), -a], 1)
            else:
                ev = a[:, None] * np.ones((m, 3)) + 1e-8 * noise
            ev = -np.sort(-ev, axis=1)
            E = _from_principal(rng, ev)
            br = classify_3d(E, c_bar, sin_phi, shear, lame)
            sel = E[br == kind][: k - have]
            got.append(sel)
            have += sel.shape[0]
        out_E.append(np.concatenate(got) if got else np.empty((0, 6)))
        out_b.append(np.full(k, kind, dtype=np.int32))
    E = np.concatenate(out_E)
    b = np.concatenate(out_b)
    perm = rng.permutation(n)
    return np.ascontiguousarray(E[perm]), b[perm]
