"""CPU oracle for the ssx hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product package never imports it.

Contents
  * ``lib()``      -- ctypes handle on oracle/_build/libssx_oracle.so, this project's C restatement
                      (oracle/ssx_oracle.c) of the 3D return mapping, strain, force and tangent-rows
                      kernels (cython/constitutive_3D_kernel.h, cython/assemble_tangent_values_3d.c).
  * ``ref()``      -- ctypes handle on oracle/_ref/libref_kernels.so, the reference's own C kernels
                      compiled from their sources (oracle/Makefile `ref`); None where not built.
  * NumPy restatements below: Davis reduction, 2D plane-strain return mapping, brute-force strain
    operator B, brute-force CSR pattern, dense element tangent.
  * ``oracle.ssr2d`` -- the plain 2D SSR end to end on the host (vectorised return mapping, sparse
    B^T D B, Newton with damping, the SSR continuation; SuperLU solves): the checker of the 2D SSR
    and bench.py's CPU baseline of that leg.

Pinning summary (what checks each part beyond retyping):
  davis_reduction   closed forms: for psi = 0 Davis B gives tan(phi_l) = sin(phi_1), phi_1 =
                    atan(tan(phi)/lambda), c_bar = 2 cos(phi_1) (c0/lambda) cos(phi_l); at c0 = 6,
                    phi = 45 deg, lambda = 1: c_bar = 4 sqrt(3), sin_phi = 1/sqrt(3).
  mc2d              the closest-point projection of the trial stress onto the six Mohr-Coulomb
                    planes (exact active-set enumeration, no sort and no branch formulas) over
                    all five branches and the tie cases of the eigenvalue sort; elastic (with
                    sigma_33) and apex closed forms; yield surface and edge equalities; DS vs
                    central differences; planted mistakes (``_mutation``) that these checks catch
                    (tests/test_oracle_mc2d.py); end to end, the reference's 2D SSR history.
  spmv              scipy CSR product (tests/test_oracle_mc2d.py).
  strain/force      brute-force B from the reference's strain-operator rule (fem/assembly.py:203-230).
  pattern           set-based restatement of fem/distributed_tangent.py:440-478 and :369-437.
  ssr2d             its return mapping equals mc2d entry for entry; its sparse tangent is the
                    central-difference derivative of its force; its SSR run reproduces the
                    reference's committed 2D lambda history (tests/test_oracle_ssr2d.py).
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_ORACLE_SO = HERE / "_build" / "libssx_oracle.so"
_REF_SO = HERE / "_ref" / "libref_kernels.so"

_P = C.c_void_p
_I64 = C.c_int64
_I = C.c_int

_lib = None
_ref = None


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not _ORACLE_SO.exists():
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(str(_ORACLE_SO))
        L.ssx_oracle_mc3d_point.argtypes = [_P, C.c_double, C.c_double, C.c_double, C.c_double, _P, _P]
        L.ssx_oracle_mc3d_point.restype = C.c_int
        L.ssx_oracle_constitutive_3d.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P]
        L.ssx_oracle_strain_3d.argtypes = [_P, _P, _P, _P, _P, _I64, _I, _I, _P]
        L.ssx_oracle_force_3d_rows.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _I64, _I, _I, _P]
        L.ssx_oracle_tangent_3d_rows.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I, _I, _P]
        L.ssx_oracle_element_matrices_3d.argtypes = [_P, _P, _P, _P, _P, _I64, _I, _I, _P]
        L.ssx_oracle_spmv_csr.argtypes = [_I64, _P, _P, _P, _P, _P]
        _lib = L
    return _lib


def ref():
    """The reference's compiled C kernels, or None where oracle/_ref was not built."""
    global _ref
    if _ref is None and _REF_SO.exists():
        R = C.CDLL(str(_REF_SO))
        R.constitutive_problem_3d_sds_batch_c.argtypes = [_P, _P, _P, _P, _P, _P, _I, _P, _P]
        R.constitutive_problem_3d_s_batch_c.argtypes = [_P, _P, _P, _P, _P, _P, _I, _P]
        R.assemble_overlap_strain_3d_c.argtypes = [_P, _P, _P, _P, _P, _I, _I, _I, _I, _P]
        R.assemble_force_3d_rows_c.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P]
        R.assemble_tangent_values_3d_rows_c.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I,
                                                        _I, _P]
        R.assemble_tangent_values_3d_p2_c.argtypes = [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P]
        _ref = R
    return _ref


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------------------------------------
# C restatement wrappers (point-major layouts)
# ---------------------------------------------------------------------------------------------

def constitutive_3d(E, c_bar, sin_phi, shear, lame, *, tangent=True):
    """E (n, 6) -> S (n, 6), DS (n, 36) or None, branch (n,)."""
    E = _c(E)
    n = E.shape[0]
    S = np.empty((n, 6))
    DS = np.empty((n, 36)) if tangent else None
    br = np.empty(n, dtype=np.int32)
    a = [_c(x) for x in (c_bar, sin_phi, shear, lame)]
    lib().ssx_oracle_constitutive_3d(n, _p(E), *(_p(x) for x in a), _p(S), _p(DS), _p(br))
    return S, DS, br


def strain_3d(dphi, u, edofs, n_q):
    dphi = _c(dphi)
    u = _c(u)
    edofs = _c(edofs, np.int32)
    n_e, nld = edofs.shape
    n_p = nld // 3
    E = np.empty((n_e * n_q, 6))
    lib().ssx_oracle_strain_3d(_p(dphi[0]), _p(dphi[1]), _p(dphi[2]), _p(u), _p(edofs), n_e, n_p, n_q, _p(E))
    return E


def force_3d_rows(dphi, S, w, pat, n_q):
    dphi = _c(dphi)
    S = _c(S)
    w = _c(w)
    n_rows = pat.indptr.size - 1
    F = np.empty(n_rows)
    lib().ssx_oracle_force_3d_rows(_p(dphi[0]), _p(dphi[1]), _p(dphi[2]), _p(S), _p(w), _p(_c(pat.row_slot_ptr, np.int32)),
                                   _p(_c(pat.slot_elem, np.int32)), _p(_c(pat.slot_lrow, np.uint8)), n_rows,
                                   dphi.shape[2], n_q, _p(F))
    F[pat.row_free == 0] = 0.0
    return F


def tangent_3d_rows(dphi, DS, w, pat, n_q):
    dphi = _c(dphi)
    DS = _c(DS)
    w = _c(w)
    n_rows = pat.indptr.size - 1
    out = np.zeros(pat.indices.size)
    lib().ssx_oracle_tangent_3d_rows(_p(dphi[0]), _p(dphi[1]), _p(dphi[2]), _p(DS), _p(w),
                                     _p(_c(pat.row_slot_ptr, np.int32)), _p(_c(pat.slot_elem, np.int32)),
                                     _p(_c(pat.slot_lrow, np.uint8)), _p(_c(pat.slot_pos, np.int32)), n_rows,
                                     dphi.shape[2], n_q, _p(out))
    out[pat.constrained_diag_positions] = 1.0
    return out


def element_matrices_3d(dphi, DS, w, n_elem, n_q):
    dphi = _c(dphi)
    n_p = dphi.shape[2]
    ke = np.empty((n_elem, 3 * n_p, 3 * n_p))
    lib().ssx_oracle_element_matrices_3d(_p(dphi[0]), _p(dphi[1]), _p(dphi[2]), _p(_c(DS)), _p(_c(w)), n_elem, n_p,
                                         n_q, _p(ke))
    return ke


def spmv(indptr, indices, val, x):
    indptr = _c(indptr, np.int32)
    y = np.empty(indptr.size - 1)
    lib().ssx_oracle_spmv_csr(indptr.size - 1, _p(indptr), _p(_c(indices, np.int32)), _p(_c(val)), _p(_c(x)), _p(y))
    return y


# ---------------------------------------------------------------------------------------------
# Reference C (compiled from the reference's sources) wrappers
# ---------------------------------------------------------------------------------------------

def ref_constitutive_3d(E, c_bar, sin_phi, shear, lame, bulk=None, *, tangent=True):
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref not built")
    E = _c(E)
    n = E.shape[0]
    bulk = np.zeros(n) if bulk is None else bulk
    a = [_c(x) for x in (c_bar, sin_phi, shear, bulk, lame)]
    S = np.empty((n, 6))
    if tangent:
        DS = np.empty((n, 36))
        R.constitutive_problem_3d_sds_batch_c(_p(E), *(_p(x) for x in a), n, _p(S), _p(DS))
        return S, DS
    R.constitutive_problem_3d_s_batch_c(_p(E), *(_p(x) for x in a), n, _p(S))
    return S, None


def ref_strain_3d(dphi, u, edofs, n_q):
    R = ref()
    dphi = _c(dphi)
    edofs = _c(edofs, np.int32)
    n_e, nld = edofs.shape
    E = np.empty((n_e * n_q, 6))
    R.assemble_overlap_strain_3d_c(_p(dphi[0]), _p(dphi[1]), _p(dphi[2]), _p(_c(u)), _p(edofs), n_e * n_q, n_e,
                                   nld // 3, n_q, _p(E))
    return E


def ref_force_3d_rows(dphi, S, w, pat, n_q):
    R = ref()
    dphi = _c(dphi)
    n_rows = pat.indptr.size - 1
    F = np.empty(n_rows)
    R.assemble_force_3d_rows_c(_p(dphi[0]), _p(dphi[1]), _p(dphi[2]), _p(_c(S)), _p(_c(w)),
                               _p(_c(pat.row_slot_ptr, np.int32)), _p(_c(pat.slot_elem, np.int32)),
                               _p(_c(pat.slot_lrow, np.uint8)), dphi.shape[1], n_rows, pat.slot_elem.size,
                               dphi.shape[2], n_q, _p(F))
    F[pat.row_free == 0] = 0.0
    return F


def ref_tangent_3d_rows(dphi, DS, w, pat, n_q):
    R = ref()
    dphi = _c(dphi)
    n_rows = pat.indptr.size - 1
    out = np.zeros(pat.indices.size)
    R.assemble_tangent_values_3d_rows_c(_p(dphi[0]), _p(dphi[1]), _p(dphi[2]), _p(_c(DS)), _p(_c(w)),
                                        _p(_c(pat.row_slot_ptr, np.int32)), _p(_c(pat.slot_elem, np.int32)),
                                        _p(_c(pat.slot_lrow, np.uint8)), _p(_c(pat.slot_pos, np.int32)),
                                        dphi.shape[1], n_rows, pat.slot_elem.size, dphi.shape[2], n_q,
                                        pat.indices.size, _p(out))
    out[pat.constrained_diag_positions] = 1.0
    return out


# ---------------------------------------------------------------------------------------------
# NumPy restatements
# ---------------------------------------------------------------------------------------------

def davis_reduction(c0, phi, psi, lam, davis_type="B"):
    """constitutive/reduction.py:14-78."""
    c0, phi, psi = (np.asarray(a, dtype=np.float64) for a in (c0, phi, psi))
    t = davis_type.upper()
    if t == "A":
        beta = np.cos(phi) * np.cos(psi) / (1.0 - np.sin(phi) * np.sin(psi))
        phil = np.arctan(beta * np.tan(phi) / lam)
        return 2.0 * (beta * c0 / lam) * np.cos(phil), np.sin(phil)
    if t == "B":
        phi1 = np.arctan(np.tan(phi) / lam)
        psi1 = np.arctan(np.tan(psi) / lam)
        beta = np.cos(phi1) * np.cos(psi1) / (1.0 - np.sin(phi1) * np.sin(psi1))
        phil = np.arctan(beta * np.tan(phi1))
        return 2.0 * (beta * (c0 / lam)) * np.cos(phil), np.sin(phil)
    if t == "C":
        phi1 = np.arctan(np.tan(phi) / lam)
        beta = np.where(phi1 > psi, np.cos(phi1) * np.cos(psi) / (1.0 - np.sin(phi1) * np.sin(psi)), 1.0)
        phil = np.arctan(beta * np.tan(phi1))
        return 2.0 * (beta * (c0 / lam)) * np.cos(phil), np.sin(phil)
    raise ValueError("Incorrect choice of the Davis approach")


def mc2d(E, c_bar, sin_phi, shear, lame, *, _mutation: str | None = None):
    """Plane-strain return mapping, loop form of constitutive/problem.py:143-475.
    E (n, 3) -> S (n, 4), DS (n, 9) column-major 3x3, branch (n,).  Small inputs only.

    ``_mutation`` plants one named mistake (tests only: it shows that the pinning checks of
    tests/test_oracle_mc2d.py fail on it): "sort_strict" (`>` for `>=` in the sort test of :224),
    "no_sort" (the eps33-between swap of :224-230 dropped), "denom_lr" (denom_l and denom_r of
    :250-253 exchanged), "sign_s3" (sign of the (1 - sin phi) term of sigma_3 on the smooth face,
    :281), "shear_sign" (sign of the in-plane shear projection, :175), "no_s33" (sigma_33 of the
    elastic branch dropped, :270), "mat3_half" (factor 2 of mat3 on the smooth face dropped,
    :289)."""
    E = np.asarray(E, dtype=np.float64)
    n = E.shape[0]
    S = np.zeros((n, 4))
    DS = np.zeros((n, 9))
    br = np.zeros(n, dtype=np.int32)
    v3 = np.array([1.0, 1.0, 0.0])
    for p in range(n):
        ex, ey, g = E[p]
        cb, sp, G, lam = c_bar[p], sin_phi[p], shear[p], lame[p]
        I1 = ex + ey
        I2 = np.sqrt((ex - ey) ** 2 + g * g)
        a1, a2, a3 = 0.5 * (I1 + I2), 0.5 * (I1 - I2), 0.0
        if I2 == 0.0:
            Pa = np.array([1.0, 1.0, 0.0, 0.0])
            Da = np.zeros(9)
        else:
            Pa = np.array([(ex - a2) / I2, (ey - a2) / I2, (g / 2.0) / I2, 0.0])
            if _mutation == "shear_sign":
                Pa[2] = -Pa[2]
        Pb = np.array([1.0, 1.0, 0.0, 0.0]) - Pa
        if I2 != 0.0:
            x, y = Pa[:3], Pb[:3]
            M = np.zeros((3, 3))
            for j in range(3):
                for i in range(3):
                    M[i, j] = (0.5 if (i == j == 2) else (1.0 if i == j else 0.0)) - x[i] * x[j] - y[i] * y[j]
            Da = (M / I2).reshape(-1, order="F")
        Pc = np.array([0.0, 0.0, 0.0, 1.0])
        eigs = [(a1, Pa, Da), (a2, Pb, -Da), (a3, Pc, np.zeros(9))]
        e1, e2, e3 = eigs[0], eigs[1], eigs[2]
        if ((a1 > a3) if _mutation == "sort_strict" else (a1 >= a3)) and (a3 > a2) and _mutation != "no_sort":
            e2, e3 = eigs[2], eigs[1]
        if a3 > a1:
            e1, e2, e3 = eigs[2], eigs[0], eigs[1]
        (l1, P1, D1), (l2, P2, D2), (l3, P3, D3) = e1, e2, e3
        tr = l1 + l2 + l3
        f = 2 * G * ((1 + sp) * l1 - (1 - sp) * l3) + 2 * lam * sp * tr - cb
        gsl, gsr = (l1 - l2) / (1 + sp), (l2 - l3) / (1 - sp)
        gla, gra = (l1 + l2 - 2 * l3) / (3 - sp), (2 * l1 - l2 - l3) / (3 + sp)
        ds_ = 4 * lam * sp ** 2 + 2 * G * (1 + sp) ** 2 + 2 * G * (1 - sp) ** 2
        dl_ = 4 * lam * sp ** 2 + G * (1 + sp) ** 2 + 2 * G * (1 - sp) ** 2
        dr_ = 4 * lam * sp ** 2 + 2 * G * (1 + sp) ** 2 + G * (1 - sp) ** 2
        if _mutation == "denom_lr":
            dl_, dr_ = dr_, dl_
        ls = f / ds_
        ll = (G * ((1 + sp) * (l1 + l2) - 2 * (1 - sp) * l3) + 2 * lam * sp * tr - cb) / dl_
        lr = (G * (2 * (1 + sp) * l1 - (1 - sp) * (l2 + l3)) + 2 * lam * sp * tr - cb) / dr_
        vol = np.outer(v3, v3)
        if f <= 0:
            S[p] = [lam * I1 + 2 * G * ex, lam * I1 + 2 * G * ey, G * g, 0.0 if _mutation == "no_s33" else lam * I1]
            DS[p] = (vol * lam + 2 * np.diag([1.0, 1.0, 0.5]) * G).reshape(-1, order="F")
            br[p] = 0
        elif ls <= min(gsl, gsr):
            s1 = lam * tr + 2 * G * l1 - ls * (2 * lam * sp + 2 * G * (1 + sp))
            s2 = lam * tr + 2 * G * l2 - ls * (2 * lam * sp)
            s3 = lam * tr + 2 * G * l3 - ls * (2 * lam * sp + (2 if _mutation == "sign_s3" else -2) * G * (1 - sp))
            S[p] = s1 * P1 + s2 * P2 + s3 * P3
            q6 = 2 * G * (1 + sp) * P1[:3] - 2 * G * (1 - sp) * P3[:3] + v3 * 2 * lam * sp
            M = (s1 * D1 + s2 * D2 + s3 * D3).reshape(3, 3, order="F") + vol * lam \
                + (1 if _mutation == "mat3_half" else 2) * G * np.outer(P1[:3], P1[:3]) \
                + 2 * G * (np.outer(P2[:3], P2[:3]) + np.outer(P3[:3], P3[:3])) \
                - np.outer(q6, q6) / ds_
            DS[p] = M.reshape(-1, order="F")
            br[p] = 1
        elif gsl < gsr and gsl <= ll <= gla:
            s1 = lam * tr + G * (l1 + l2) - ll * (2 * lam * sp + G * (1 + sp))
            s3 = lam * tr + 2 * G * l3 - ll * (2 * lam * sp - 2 * G * (1 - sp))
            S[p] = s1 * (P1 + P2) + s3 * P3
            P12 = P1[:3] + P2[:3]
            q6 = G * (1 + sp) * P12 - 2 * G * (1 - sp) * P3[:3] + v3 * 2 * lam * sp
            M = (s1 * (D1 + D2) + s3 * D3).reshape(3, 3, order="F") + vol * lam + G * np.outer(P12, P12) \
                + 2 * G * np.outer(P3[:3], P3[:3]) - np.outer(q6, q6) / dl_
            DS[p] = M.reshape(-1, order="F")
            br[p] = 2
        elif gsl > gsr and gsr <= lr <= gra:
            s1 = lam * tr + 2 * G * l1 - lr * (2 * lam * sp + 2 * G * (1 + sp))
            s3 = lam * tr + G * (l2 + l3) - lr * (2 * lam * sp - G * (1 - sp))
            S[p] = s1 * P1 + s3 * (P2 + P3)
            P23 = P2[:3] + P3[:3]
            q6 = 2 * G * (1 + sp) * P1[:3] - G * (1 - sp) * P23 + v3 * 2 * lam * sp
            M = (s1 * D1 + s3 * (D2 + D3)).reshape(3, 3, order="F") + vol * lam + 2 * G * np.outer(P1[:3], P1[:3]) \
                + G * np.outer(P23, P23) - np.outer(q6, q6) / dr_
            DS[p] = M.reshape(-1, order="F")
            br[p] = 3
        else:
            sa = cb / (2 * sp)
            S[p] = [sa, sa, 0.0, sa]
            br[p] = 4
    return S, DS, br


def strain_operator(dphi: np.ndarray, elem: np.ndarray, n_nodes: int, n_q: int):
    """Brute-force sparse B (n_int*n_strain x dim*n_nodes), strain rows point-major
    [e11 e22 e33 g12 g23 g13] (3D) / [e11 e22 g12] (2D), dof = dim*node + comp
    (fem/assembly.py:99-116, :203-230)."""
    import scipy.sparse as sp
    dim, n_int, n_p = dphi.shape
    ns = 6 if dim == 3 else 3
    rows, cols, vals = [], [], []
    for g in range(n_int):
        e = g // n_q
        for j in range(n_p):
            node = int(elem[j, e])
            gx, gy = dphi[0, g, j], dphi[1, g, j]
            if dim == 3:
                gz = dphi[2, g, j]
                ent = [(0, 0, gx), (1, 1, gy), (2, 2, gz), (3, 0, gy), (3, 1, gx), (4, 1, gz), (4, 2, gy),
                       (5, 0, gz), (5, 2, gx)]
            else:
                ent = [(0, 0, gx), (1, 1, gy), (2, 0, gy), (2, 1, gx)]
            for s, c, v in ent:
                rows.append(ns * g + s)
                cols.append(dim * node + c)
                vals.append(v)
    return sp.csr_matrix((vals, (rows, cols)), shape=(ns * n_int, dim * n_nodes))


def brute_pattern(elem: np.ndarray, q_mask: np.ndarray):
    """(indptr, indices) by the reference's set rule (fem/distributed_tangent.py:440-478)."""
    dim, n_nodes = q_mask.shape
    free = q_mask.T.reshape(-1)
    rows = [set() for _ in range(dim * n_nodes)]
    for e in range(elem.shape[1]):
        dofs = [dim * int(v) + c for v in elem[:, e] for c in range(dim)]
        fcols = [d for d in dofs if free[d]]
        for r in dofs:
            if free[r]:
                rows[r].update(fcols)
    for r in range(dim * n_nodes):
        if not free[r]:
            rows[r] = {r}
    indptr = np.zeros(dim * n_nodes + 1, dtype=np.int64)
    chunks = []
    for r, s in enumerate(rows):
        a = np.array(sorted(s), dtype=np.int64)
        chunks.append(a)
        indptr[r + 1] = indptr[r] + a.size
    return indptr, np.concatenate(chunks)


def a_orthogonalize(B: np.ndarray, A) -> tuple[np.ndarray, np.ndarray]:
    """A-orthonormalisation of the rows of B newest-first with the drop rule of
    linear/orthogonalize.py:124-170: v <- v - sum_j s_j (w_j . A v0) w_j with A v0 taken before the
    projection, kept when |v . A v0| > 1e-3, scaled by 1/sqrt|v . A v0|, sign s = sign(v . A v0).
    Returns (W oldest-first, signs).  Pinned by the closed-form checks in its tests (W^T A W =
    diag(s) on a well-conditioned basis; exact dependence is dropped)."""
    k = B.shape[0]
    W, S = [], []
    for i in range(k):
        v = np.array(B[k - 1 - i], dtype=np.float64)
        av = A @ v
        for w, s in zip(W, S):
            v = v - s * float(w @ av) * w
        nv = float(v @ av)
        if abs(nv) > 1e-3:
            W.append(v / np.sqrt(abs(nv)))
            S.append(1.0 if nv > 0 else -1.0)
    if not W:
        return np.zeros((0, B.shape[1])), np.zeros(0)
    return np.array(W[::-1]), np.array(S[::-1])
