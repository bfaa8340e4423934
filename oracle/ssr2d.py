"""CPU restatement of the plain SSR continuation for plane-strain P2 slopes (NumPy + scipy).

TEST INFRASTRUCTURE ONLY (the CPU checker and bench.py's CPU baseline of the end-to-end SSR leg);
the product package never imports it.

What it restates, all on the host:
  * Davis B reduction (constitutive/reduction.py:14-78) -> `oracle.davis_reduction`;
  * the plane-strain return mapping (constitutive/problem.py:143-475), vectorised (`mc2d_vec`,
    checked entry for entry against the loop form `oracle.mc2d`, which is pinned by the
    closest-point projection in tests/test_oracle_mc2d.py);
  * strain E = B u, force F = B^T (w S), tangent K = B^T diag(w DS) B with sparse B
    (fem/assembly.py:50-136), K_r = (1 - r) K + r K_el (constitutive/problem.py:1595-1644);
  * plain Newton with its damping (nonlinear/newton.py:345-626, nonlinear/damping.py:45-134), the
    nested SSR Newton with its damping (nonlinear/newton.py:629-1128, nonlinear/damping.py:137-214),
    the initial phase (continuation/indirect.py:1627-1736) and the SSR continuation with the secant
    predictor and the omega-doubling rule (continuation/indirect.py:550-566, :2267-3557).
Linear systems are solved exactly on the free dofs (SuperLU); the reference solves them with
deflated FGMRES to a relative 1e-1, so Newton counts differ while the FoS agrees to 1e-3.
"""

from __future__ import annotations

import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def mc2d_vec(E, c_bar, sin_phi, shear, lame):
    """Vectorised plane-strain return mapping: E (n, 3) -> S (n, 4), DS (n, 9) column-major 3x3,
    branch (n,) -- the loop form `oracle.mc2d` over arrays."""
    E = np.asarray(E, dtype=np.float64)
    n = E.shape[0]
    ex, ey, g = E[:, 0], E[:, 1], E[:, 2]
    cb, s_, G, L = (np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)) for a in (c_bar, sin_phi, shear, lame))
    I1 = ex + ey
    I2 = np.sqrt((ex - ey) ** 2 + g * g)
    a1, a2 = 0.5 * (I1 + I2), 0.5 * (I1 - I2)
    z = I2 == 0.0
    I2s = np.where(z, 1.0, I2)
    Pa = np.zeros((n, 4))
    Pa[:, 0] = np.where(z, 1.0, (ex - a2) / I2s)
    Pa[:, 1] = np.where(z, 1.0, (ey - a2) / I2s)
    Pa[:, 2] = np.where(z, 0.0, (g / 2.0) / I2s)
    Pb = np.zeros((n, 4))
    Pb[:, 0], Pb[:, 1] = 1.0 - Pa[:, 0], 1.0 - Pa[:, 1]
    Pb[:, 2] = -Pa[:, 2]
    Pc = np.zeros((n, 4))
    Pc[:, 3] = 1.0
    # derivative of the first eigenprojection: (diag(1, 1, 1/2) - x x^T - y y^T) / I2
    x, y = Pa[:, :3], Pb[:, :3]
    M = np.zeros((n, 3, 3))
    M[:, 0, 0] = M[:, 1, 1] = 1.0
    M[:, 2, 2] = 0.5
    M -= x[:, :, None] * x[:, None, :] + y[:, :, None] * y[:, None, :]
    Da = np.where(z[:, None, None], 0.0, M / I2s[:, None, None])
    Dz = np.zeros_like(Da)
    t2 = (a1 >= 0.0) & (0.0 > a2)
    t3 = 0.0 > a1
    # ordered eigenvalues/projections: default (1, 2, 3); t2: (1, 3, 2); t3: (3, 1, 2)
    l1 = np.where(t3, 0.0, a1)
    l2 = np.where(t3, a1, np.where(t2, 0.0, a2))
    l3 = np.where(t3 | t2, a2, 0.0)
    P1 = np.where(t3[:, None], Pc, Pa)
    P2 = np.where(t3[:, None], Pa, np.where(t2[:, None], Pc, Pb))
    P3 = np.where((t3 | t2)[:, None], Pb, Pc)
    D1 = np.where(t3[:, None, None], Dz, Da)
    D2 = np.where(t3[:, None, None], Da, np.where(t2[:, None, None], Dz, -Da))
    D3 = np.where((t3 | t2)[:, None, None], -Da, Dz)
    sp_, lam = s_, L
    tr = l1 + l2 + l3
    f = 2 * G * ((1 + sp_) * l1 - (1 - sp_) * l3) + 2 * lam * sp_ * tr - cb
    gsl, gsr = (l1 - l2) / (1 + sp_), (l2 - l3) / (1 - sp_)
    gla, gra = (l1 + l2 - 2 * l3) / (3 - sp_), (2 * l1 - l2 - l3) / (3 + sp_)
    ds_ = 4 * lam * sp_ ** 2 + 2 * G * (1 + sp_) ** 2 + 2 * G * (1 - sp_) ** 2
    dl_ = 4 * lam * sp_ ** 2 + G * (1 + sp_) ** 2 + 2 * G * (1 - sp_) ** 2
    dr_ = 4 * lam * sp_ ** 2 + 2 * G * (1 + sp_) ** 2 + G * (1 - sp_) ** 2
    ls = f / ds_
    ll = (G * ((1 + sp_) * (l1 + l2) - 2 * (1 - sp_) * l3) + 2 * lam * sp_ * tr - cb) / dl_
    lr = (G * (2 * (1 + sp_) * l1 - (1 - sp_) * (l2 + l3)) + 2 * lam * sp_ * tr - cb) / dr_
    el = f <= 0
    sm = ~el & (ls <= np.minimum(gsl, gsr))
    le = ~el & ~sm & (gsl < gsr) & (ll >= gsl) & (ll <= gla)
    ri = ~el & ~sm & (gsl > gsr) & (lr >= gsr) & (lr <= gra)
    ap = ~(el | sm | le | ri)
    S = np.zeros((n, 4))
    DS = np.zeros((n, 3, 3))
    v3 = np.array([1.0, 1.0, 0.0])
    vol = np.outer(v3, v3)
    outer = lambda a, b: a[:, :, None] * b[:, None, :]  # noqa: E731
    # elastic
    S[el] = np.stack([lam * I1 + 2 * G * ex, lam * I1 + 2 * G * ey, G * g, lam * I1], axis=1)[el]
    Del = vol[None] * lam[:, None, None] + 2 * G[:, None, None] * np.diag([1.0, 1.0, 0.5])[None]
    DS[el] = Del[el]
    # smooth face
    s1 = lam * tr + 2 * G * l1 - ls * (2 * lam * sp_ + 2 * G * (1 + sp_))
    s2 = lam * tr + 2 * G * l2 - ls * (2 * lam * sp_)
    s3 = lam * tr + 2 * G * l3 - ls * (2 * lam * sp_ - 2 * G * (1 - sp_))
    Ss = s1[:, None] * P1 + s2[:, None] * P2 + s3[:, None] * P3
    q6 = 2 * G[:, None] * (1 + sp_)[:, None] * P1[:, :3] - 2 * G[:, None] * (1 - sp_)[:, None] * P3[:, :3] \
        + v3[None] * (2 * lam * sp_)[:, None]
    Ms = s1[:, None, None] * D1 + s2[:, None, None] * D2 + s3[:, None, None] * D3 + vol[None] * lam[:, None, None] \
        + 2 * G[:, None, None] * (outer(P1[:, :3], P1[:, :3]) + outer(P2[:, :3], P2[:, :3]) + outer(P3[:, :3], P3[:, :3])) \
        - outer(q6, q6) / ds_[:, None, None]
    S[sm], DS[sm] = Ss[sm], Ms[sm]
    # left edge
    s1l = lam * tr + G * (l1 + l2) - ll * (2 * lam * sp_ + G * (1 + sp_))
    s3l = lam * tr + 2 * G * l3 - ll * (2 * lam * sp_ - 2 * G * (1 - sp_))
    P12 = P1[:, :3] + P2[:, :3]
    Sl = s1l[:, None] * (P1 + P2) + s3l[:, None] * P3
    q6l = G[:, None] * (1 + sp_)[:, None] * P12 - 2 * G[:, None] * (1 - sp_)[:, None] * P3[:, :3] \
        + v3[None] * (2 * lam * sp_)[:, None]
    Ml = s1l[:, None, None] * (D1 + D2) + s3l[:, None, None] * D3 + vol[None] * lam[:, None, None] \
        + G[:, None, None] * outer(P12, P12) + 2 * G[:, None, None] * outer(P3[:, :3], P3[:, :3]) \
        - outer(q6l, q6l) / dl_[:, None, None]
    S[le], DS[le] = Sl[le], Ml[le]
    # right edge
    s1r = lam * tr + 2 * G * l1 - lr * (2 * lam * sp_ + 2 * G * (1 + sp_))
    s3r = lam * tr + G * (l2 + l3) - lr * (2 * lam * sp_ - G * (1 - sp_))
    P23 = P2[:, :3] + P3[:, :3]
    Sr = s1r[:, None] * P1 + s3r[:, None] * (P2 + P3)
    q6r = 2 * G[:, None] * (1 + sp_)[:, None] * P1[:, :3] - G[:, None] * (1 - sp_)[:, None] * P23 \
        + v3[None] * (2 * lam * sp_)[:, None]
    Mr = s1r[:, None, None] * D1 + s3r[:, None, None] * (D2 + D3) + vol[None] * lam[:, None, None] \
        + 2 * G[:, None, None] * outer(P1[:, :3], P1[:, :3]) + G[:, None, None] * outer(P23, P23) \
        - outer(q6r, q6r) / dr_[:, None, None]
    S[ri], DS[ri] = Sr[ri], Mr[ri]
    # apex
    sa = cb / (2 * sp_)
    S[ap] = np.stack([sa, sa, np.zeros(n), sa], axis=1)[ap]
    br = np.select([el, sm, le, ri], [0, 1, 2, 3], default=4).astype(np.int32)
    # (n, 3, 3) row i, column j -> column-major 9 (index i + 3 j)
    return S, DS.transpose(0, 2, 1).reshape(n, 9), br


class Problem2D:
    """Sparse operators of one plane-strain P2 slope on the host."""

    def __init__(self, mesh, geom, material):
        import oracle
        self.dim = 2
        self.n_nodes = mesh.n_nodes
        self.n_int = geom.n_int
        self.B = oracle.strain_operator(geom.dphi, mesh.elem, mesh.n_nodes, geom.n_q).tocsr()
        self.w = np.asarray(geom.weight, dtype=np.float64)
        self.free = np.flatnonzero(mesh.q_mask.T.reshape(-1))
        m = material
        self.c0, self.phi, self.psi = m.c0, np.deg2rad(m.phi), np.deg2rad(m.psi)
        self.G = m.young / (2 * (1 + m.poisson))
        self.lame = m.young * m.poisson / ((1 + m.poisson) * (1 - 2 * m.poisson))
        n = self.n_int
        rows = np.repeat(np.arange(3 * n).reshape(n, 3), 3, axis=1).reshape(-1)
        cols = np.tile(np.arange(3 * n).reshape(n, 3), (1, 3)).reshape(-1)
        self._blk = (rows, cols)
        Del = np.zeros((n, 9))
        Del[:, [0, 4]] = self.lame + 2 * self.G
        Del[:, [1, 3]] = self.lame
        Del[:, 8] = self.G
        self.K_el = self._tangent(Del)
        self.lam = None

    def _tangent(self, DS):
        # DS column-major per point -> row-major block entries (i, j) = DS[i + 3 j]
        D = DS.reshape(-1, 3, 3).transpose(0, 2, 1) * self.w[:, None, None]
        Dt = sp.csr_matrix((D.reshape(-1), self._blk), shape=(3 * self.n_int, 3 * self.n_int))
        return (self.B.T @ Dt @ self.B).tocsr()

    def reduction(self, lam):
        import oracle
        if lam <= 0:
            raise ValueError("Reduction parameter lambda must be positive")
        self.lam = lam
        cb, sp_ = oracle.davis_reduction(self.c0, self.phi, self.psi, lam, "B")
        self.cb, self.sp = float(cb), float(sp_)

    def stress(self, U, tangent=False):
        E = (self.B @ U).reshape(-1, 3)
        S, DS, _ = mc2d_vec(E, self.cb, self.sp, self.G, self.lame)
        F = self.B.T @ (self.w[:, None] * S[:, :3]).reshape(-1)
        F[np.setdiff1d(np.arange(F.size), self.free, assume_unique=True)] = 0.0
        return (F, DS) if tangent else F

    def F_all(self, lam, U):
        self.reduction(lam)
        return self.stress(U)

    def K_r(self, DS, r):
        return ((1.0 - r) * self._tangent(DS) + r * self.K_el).tocsr()

    def solve(self, K, rhs_list):
        fr = self.free
        lu = spla.splu(K[fr][:, fr].tocsc())
        out = []
        for b in rhs_list:
            x = np.zeros(K.shape[0])
            x[fr] = lu.solve(b[fr])
            out.append(x)
        return out


def _damping(it_damp_max, U, dU, F, f, P):
    res = F - f
    dec0 = float(res @ dU)
    if not np.isfinite(dec0) or dec0 >= 0.0:
        return 0.0
    alpha, a_min, a_max = 1.0, 0.0, 1.0
    for _ in range(int(it_damp_max)):
        Fa = P.stress(U + alpha * dU)
        dec = float((Fa - f) @ dU)
        if dec < 0.0:
            if alpha == 1.0:
                break
            a_min = alpha
        else:
            a_max = alpha
        alpha = 0.5 * (a_min + a_max)
    return float(alpha)


def _damping_alg5(it_damp_max, U, lam, dU, dl, f, crit, P):
    if not np.isfinite(dl) or not np.isfinite(crit) or it_damp_max <= 0:
        return 0.0
    alpha, last = 1.0, None
    for _ in range(int(it_damp_max)):
        la = lam + alpha * dl
        if la <= 0.0:
            alpha *= 0.5
            continue
        c = float(np.linalg.norm(P.F_all(la, U + alpha * dU) - f))
        last = float(alpha)
        if c < crit:
            break
        alpha *= 0.5
    return 0.0 if last is None else float(alpha if alpha == last else last)


def newton(U, tol, it_max, it_damp_max, r_min, f, P, stats):
    U = U.copy()
    nf = float(np.linalg.norm(f)) or 1.0
    r, it, flag, diffs = float(r_min), 0, 0, True
    while True:
        it += 1
        if diffs:
            F, DS = P.stress(U, tangent=True)
        else:
            F = P.stress(U)
        crit = float(np.linalg.norm(F - f)) / nf
        if diffs and crit < tol:
            break
        (dU,) = P.solve(P.K_r(DS, r), [f - F])
        alpha = _damping(it_damp_max, U, dU, F, f, P)
        diffs = True
        if alpha < 1e-1:
            if alpha == 0.0:
                diffs = False
                r *= 2.0
            else:
                r *= 2.0 ** 0.25
        elif alpha > 0.5:
            r = max(r / np.sqrt(2.0), r_min)
        stats["newton"] += 1
        if alpha == 0.0 and r > 1.0:
            flag = 1
            break
        U += alpha * dU
        if not np.isfinite(crit) or it == it_max:
            flag = 1
            break
    return U, flag


def newton_ind_ssr(U, omega, lam, it_max, it_damp_max, tol, r_min, f, P, stats):
    U = U.copy()
    nf = float(np.linalg.norm(f)) or 1.0
    eps = tol / 1000.0
    r, it, flag, diffs, rel = float(r_min), 0, 0, True, np.nan
    while True:
        it += 1
        P.reduction(lam)
        if diffs:
            F, DS = P.stress(U, tangent=True)
        else:
            F = P.stress(U)
        crit = float(np.linalg.norm(F - f))
        rel = crit / nf
        if diffs and rel < tol and it > 1:
            break
        G = (P.F_all(lam + eps, U) - F) / eps
        dW, dV = P.solve(P.K_r(DS, r), [-G, f - F])
        den = float(f @ dW)
        dl = 0.0 if abs(den) < 1e-30 else -float(f @ dV) / den
        dU = dV + dl * dW
        alpha = _damping_alg5(it_damp_max, U, lam, dU, dl, f, crit, P)
        diffs = True
        if alpha < 1e-1:
            if alpha == 0.0:
                diffs = False
                r *= 2.0
            else:
                r *= 2.0 ** 0.25
        elif alpha > 0.5:
            r = max(r / np.sqrt(2.0), r_min)
        stats["newton"] += 1
        if alpha == 0.0 and r > 1.0:
            flag = 1 if rel > 10.0 * tol else 0
            break
        U += alpha * dU
        d = float(f @ U)
        if d != 0.0:
            U *= omega / d
        lam = lam + alpha * dl
        if not np.isfinite(rel) or it == it_max:
            flag = 1 if rel > 10.0 * tol else 0
            break
    return U, lam, flag


def ssr(P, f, *, lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max, omega_max_stop,
        it_newt_max, it_damp_max=10, tol=1e-4, r_min=1e-4):
    """Plain SSR to the factor of safety.  Returns dict(fos, lambda_hist, omega_hist, newton, wall)."""
    t0 = time.perf_counter()
    stats = {"newton": 0}
    dl = float(d_lambda_init)
    lam1 = float(lambda_init)
    while True:
        P.reduction(lam1)
        U1, flag = newton(np.zeros_like(f), tol, it_newt_max, it_damp_max, r_min, f, P, stats)
        if flag == 0:
            break
        lam1 *= 0.5
        dl *= 0.5
        if dl < d_lambda_min:
            raise RuntimeError("Initial choice of lambda seems to be too large.")
    om1 = float(f @ U1)
    while True:
        lam_it = lam1 + dl
        P.reduction(lam_it)
        U_it, flag = newton(U1, tol, it_newt_max, it_damp_max, r_min, f, P, stats)
        if flag == 1:
            dl /= 2.0
        else:
            U2, om2, lam2 = U_it, float(f @ U_it), lam_it
            if (om2 - om1) / max(1.0, om1) < 1e-5:
                U1, lam1, om1 = U2, lam2, om2
            else:
                break
        if dl < d_lambda_min:
            raise RuntimeError("It seems that the FoS is equal to lambda_init.")
        if lam1 > 10.0:
            raise RuntimeError("It seems that the FoS is greater than 10.")
    U_old, U, omega_old, omega, lam_old, lam = U1, U2, om1, om2, lam1, lam2
    lh, oh = [lam_old, lam], [omega_old, omega]
    d_omega = omega - omega_old
    if omega_max_stop < omega + d_omega:
        raise ValueError("Too small value of omega_max_stop. Increase it and rerun.")
    step, n_omega, stop = 2, 0, "completed"
    while True:
        omega_it = min(omega + d_omega, omega_max_stop)
        d_omega = omega_it - omega
        a = 0.0 if abs(omega - omega_old) <= 1e-30 else (omega_it - omega) / (omega - omega_old)
        U_it, lam_c, flag = newton_ind_ssr(U + a * (U - U_old), omega_it, lam, it_newt_max, it_damp_max, tol, r_min,
                                           f, P, stats)
        if flag != 0:
            d_omega /= 2.0
            n_omega += 1
        else:
            step += 1
            U_old, U = U, U_it
            omega_old, omega = omega, omega_it
            d_lam = lam_c - lam
            lam_old, lam = lam, lam_c
            dls = np.nan if d_omega == 0.0 else (d_lam / d_omega) * (omega - oh[0])
            n_omega = 0
            lh.append(lam)
            oh.append(omega)
            if (lh[-1] - lh[-2]) < 0.9 * (lh[-2] - lh[-3]):
                d_omega *= 2.0
            if d_lambda_diff_scaled_min > 0.0 and dls < d_lambda_diff_scaled_min:
                stop = "d_lambda_diff_scaled_min"
                break
            if omega >= omega_max_stop:
                stop = "omega_max_stop"
                break
        if n_omega >= 5:
            stop = "omega_reduction_limit"
            break
        if step >= step_max:
            stop = "step_max"
            break
    return {"fos": float(lh[-1]), "lambda_hist": np.array(lh), "omega_hist": np.array(oh), "newton": stats["newton"],
            "wall": time.perf_counter() - t0, "stop": stop, "U": U}
