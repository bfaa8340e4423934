"""Newton solvers on the device: plain Newton for F(U) = f, the nested SSR Newton for
F_lambda(U) = f with f^T U = omega, the nested LL Newton, and their damping line searches.

Control flow follows the reference line by line in meaning (nonlinear/newton.py:345-626 `newton`,
:629-1128 `newton_ind_ssr`, :1131-1487 `newton_ind_ll`; nonlinear/damping.py:45-134 `damping`,
:137-214 `damping_alg5`): regularisation r starts at r_min, doubles when the line search returns
alpha = 0 (then F is recomputed without a new tangent), grows by 2^(1/4) for alpha < 0.1, shrinks
by sqrt(2) (not below r_min) for alpha > 0.5; the deflation basis grows with accepted directions;
the FD derivative G = dF/dlambda uses eps = tol/1000; iterates are rescaled to f^T U = omega.

Vectors are flat device tensors over all dofs (dof = dim*node + comp) holding zeros at
constrained dofs, so norms and dots over them equal the reference's free-dof norms and dots.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from .linear import _Scratch, axpby, dot


def _norm(x, scr) -> float:
    return float(np.sqrt(max(dot(x, x, scr), 0.0)))


def _is_invalid_lambda(exc: Exception) -> bool:
    return isinstance(exc, ValueError) and "lambda must be positive" in str(exc)


@dataclass
class NewtonStats:
    newton_iterations: int = 0
    linear_iterations: int = 0
    residual_evals: int = 0
    history: list = field(default_factory=list)


def damping(it_damp_max, U, dU, F, f, builder, scr) -> float:
    """Line search for plain Newton (nonlinear/damping.py:45-134)."""
    if it_damp_max < 0:
        return 0.0
    res = F - f
    decrease0 = dot(res, dU, scr)
    dnorm = _norm(dU, scr)
    if not np.isfinite(decrease0) or not np.isfinite(dnorm) or decrease0 >= 0.0:
        return 0.0
    alpha, a_min, a_max = 1.0, 0.0, 1.0
    Ua = torch.empty_like(U)
    for _ in range(int(it_damp_max)):
        torch.add(U, dU, alpha=alpha, out=Ua)
        Fa = builder.build_F_reduced_vec(Ua)
        dec = dot(Fa - f, dU, scr)
        if dec < 0.0:
            if alpha == 1.0:
                break
            a_min = alpha
        else:
            a_max = alpha
        alpha = 0.5 * (a_min + a_max)
    return float(alpha)


def damping_alg5(it_damp_max, U, lam, dU, dl, f, criterion, builder, scr) -> float:
    """Line search for the nested Newton (nonlinear/damping.py:137-214)."""
    if not np.isfinite(dl) or not np.isfinite(criterion) or it_damp_max <= 0:
        return 0.0
    alpha = 1.0
    last = None
    Ua = torch.empty_like(U)
    for _ in range(int(it_damp_max)):
        la = lam + alpha * dl
        if la <= 0.0:
            alpha *= 0.5
            continue
        torch.add(U, dU, alpha=alpha, out=Ua)
        try:
            Fa = builder.build_F_all_vec(la, Ua)
        except ValueError as exc:
            if not _is_invalid_lambda(exc):
                raise
            alpha *= 0.5
            continue
        crit = _norm(Fa - f, scr)
        last = float(alpha)
        if crit < criterion:
            break
        alpha *= 0.5
    if last is None:
        return 0.0
    return float(alpha if alpha == last else last)


def _free(v, mask):
    return v * mask


def newton(U_ini, tol, it_newt_max, it_damp_max, r_min, f, builder, solver, *, free_mask, stats=None):
    """Plain Newton for F(U) = f at the current reduction.  Returns (U, flag, it)."""
    scr = _Scratch(U_ini.device)
    U = U_ini.clone()
    norm_f = _norm(f, scr) or 1.0
    r = float(r_min)
    it = 0
    flag = 0
    compute_diffs = True
    while True:
        it += 1
        if compute_diffs:
            builder.constitutive_problem_stress_tangent(U)
            F = builder.build_F_vec().clone()
            K_r = builder.build_K_regularized(r)
        else:
            F = builder.build_F_reduced_vec(U).clone()
        crit = _norm(F - f, scr) / norm_f
        if stats is not None:
            stats.residual_evals += 1
        if compute_diffs and crit < tol:
            break
        rhs = f - F
        if not compute_diffs:
            K_r = builder.build_K_regularized(r)
        solver.setup_preconditioner(K_r)
        solver.A_orthogonalize(K_r)
        dU = solver.solve(K_r, rhs)
        dU.mul_(free_mask)
        alpha = damping(it_damp_max, U, dU, F, f, builder, scr)
        compute_diffs = True
        if alpha < 1e-1:
            if alpha == 0.0:
                compute_diffs = False
                r *= 2.0
            else:
                r *= 2.0 ** 0.25
        else:
            solver.expand_deflation_basis(dU)
            if alpha > 0.5:
                r = max(r / np.sqrt(2.0), r_min)
        if stats is not None:
            stats.newton_iterations += 1
        if alpha == 0.0 and r > 1.0:
            flag = 1
            break
        U.add_(dU, alpha=alpha)
        if not np.isfinite(crit) or it == it_newt_max:
            flag = 1
            break
    return U, flag, it


def newton_ind_ssr(U_ini, omega, lambda_it, it_newt_max, it_damp_max, tol, r_min, f, builder, solver, *,
                   free_mask, stats: NewtonStats | None = None, first_residual_gate: float | None = None,
                   progress_gate: bool = False):
    """Nested Newton for F_lambda(U) = f with f^T U = omega.  Returns (U, lambda, flag, it, info).

    ``first_residual_gate``: if given, the attempt is abandoned (flag 2) when the first relative
    residual exceeds it.  ``progress_gate``: the attempt is abandoned (flag 2) when the first Newton
    step does not reduce the residual.  Both are acceptance tests of a learned initial guess."""
    scr = _Scratch(U_ini.device)
    U = U_ini.clone()
    norm_f = _norm(f, scr) or 1.0
    eps = tol / 1000.0
    r = float(r_min)
    it = 0
    flag = 0
    compute_diffs = True
    rel = np.nan
    info = {"first_rel_residual": None, "rel_residual": []}
    while True:
        it += 1
        if compute_diffs:
            builder.reduction(lambda_it)
            builder.constitutive_problem_stress_tangent(U)
            F = builder.build_F_vec().clone()
            K_r = builder.build_K_regularized(r)
        else:
            F = builder.build_F_all_vec(lambda_it, U).clone()
        crit = _norm(F - f, scr)
        rel = crit / norm_f
        info["rel_residual"].append(rel)
        if it == 1:
            info["first_rel_residual"] = rel
            if first_residual_gate is not None and not (rel < first_residual_gate):
                flag = 2
                break
        elif it == 2 and progress_gate and not (rel < info["first_rel_residual"]):
            flag = 2
            break
        if stats is not None:
            stats.residual_evals += 1
        if compute_diffs and rel < tol and it > 1:
            break
        if not compute_diffs:
            K_r = builder.build_K_regularized(r)  # tangent cached from the last evaluation at this U
        # G = dF/dlambda by forward difference
        F_eps = builder.build_F_all_vec(lambda_it + eps, U)
        G = (F_eps - F) / eps
        solver.setup_preconditioner(K_r)
        solver.A_orthogonalize(K_r)
        dW = solver.solve(K_r, -G).mul_(free_mask)
        dV = solver.solve(K_r, f - F).mul_(free_mask)
        denom = dot(f, dW, scr)
        d_l = 0.0 if abs(denom) < 1e-30 else -dot(f, dV, scr) / denom
        dU = dV + d_l * dW
        alpha = damping_alg5(it_damp_max, U, lambda_it, dU, d_l, f, crit, builder, scr)
        compute_diffs = True
        if alpha < 1e-1:
            if alpha == 0.0:
                compute_diffs = False
                r *= 2.0
            else:
                r *= 2.0 ** 0.25
        else:
            solver.expand_deflation_basis(dW)
            solver.expand_deflation_basis(dV)
            if alpha > 0.5:
                r = max(r / np.sqrt(2.0), r_min)
        if stats is not None:
            stats.newton_iterations += 1
        if alpha == 0.0 and r > 1.0:
            if rel > 10.0 * tol:
                flag = 1
            break
        U.add_(dU, alpha=alpha)
        d = dot(f, U, scr)
        if d != 0.0:
            U.mul_(omega / d)
        lambda_it = lambda_it + alpha * d_l
        if not np.isfinite(rel) or it == it_newt_max:
            if rel > 10.0 * tol:
                flag = 1
            break
    info["final_rel_residual"] = rel
    return U, float(lambda_it), flag, it, info


def newton_ind_ll(U_ini, t_ini, omega, it_newt_max, it_damp_max, tol, r_min, f, builder, solver, *, free_mask,
                  stats: NewtonStats | None = None):
    """Nested Newton for F(U) = t f with f^T U = omega (nonlinear/newton.py:1131-1487): residual
    |t f - F(U)| / |f|; solves K dW = f and K dV = t f - F; d_t = -f^T dV / f^T dW; damping of the
    plain-Newton kind on (F, 0) (the reference passes a zero load to `damping`); U is rescaled to
    omega and t advances by the full d_t.  Returns (U, t, flag, it)."""
    scr = _Scratch(U_ini.device)
    U = U_ini.clone()
    norm_f = _norm(f, scr) or 1.0
    r = float(r_min)
    it = 0
    flag = 0
    t = float(t_ini)
    compute_diffs = True
    zero = torch.zeros_like(f)
    while True:
        it += 1
        if compute_diffs:
            builder.constitutive_problem_stress_tangent(U)
            F = builder.build_F_vec().clone()
            K_r = builder.build_K_regularized(r)
        else:
            F = builder.build_F_reduced_vec(U).clone()
        rel = _norm(t * f - F, scr) / norm_f
        if stats is not None:
            stats.residual_evals += 1
        if compute_diffs and rel < tol and it > 1:
            break
        if not compute_diffs:
            K_r = builder.build_K_regularized(r)
        solver.setup_preconditioner(K_r)
        solver.A_orthogonalize(K_r)
        dW = solver.solve(K_r, f.clone()).mul_(free_mask)
        dV = solver.solve(K_r, t * f - F).mul_(free_mask)
        denom = dot(f, dW, scr)
        d_t = 0.0 if abs(denom) < 1e-30 else -dot(f, dV, scr) / denom
        dU = dV + d_t * dW
        alpha = damping(it_damp_max, U, dU, F, zero, builder, scr)
        compute_diffs = True
        if alpha < 1e-1:
            if alpha == 0.0:
                compute_diffs = False
                r *= 2.0
            else:
                r *= 2.0 ** 0.25
        else:
            solver.expand_deflation_basis(dW)
            solver.expand_deflation_basis(dV)
            if alpha > 0.5:
                r = max(r / np.sqrt(2.0), r_min)
        if stats is not None:
            stats.newton_iterations += 1
        if alpha == 0.0 and r > 1.0:
            flag = 1
            break
        U.add_(dU, alpha=alpha)
        d = dot(f, U, scr)
        if d != 0.0:
            U.mul_(omega / d)
        t = t + d_t
        if not np.isfinite(rel) or it == it_newt_max:
            flag = 1
            break
    return U, float(t), flag, it
