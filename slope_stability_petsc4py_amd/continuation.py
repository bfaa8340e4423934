"""Indirect continuation for the factor of safety (SSR) and the limit load (LL), driving the device
Newton solvers.

SSR (continuation/indirect.py:1627-1736 init phase, :2267-3557 main loop, classic mode with the
legacy omega controller): two converged states at lambda_init and lambda_init + d_lambda start the
branch; each step targets omega + d_omega, predicts U by the secant in omega
(continuation/indirect.py:550-566) or by a supplied predictor, runs the nested Newton, halves
d_omega on failure (at most 5 times in a row), doubles it after a step whose lambda increment
shrank by more than 10%, and stops when d_lambda/d_omega * (omega - omega_0) falls below
d_lambda_diff_scaled_min, omega reaches omega_max_stop, or step_max steps were taken.
The factor of safety is the last accepted lambda.

A predictor slot takes a learned initial guess; it is accepted only when the nested Newton's
first residual is no larger than the secant guess would have been allowed to start from (the
attempt is gated, and the secant is used when the gate rejects it).

LL (continuation/limit_load.py:66-374): same structure over the load factor t with F(U) = t f.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from .linear import _Scratch, dot
from .nonlinear import NewtonStats, newton, newton_ind_ll, newton_ind_ssr


@dataclass
class SSRResult:
    lambda_hist: np.ndarray
    omega_hist: np.ndarray
    Umax_hist: np.ndarray
    U: torch.Tensor
    stop_reason: str
    wall_time: float
    newton_iterations: int
    linear_iterations: int
    attempts: int
    step_U: list = field(default_factory=list)
    step_info: list = field(default_factory=list)

    @property
    def fos(self) -> float:
        return float(self.lambda_hist[-1])


def _umax(U: torch.Tensor, dim: int) -> float:
    return float(torch.linalg.vector_norm(U.view(-1, dim), dim=1).max().item())


def init_phase_ssr(lambda_init, d_lambda_init, d_lambda_min, it_newt_max, it_damp_max, tol, r_min, f, builder,
                   solver, *, free_mask, stats):
    """Two converged states on the branch (continuation/indirect.py:1627-1736)."""
    scr = _Scratch(f.device)
    U_ini = torch.zeros_like(f)
    d_lambda = float(d_lambda_init)
    lambda1 = float(lambda_init)
    while True:
        builder.reduction(lambda1)
        U_it, flag, _ = newton(U_ini, tol, it_newt_max, it_damp_max, r_min, f, builder, solver, free_mask=free_mask,
                               stats=stats)
        if flag == 0:
            break
        lambda1 *= 0.5
        d_lambda *= 0.5
        if d_lambda < d_lambda_min:
            raise RuntimeError("Initial choice of lambda seems to be too large.")
    U1 = U_it
    omega1 = dot(f, U1, scr)
    solver.expand_deflation_basis(U1)
    while True:
        lambda_it = lambda1 + d_lambda
        builder.reduction(lambda_it)
        U_it, flag, _ = newton(U1, tol, it_newt_max, it_damp_max, r_min, f, builder, solver, free_mask=free_mask,
                               stats=stats)
        if flag == 1:
            d_lambda /= 2.0
        else:
            U2 = U_it
            omega2 = dot(f, U2, scr)
            lambda2 = lambda_it
            if (omega2 - omega1) / max(1.0, omega1) < 1e-5:
                U1, lambda1, omega1 = U2, lambda2, omega2
            else:
                break
        if d_lambda < d_lambda_min:
            raise RuntimeError("It seems that the FoS is equal to lambda_init.")
        if lambda1 > 10.0:
            raise RuntimeError("It seems that the FoS is greater than 10.")
    return U1, U2, omega1, omega2, lambda1, lambda2


def ssr_indirect_continuation(lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max,
                              omega_max_stop, it_newt_max, it_damp_max, tol, r_min, f, builder, solver, *,
                              free_mask, dim, predictor=None, store_step_u=False, progress=None,
                              predictor_weights=(1.0, 0.5, 0.25), predictor_gate: str = "first_residual",
                              predictor_patience: int = 3) -> SSRResult:
    """SSR indirect continuation; returns the lambda/omega/Umax histories (FoS = lambda_hist[-1]).

    ``predictor(ctx) -> U_guess or None`` is the learned-predictor slot: ctx holds the accepted
    states, histories and the omega target; it returns U_guess or (U_guess, lambda_guess).  The
    guess is blended with the secant guess at the weights ``predictor_weights`` (1 = the guess
    itself), the blend with the smallest first residual (at lambda_guess, default the current
    lambda) is tried first.  After ``predictor_patience`` learned guesses in a row were gated off or
    rejected, the predictor is no longer called (its inference and the gate's residual evaluations
    then cost nothing).  ``predictor_gate`` "first_residual": it is kept only if that residual
    is below the secant's; "newton_progress": Newton starts from it and the attempt is dropped
    (the secant then used) when the first Newton step does not reduce the residual."""
    t0 = perf_counter()
    stats = NewtonStats()
    scr = _Scratch(f.device)
    lin0 = solver.iteration_collector.get_total_iterations()
    U_old, U, omega_old, omega, lambda_old, lambda_value = init_phase_ssr(
        lambda_init, d_lambda_init, d_lambda_min, it_newt_max, it_damp_max, tol, r_min, f, builder, solver,
        free_mask=free_mask, stats=stats)
    solver.expand_deflation_basis(U_old)
    lambda_hist = [lambda_old, lambda_value]
    omega_hist = [omega_old, omega]
    umax_hist = [_umax(U_old, dim), _umax(U, dim)]
    step_U = [U_old.clone(), U.clone()] if store_step_u else []
    states = [(U_old, omega_old, lambda_old), (U, omega, lambda_value)]  # last accepted states (<= 3)
    step_info = []
    d_omega = omega - omega_old
    if omega_max_stop < omega + d_omega:
        raise ValueError("Too small value of omega_max_stop. Increase it and rerun.")
    step = 2
    n_omega = 0
    attempts = 0
    misses = 0          # learned guesses gated off or rejected in a row
    stop_reason = "completed"
    while True:
        omega_it = min(omega + d_omega, omega_max_stop)
        d_omega = omega_it - omega
        alpha_sec = 0.0 if abs(omega - omega_old) <= 1e-30 else (omega_it - omega) / (omega - omega_old)
        U_sec = U + alpha_sec * (U - U_old)
        kind = "secant"
        learned = None
        if predictor is not None and misses >= predictor_patience > 0:
            predictor = None
        if predictor is not None:
            learned = predictor({"U": U, "U_old": U_old, "omega": omega, "omega_old": omega_old,
                                 "omega_target": omega_it, "lambda": lambda_value, "lambda_old": lambda_old,
                                 "lambda_hist": lambda_hist, "omega_hist": omega_hist, "states": states})
        basis_snap = solver.get_deflation_basis_snapshot()
        t_att = perf_counter()
        newton_it_before = stats.newton_iterations
        result = None
        gate_info = None
        lam_guess = lambda_value
        if isinstance(learned, tuple):
            learned, lam_guess = learned[0], float(learned[1])
        if learned is not None:
            # rescale to the omega constraint, then gate on the first residual vs the secant's; the
            # learned correction is also tried at reduced weight (a residual line search along
            # learned - secant), so a correction that helps only in part is still used
            d = dot(f, learned, scr)
            if d != 0.0:
                learned = learned * (omega_it / d)
            builder.reduction(lambda_value)
            F_sec = builder.build_F_reduced_vec(U_sec).clone()
            r_sec = float(torch.linalg.vector_norm(F_sec - f).item())
            if lam_guess != lambda_value:
                builder.reduction(lam_guess)
            best, r_l, gamma_best = None, np.inf, 0.0
            for gamma in predictor_weights:
                cand = learned if gamma == 1.0 else U_sec + gamma * (learned - U_sec)
                F_l = builder.build_F_reduced_vec(cand)
                r_c = float(torch.linalg.vector_norm(F_l - f).item())
                if r_c < r_l:
                    best, r_l, gamma_best = cand, r_c, gamma
            learned = best
            gate_info = {"secant_first_residual": r_sec, "learned_first_residual": r_l, "learned_weight": gamma_best}
            if r_l < r_sec or predictor_gate == "newton_progress":
                kind = "learned"
                result = newton_ind_ssr(learned, omega_it, lam_guess, it_newt_max, it_damp_max, tol, r_min, f,
                                        builder, solver, free_mask=free_mask, stats=stats,
                                        progress_gate=predictor_gate == "newton_progress")
                if result[2] != 0:
                    solver.restore_deflation_basis(basis_snap)
                    result = None
                    kind = "learned_rejected"
            else:
                kind = "learned_gated"
        if result is None:
            result = newton_ind_ssr(U_sec, omega_it, lambda_value, it_newt_max, it_damp_max, tol, r_min, f, builder,
                                    solver, free_mask=free_mask, stats=stats)
            if kind == "learned_rejected":
                kind = "secant_after_rejection"
        U_it, lambda_cand, flag, it_newt, info = result
        solver.restore_deflation_basis(basis_snap)
        if learned is not None:
            misses = 0 if (kind == "learned" and flag == 0) else misses + 1
        attempts += 1
        rec = {"step": step + 1, "omega_target": omega_it, "flag": flag, "newton_it": it_newt, "predictor": kind,
               "attempt_wall": perf_counter() - t_att, "newton_total": stats.newton_iterations - newton_it_before,
               "first_rel_residual": info.get("first_rel_residual")}
        if gate_info:
            rec.update(gate_info)
        step_info.append(rec)
        if progress is not None:
            progress(dict(rec, lam=lambda_cand))
        if flag != 0:
            d_omega /= 2.0
            n_omega += 1
        else:
            step += 1
            U_old, U = U, U_it
            solver.expand_deflation_basis(U)
            omega_old, omega = omega, omega_it
            d_lambda = lambda_cand - lambda_value
            lambda_old, lambda_value = lambda_value, lambda_cand
            d_lambda_diff_scaled = np.nan if d_omega == 0.0 else (d_lambda / d_omega) * (omega - omega_hist[0])
            n_omega = 0
            lambda_hist.append(lambda_value)
            omega_hist.append(omega)
            umax_hist.append(_umax(U, dim))
            if store_step_u:
                step_U.append(U.clone())
            states = (states + [(U, omega, lambda_value)])[-3:]
            if (lambda_hist[-1] - lambda_hist[-2]) < 0.9 * (lambda_hist[-2] - lambda_hist[-3]):
                d_omega *= 2.0
            if d_lambda_diff_scaled_min > 0.0 and d_lambda_diff_scaled < d_lambda_diff_scaled_min:
                stop_reason = "d_lambda_diff_scaled_min"
                break
            if omega >= omega_max_stop:
                stop_reason = "omega_max_stop"
                break
        if n_omega >= 5:
            stop_reason = "omega_reduction_limit"
            break
        if step >= step_max:
            stop_reason = "step_max"
            break
    if f.is_cuda:
        torch.cuda.synchronize()
    return SSRResult(np.asarray(lambda_hist), np.asarray(omega_hist), np.asarray(umax_hist), U, stop_reason,
                     perf_counter() - t0, stats.newton_iterations,
                     solver.iteration_collector.get_total_iterations() - lin0, attempts, step_U, step_info)


@dataclass
class LLResult:
    t_hist: np.ndarray
    omega_hist: np.ndarray
    Umax_hist: np.ndarray
    U: torch.Tensor
    stop_reason: str
    wall_time: float
    newton_iterations: int
    linear_iterations: int
    step_U: list = field(default_factory=list)
    step_info: list = field(default_factory=list)

    @property
    def limit_load(self) -> float:
        return float(self.t_hist[-1])


class LoadOffsetBuilder:
    """A constitutive builder whose internal forces are F(U) - f_const: the limit-load continuation
    then solves F(U) = f_const + t f (a constant load, e.g. gravity, under a growing one, e.g. a
    footing pressure) with the reference's algorithm unchanged.  Every other call is forwarded."""

    def __init__(self, builder, f_const: torch.Tensor):
        self._b = builder
        self._f0 = f_const

    def __getattr__(self, name):
        return getattr(self._b, name)

    def build_F_vec(self, *a, **k):
        return self._b.build_F_vec(*a, **k) - self._f0

    def build_F_reduced_vec(self, *a, **k):
        return self._b.build_F_reduced_vec(*a, **k) - self._f0

    def build_F_all_vec(self, *a, **k):
        return self._b.build_F_all_vec(*a, **k) - self._f0


def ll_indirect_continuation(d_omega_ini_scale, d_t_min, step_max, omega_max_stop, it_newt_max, it_damp_max, tol,
                             r_min, f, builder, solver, *, free_mask, dim, lambda_ell: float = 1.0,
                             progress=None, f_const: torch.Tensor | None = None, predictor=None,
                             store_step_u: bool = False, predictor_patience: int = 3) -> LLResult:
    """Limit-load continuation in omega = f^T U (continuation/limit_load.py:66-374 with the driver's
    elastic start, cli/run_3D_hetero_SSR_capture.py:1224-1305): one elastic solve K_el U_el = f
    (strength reduced by lambda_ell), omega_el = f^T U_el, start from U = scale*U_el, t = 0, first
    omega step scale*omega_el; secant predictor in omega; d_omega halves on failure (5 in a row
    stop), doubles after an easy step (< 20 Newton iterations) whose t increment shrank by > 10%;
    stop at d_t < d_t_min, omega_max_stop or step_max.  The limit load factor is t_hist[-1].

    ``f_const`` (not in the reference, whose LL scales its one load): a constant load carried
    throughout, F(U) = f_const + t f.  The branch then starts from the equilibrium U_0 under f_const
    alone (plain Newton at t = 0), omega_0 = f^T U_0, and the elastic increment is added to U_0.

    ``predictor(ctx) -> U_guess or None``: the learned-predictor slot, as in the SSR continuation
    (ctx["states"] holds the last accepted (U, omega, t)); the guess is rescaled to the omega target
    and used only when its first residual |t f - F(U)| is below the secant guess's (the gate)."""
    t0 = perf_counter()
    stats = NewtonStats()
    scr = _Scratch(f.device)
    lin0 = solver.iteration_collector.get_total_iterations()
    builder.reduction(float(lambda_ell))
    U_0 = None
    if f_const is not None:
        U_0, flag0, _ = newton(torch.zeros_like(f), tol, it_newt_max, it_damp_max, r_min, f_const, builder, solver,
                               free_mask=free_mask, stats=stats)
        if flag0 != 0:
            raise RuntimeError("limit load: no equilibrium under the constant load alone")
        solver.expand_deflation_basis(U_0)
        builder = LoadOffsetBuilder(builder, f_const)
    K_el = builder.elastic_matrix()
    solver.setup_preconditioner(K_el)
    solver.A_orthogonalize(K_el)
    U_el = solver.solve(K_el, f.clone()).mul_(free_mask)
    solver.expand_deflation_basis(U_el)
    omega_el = dot(f, U_el, scr)
    d_omega = omega_el * float(d_omega_ini_scale)
    U = U_el * float(d_omega_ini_scale)
    omega = omega_old = 0.0
    if U_0 is not None:
        U = U + U_0
        omega = omega_old = dot(f, U_0, scr)
    U_old = U.clone()
    t = 0.0
    t_hist, omega_hist, umax_hist = [0.0], [omega], [_umax(U if U_0 is None else U_0, dim)]
    step_U = [(U_0 if U_0 is not None else torch.zeros_like(U)).clone()] if store_step_u else []
    states = [((U_0 if U_0 is not None else torch.zeros_like(U)), omega, 0.0)]
    step_info = []
    step = 1
    n_omega = 0
    misses = 0          # learned guesses gated off or rejected in a row (predictor_patience)
    stop_reason = "completed"
    while True:
        omega_it = omega + d_omega
        if omega_it > omega_max_stop:
            omega_it = omega_max_stop
            d_omega = omega_it - omega
        if step > 1:
            den = omega - omega_old
            U_ini = U if den == 0.0 else U + (d_omega / den) * (U - U_old)
        else:
            U_ini = U
        kind = "secant"
        gate_info = {}
        if predictor is not None and misses >= predictor_patience > 0:
            predictor = None
        if predictor is not None and step > 1:
            guess = predictor({"U": U, "U_old": U_old, "omega": omega, "omega_old": omega_old,
                               "omega_target": omega_it, "lambda": t, "states": states})
            if guess is not None:
                d = dot(f, guess, scr)
                if d != 0.0:
                    guess = guess * (omega_it / d)
                r_sec = float(torch.linalg.vector_norm(t * f - builder.build_F_reduced_vec(U_ini)).item())
                r_l = float(torch.linalg.vector_norm(t * f - builder.build_F_reduced_vec(guess)).item())
                gate_info = {"secant_first_residual": r_sec, "learned_first_residual": r_l}
                if r_l < r_sec:
                    U_ini, kind = guess, "learned"
                else:
                    kind = "learned_gated"
        basis_snap = solver.get_deflation_basis_snapshot()
        U_it, t_it, flag, it_newt = newton_ind_ll(U_ini, t, omega_it, it_newt_max, it_damp_max, tol, r_min, f,
                                                  builder, solver, free_mask=free_mask, stats=stats)
        if flag != 0 and kind == "learned":
            # the learned start failed: the secant start is tried before the step is halved
            solver.restore_deflation_basis(basis_snap)
            den = omega - omega_old
            U_sec = U if den == 0.0 else U + (d_omega / den) * (U - U_old)
            U_it, t_it, flag, it_newt = newton_ind_ll(U_sec, t, omega_it, it_newt_max, it_damp_max, tol, r_min, f,
                                                      builder, solver, free_mask=free_mask, stats=stats)
            kind = "secant_after_rejection"
        if kind != "secant":
            misses = 0 if kind == "learned" else misses + 1
        solver.restore_deflation_basis(basis_snap)
        step_info.append({"step": step, "omega_target": omega_it, "flag": flag, "newton_it": it_newt,
                          "predictor": kind, **gate_info})
        if progress is not None:
            progress({"omega_target": omega_it, "flag": flag, "newton_it": it_newt, "t": t_it})
        if flag == 1:
            d_omega /= 2.0
            n_omega += 1
        else:
            step += 1
            U_old, U = U, U_it
            solver.expand_deflation_basis(U)
            omega_old, omega = omega, omega_it
            d_t = t_it - t
            t = t_it
            n_omega = 0
            t_hist.append(t)
            omega_hist.append(omega)
            umax_hist.append(_umax(U, dim))
            if store_step_u:
                step_U.append(U.clone())
            states = (states + [(U, omega, t)])[-3:]
            if d_t < d_t_min:
                stop_reason = "d_t_min"
                break
            if it_newt < 20 and step > 2 and (t_hist[-1] - t_hist[-2]) < 0.9 * (t_hist[-2] - t_hist[-3]):
                d_omega *= 2.0
        if n_omega >= 5:
            stop_reason = "omega_reduction_limit"
            break
        if omega >= omega_max_stop:
            stop_reason = "omega_max"
            break
        if step >= step_max:
            stop_reason = "step_max"
            break
    if f.is_cuda:
        torch.cuda.synchronize()
    return LLResult(np.asarray(t_hist), np.asarray(omega_hist), np.asarray(umax_hist), U, stop_reason,
                    perf_counter() - t0, stats.newton_iterations,
                    solver.iteration_collector.get_total_iterations() - lin0, step_U, step_info)
