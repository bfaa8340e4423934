"""Newton / continuation control flow on the CPU with fake builders and solvers (the reference tests
its drivers the same way: tests/test_newton_cleanup.py, tests/test_ssr_indirect_init.py,
tests/test_damping.py).  The device dot kernel is replaced by a host dot for these tests only.

Model problem, per dof: F_lambda(U) = (k c / lambda) tanh(U / u0), load f.  Its limit strength
factor is lambda* = k c / f exactly (tanh -> 1), so the continuation must approach it from below.
"""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from slope_stability_petsc4py_amd import continuation, nonlinear


class _Scr:
    def __init__(self, dev=None):
        pass


def _dot(x, y, scr=None):
    return float(torch.dot(x, y).item())


@pytest.fixture(autouse=True)
def host_dot(monkeypatch):
    monkeypatch.setattr(nonlinear, "dot", _dot)
    monkeypatch.setattr(nonlinear, "_Scratch", _Scr)
    monkeypatch.setattr(continuation, "dot", _dot)
    monkeypatch.setattr(continuation, "_Scratch", _Scr)


class TanhBuilder:
    def __init__(self, n=6, k=2.0, c=3.0, u0=0.1):
        self.k, self.c, self.u0 = k, c, u0
        self.lam = None
        self.U = None
        self.calls = {"tangent": 0, "stress": 0}

    def reduction(self, lam):
        if lam <= 0:
            raise ValueError("Reduction parameter lambda must be positive")
        self.lam = lam

    def _F(self, U):
        return self.k * self.c / self.lam * torch.tanh(U / self.u0)

    def constitutive_problem_stress_tangent(self, U):
        self.U = U.clone()
        self.calls["tangent"] += 1

    def constitutive_problem_stress(self, U):
        self.U = U.clone()
        self.calls["stress"] += 1

    def build_F_vec(self):
        return self._F(self.U)

    def build_F_all_vec(self, lam, U):
        self.reduction(lam)
        self.constitutive_problem_stress(U)
        return self._F(U)

    def build_F_reduced_vec(self, U):
        self.constitutive_problem_stress(U)
        return self._F(U)

    def build_K_regularized(self, r):
        d = self.k * self.c / self.lam / self.u0 * (1 - torch.tanh(self.U / self.u0) ** 2)
        el = self.k * self.c / self.lam / self.u0
        return SimpleNamespace(diag=(1 - r) * d + r * el)


    def elastic_matrix(self):
        return SimpleNamespace(diag=torch.zeros(6, dtype=torch.float64) + self.k * self.c / self.lam / self.u0)


class DiagSolver:
    def __init__(self):
        self.basis = []
        self.iteration_collector = SimpleNamespace(get_total_iterations=lambda: 0)

    def setup_preconditioner(self, A, **kw):
        self.A = A

    def A_orthogonalize(self, A):
        pass

    def solve(self, A, b, **kw):
        return b / A.diag

    def expand_deflation_basis(self, v):
        self.basis.append(v.clone())

    def get_deflation_basis_snapshot(self):
        return list(self.basis)

    def restore_deflation_basis(self, snap):
        self.basis = list(snap)


def test_plain_newton_converges_and_expands_basis():
    b = TanhBuilder()
    s = DiagSolver()
    f = torch.full((6,), 1.0, dtype=torch.float64)
    b.reduction(1.0)
    U, flag, it = nonlinear.newton(torch.zeros(6, dtype=torch.float64), 1e-10, 50, 10, 1e-4, f, b, s,
                                   free_mask=torch.ones(6, dtype=torch.float64))
    assert flag == 0
    np.testing.assert_allclose(b._F(U).numpy(), 1.0, rtol=1e-9)
    assert len(s.basis) >= 1


def test_newton_fails_beyond_limit_and_flags():
    b = TanhBuilder()
    s = DiagSolver()
    f = torch.full((6,), 1.0, dtype=torch.float64)
    b.reduction(10.0)  # k c / lambda = 0.6 < f: no solution
    U, flag, it = nonlinear.newton(torch.zeros(6, dtype=torch.float64), 1e-10, 30, 10, 1e-4, f, b, s,
                                   free_mask=torch.ones(6, dtype=torch.float64))
    assert flag == 1


def test_damping_alg5_skips_nonpositive_lambda():
    b = TanhBuilder()
    f = torch.full((4,), 1.0, dtype=torch.float64)
    U = torch.full((4,), 0.05, dtype=torch.float64)
    b.reduction(1.0)
    crit = float(torch.linalg.vector_norm(b._F(U) - f))
    # a step that would drive lambda negative at alpha = 1 must halve first
    a = nonlinear.damping_alg5(10, U, 1.0, torch.zeros(4, dtype=torch.float64), -1.5, f, crit, b, None)
    assert 0.0 < a <= 0.5


def test_ssr_continuation_approaches_limit_factor():
    k, c, fval = 2.0, 3.0, 1.0
    b = TanhBuilder(k=k, c=c)
    s = DiagSolver()
    n = 6
    f = torch.full((n,), fval, dtype=torch.float64)
    res = continuation.ssr_indirect_continuation(
        1.0, 0.5, 1e-5, 1e-4, 60, 1e9, 50, 10, 1e-9, 1e-4, f, b, s,
        free_mask=torch.ones(n, dtype=torch.float64), dim=3)
    lam_star = k * c / fval
    assert res.lambda_hist[-1] < lam_star
    assert res.lambda_hist[-1] > lam_star - 0.05
    assert np.all(np.diff(res.lambda_hist) > 0)
    assert np.all(np.diff(res.omega_hist) > 0)
    assert res.stop_reason in {"d_lambda_diff_scaled_min", "step_max"}


def test_ll_continuation_approaches_limit_load():
    """F_i(U) = k_i c tanh(U_i/u0) = t f_i has the limit load t* = min_i k_i c / f_i
    (continuation/limit_load.py); per-dof k keeps the omega constraint non-degenerate."""
    c, fval = 3.0, 1.5
    k = torch.tensor([2.0, 2.5, 3.0, 2.2, 4.0, 2.8], dtype=torch.float64)
    b = TanhBuilder(k=k, c=c)
    b.U = torch.zeros(6, dtype=torch.float64)
    s = DiagSolver()
    n = 6
    f = torch.full((n,), fval, dtype=torch.float64)
    res = continuation.ll_indirect_continuation(0.2, 1e-4, 80, 1e9, 50, 10, 1e-10, 1e-4, f, b, s,
                                                free_mask=torch.ones(n, dtype=torch.float64), dim=3)
    t_star = float(k.min()) * c / fval
    assert res.t_hist[0] == 0.0
    assert np.all(np.diff(res.t_hist) > 0)
    assert t_star - 0.02 < res.limit_load < t_star
    assert res.stop_reason in {"d_t_min", "step_max"}


def test_learned_predictor_gate_rejects_bad_guess():
    b = TanhBuilder()
    s = DiagSolver()
    n = 6
    f = torch.full((n,), 1.0, dtype=torch.float64)
    calls = {"n": 0}

    def bad_predictor(ctx):
        calls["n"] += 1
        # far off the branch (the symmetric branch has all dofs equal): first residual worse than the secant's
        return torch.tensor([5.0, -3.0, 5.0, -3.0, 5.0, -3.0], dtype=torch.float64)

    res = continuation.ssr_indirect_continuation(
        1.0, 0.5, 1e-5, 1e-3, 8, 1e9, 50, 10, 1e-9, 1e-4, f, b, s,
        free_mask=torch.ones(n, dtype=torch.float64), dim=3, predictor=bad_predictor)
    assert calls["n"] > 0
    kinds = {r["predictor"] for r in res.step_info}
    assert "learned" not in kinds and "learned_gated" in kinds
    # patience: after 3 guesses in a row gated off, the predictor is not called again
    assert calls["n"] == 3 and sum(r["predictor"] == "learned_gated" for r in res.step_info) == 3
    assert len(res.step_info) > 3
    calls["n"] = 0
    continuation.ssr_indirect_continuation(
        1.0, 0.5, 1e-5, 1e-3, 8, 1e9, 50, 10, 1e-9, 1e-4, f, TanhBuilder(), DiagSolver(),
        free_mask=torch.ones(n, dtype=torch.float64), dim=3, predictor=bad_predictor, predictor_patience=0)
    assert calls["n"] == len(res.step_info)      # patience 0: called at every attempt


def test_learned_predictor_exact_guess_is_used():
    k, c = 2.0, 3.0
    b = TanhBuilder(k=k, c=c)
    s = DiagSolver()
    n = 6
    f = torch.full((n,), 1.0, dtype=torch.float64)

    def exact(ctx):
        # the branch point at omega_target: U_i = omega/n, lambda = k c tanh(U/u0) / f
        return torch.full((n,), ctx["omega_target"] / n, dtype=torch.float64)

    res = continuation.ssr_indirect_continuation(
        1.0, 0.5, 1e-5, 1e-3, 8, 1e9, 50, 10, 1e-9, 1e-4, f, b, s,
        free_mask=torch.ones(n, dtype=torch.float64), dim=3, predictor=exact)
    assert any(r["predictor"] == "learned" for r in res.step_info)


def test_oracle_a_orthogonalize_pins():
    """The A-orthonormalisation oracle: W A W^T = diag(signs) on the kept rows, an exactly
    dependent row is dropped (newest-first processing keeps the newest), and an indefinite A gives
    a negative sign."""
    import oracle
    rng = np.random.default_rng(0)
    n = 30
    M = rng.normal(size=(n, n))
    A = M @ M.T + n * np.eye(n)
    B = rng.normal(size=(4, n))
    B = np.vstack([B, B[0] + B[2]])
    W, S = oracle.a_orthogonalize(B, A)
    assert W.shape == (4, n) and np.all(S == 1.0)
    assert np.abs(W @ A @ W.T - np.eye(4)).max() < 1e-12
    # the newest row survives unchanged up to scale
    w_new = W[-1]
    assert abs(abs(w_new @ B[-1]) / (np.linalg.norm(w_new) * np.linalg.norm(B[-1])) - 1.0) < 1e-12
    D = np.diag(np.r_[-1.0, np.ones(n - 1)])
    W2, S2 = oracle.a_orthogonalize(np.eye(n)[:1] * 3.0, D)
    assert S2.tolist() == [-1.0] and abs(W2[0, 0] - 1.0) < 1e-15


def test_load_offset_builder_subtracts_constant_load():
    """Limit load with a constant load: every internal-force call returns F(U) - f_const and every
    other call is forwarded unchanged."""
    import torch
    from slope_stability_petsc4py_amd.continuation import LoadOffsetBuilder

    class Fake:
        def __init__(self):
            self.calls = []

        def build_F_vec(self):
            return torch.tensor([1.0, 2.0])

        def build_F_reduced_vec(self, U):
            return U * 2.0

        def build_F_all_vec(self, lam, U):
            return U * lam

        def reduction(self, lam):
            self.calls.append(lam)
            return "ok"

    fb = Fake()
    f0 = torch.tensor([0.5, -1.0])
    b = LoadOffsetBuilder(fb, f0)
    U = torch.tensor([3.0, 4.0])
    assert torch.equal(b.build_F_vec(), torch.tensor([0.5, 3.0]))
    assert torch.equal(b.build_F_reduced_vec(U), torch.tensor([5.5, 9.0]))
    assert torch.equal(b.build_F_all_vec(2.0, U), torch.tensor([5.5, 9.0]))
    assert b.reduction(1.5) == "ok" and fb.calls == [1.5]


def _ll_tanh_case():
    c, fval = 3.0, 1.5
    k = torch.tensor([2.0, 2.5, 3.0, 2.2, 4.0, 2.8], dtype=torch.float64)
    b = TanhBuilder(k=k, c=c)
    b.U = torch.zeros(6, dtype=torch.float64)
    f = torch.full((6,), fval, dtype=torch.float64)
    return b, f, k, c, fval


def test_ll_learned_predictor_slot_gate():
    """Limit-load predictor slot: an exact branch point is used (and takes fewer Newton iterations
    than the secant), a guess off the branch is gated; the limit load is the plain method's."""
    b, f, k, c, fval = _ll_tanh_case()
    args = (0.2, 1e-4, 80, 1e9, 50, 10, 1e-10, 1e-4, f, b, DiagSolver())
    plain = continuation.ll_indirect_continuation(*args, free_mask=torch.ones(6, dtype=torch.float64), dim=3,
                                                  store_step_u=True)
    states = [(float(w), U.clone()) for w, U in zip(plain.omega_hist, plain.step_U)]
    assert len(plain.step_U) == len(plain.t_hist)

    def exact(ctx):
        w = ctx["omega_target"]
        hit = [U for (wo, U) in states if abs(wo - w) <= 1e-9 * max(1.0, abs(w))]
        return hit[0] if hit else None

    b2, *_ = _ll_tanh_case()
    res = continuation.ll_indirect_continuation(0.2, 1e-4, 80, 1e9, 50, 10, 1e-10, 1e-4, f, b2, DiagSolver(),
                                                free_mask=torch.ones(6, dtype=torch.float64), dim=3,
                                                predictor=exact)
    kinds = [r["predictor"] for r in res.step_info]
    assert "learned" in kinds
    assert abs(res.limit_load - plain.limit_load) < 1e-6
    assert res.newton_iterations < plain.newton_iterations

    b3, *_ = _ll_tanh_case()
    bad = continuation.ll_indirect_continuation(
        0.2, 1e-4, 80, 1e9, 50, 10, 1e-10, 1e-4, f, b3, DiagSolver(), free_mask=torch.ones(6, dtype=torch.float64),
        dim=3, predictor=lambda ctx: torch.tensor([5.0, -3.0, 5.0, -3.0, 5.0, -3.0], dtype=torch.float64))
    kinds = {r["predictor"] for r in bad.step_info}
    assert "learned" not in kinds and "learned_gated" in kinds
    assert abs(bad.limit_load - plain.limit_load) < 1e-12


def test_reference_probes_take_the_owned_free_path():
    """The reference's Newton and damping look for the owned/local force builders by name:
    nonlinear/newton.py:329-338 (callable attribute and a non-None owned_tangent_pattern) and
    nonlinear/damping.py:105-106,181-182 (callable attribute).  On ConstitutiveOperator those
    probes succeed, and the builders evaluate in the reference's order (reduction, stress, free
    rows of the force)."""
    from slope_stability_petsc4py_amd.constitutive import ConstitutiveOperator

    calls = []

    class Probe(ConstitutiveOperator):
        def __init__(self):   # no device state: the call order is what is checked
            self.owned_tangent_pattern = object()

        def reduction(self, lam):
            calls.append(("reduction", lam))

        def constitutive_problem_stress(self, U):
            calls.append(("stress", U))

        def build_F_free(self):
            calls.append(("free",))
            return "F_free"

    op = Probe()

    def supports_local(builder, name):   # nonlinear/newton.py:334-336
        return callable(getattr(builder, name, None)) and getattr(builder, "owned_tangent_pattern", None) is not None

    for name in ("build_F_free_local", "build_F_all_free_local", "build_F_reduced_free_local",
                 "build_F_K_tangent_reduced_free", "build_F_K_regularized_reduced_free"):
        assert supports_local(op, name), name
    assert op.build_F_all_free_local(1.3, "U") == "F_free"
    assert calls == [("reduction", 1.3), ("stress", "U"), ("free",)]
    calls.clear()
    assert op.build_F_reduced_free_local("U2") == "F_free"
    assert calls == [("stress", "U2"), ("free",)]
    calls.clear()
    assert op.build_F_free_local() == "F_free" and calls == [("free",)]
