"""Device MLP (bf16 MFMA GEMMs, tanh, Adam) against the CPU oracle; training reduces the loss."""

import numpy as np
import pytest
import torch

from oracle import mlp as O
from slope_stability_petsc4py_amd.learn import NodeMLP, train

pytestmark = pytest.mark.gpu


def _load(net, W, b):
    net.load_state_dict({"W": [w.astype(np.float32) for w in W], "b": [x.astype(np.float32) for x in b]})


def test_forward_backward_vs_oracle(gpu):
    print("case: MLP fwd/bwd vs oracle", flush=True)
    rng = np.random.default_rng(0)
    net = NodeMLP(n_in=16, hidden=128, layers=4, n_out=3, seed=1, zero_last=False)
    W = [w.cpu().numpy().astype(np.float64) for w in net.W]
    b = [rng.normal(size=x.numel()) * 0.1 for x in net.b]
    _load(net, W, b)
    W = [w.cpu().numpy().astype(np.float64) for w in net.W]
    b = [x.cpu().numpy().astype(np.float64) for x in net.b]
    M = 1000 + 37  # not a multiple of 16
    X = rng.normal(size=(M, 16)).astype(np.float32)
    T = rng.normal(size=(M, 3)).astype(np.float32)
    out = net.forward(torch.from_numpy(X).to(gpu))
    acts = O.forward(W, b, X)
    np.testing.assert_allclose(out.cpu().numpy(), acts[-1], rtol=0, atol=2e-2 * np.abs(acts[-1]).max())
    loss, g = net.mse(out, torch.from_numpy(T).to(gpu))
    lo, go = O.mse(acts[-1], T)
    assert abs(loss - lo) < 2e-2 * lo
    net.backward(g)
    gW, gb = O.backward(W, acts, go)
    for li in range(len(W)):
        gd = net.gW[li].cpu().numpy()
        assert np.abs(gd - gW[li]).max() < 3e-2 * np.abs(gW[li]).max(), li
        assert np.abs(net.gb[li].cpu().numpy() - gb[li]).max() < 3e-2 * np.abs(gb[li]).max(), li


def test_training_reduces_loss(gpu):
    print("case: MLP training", flush=True)
    rng = np.random.default_rng(3)
    M = 4096
    X = rng.normal(size=(M, 16)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) * X[:, 7], 0.5 * X[:, 8] * X[:, 9], np.tanh(X[:, 13])], axis=1).astype(np.float32)
    net = NodeMLP(seed=2)
    Xd, Td = torch.from_numpy(X).to(gpu), torch.from_numpy(T).to(gpu)
    losses = []
    for _ in range(150):
        out = net.forward(Xd)
        loss, g = net.mse(out, Td)
        net.backward(g)
        net.adam(3e-3)
        losses.append(loss)
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])


def test_adam_matches_oracle(gpu):
    net = NodeMLP(n_in=16, hidden=16, layers=2, n_out=3, seed=4, zero_last=False)
    w0 = net.W[0].cpu().numpy().astype(np.float64)
    g = np.random.default_rng(5).normal(size=w0.shape).astype(np.float32)
    net.gW[0].copy_(torch.from_numpy(g))
    for li in range(1, len(net.W)):
        net.gW[li].zero_()
    for li in range(len(net.b)):
        net.gb[li].zero_()
    net.adam(1e-2)
    w1, _, _ = O.adam(w0, g.astype(np.float64), np.zeros_like(w0), np.zeros_like(w0), 1, lr=1e-2)
    np.testing.assert_allclose(net.W[0].cpu().numpy(), w1, rtol=1e-5, atol=1e-6)


def test_learned_predictor_keeps_fos_and_bounds_cost_on_hetero_ssr(gpu):
    """The network is trained (device forward/backward/Adam) on the continuation histories of
    three perturbed-material runs of the 3D heterogeneous L1 slope (the loss must fall), then fills
    the predictor slot of the SSR continuation on the reference materials (FP32 form).  Checked:
    the factor of safety stays within 1e-3 of the secant form's and of the reference's, and the
    first-residual gate keeps the Newton iterations within 10% (+2) of the secant form's.  The
    Newton and wall-time ratios (what the network gains or costs here) are printed, not asserted:
    on this slope the gate keeps most steps on the secant."""
    from slope_stability_petsc4py_amd import cases
    from slope_stability_petsc4py_amd.learn import LearnedPredictor, node_info, samples_from_history, train
    from slope_stability_petsc4py_amd.learned_slopes import hetero_variant
    print("case: learned predictor on hetero3d_L1", flush=True)
    samples = []
    for seed in range(3):
        case = hetero_variant(seed)
        res = cases.run_ssr(cases.make_solver(case, precision="fp32"), store_step_u=True)
        samples += samples_from_history(node_info(case.mesh, case.materials), res.step_U, res.omega_hist,
                                        res.lambda_hist)
        del res
    assert len(samples) >= 10
    net = NodeMLP(seed=0)
    losses = train(net, samples, epochs=60, lr=2e-3)
    assert losses[-1] < losses[0]
    ref = hetero_variant(-1)
    plain = cases.run_ssr(cases.make_solver(ref, precision="fp32"))
    pred = LearnedPredictor(net, node_info(ref.mesh, ref.materials), smoothing=2)
    lrn = cases.run_ssr(cases.make_solver(ref, precision="fp32"), predictor=pred)
    kinds = [r["predictor"] for r in lrn.step_info]
    print("gain: newton", plain.newton_iterations, "->", lrn.newton_iterations,
          f"({plain.newton_iterations / max(lrn.newton_iterations, 1):.3f}x),",
          "wall", round(plain.wall_time, 3), "->", round(lrn.wall_time, 3),
          f"({plain.wall_time / lrn.wall_time:.3f}x);", "steps used / gated:", kinds.count("learned"),
          "/", kinds.count("learned_gated"), "; fos", plain.fos, lrn.fos, flush=True)
    assert pred.calls > 0
    assert abs(lrn.fos - plain.fos) < 1e-3
    assert abs(lrn.fos - 1.6661099309952119) < 1e-3
    assert lrn.newton_iterations <= 1.1 * plain.newton_iterations + 2


def test_learned_predictor_on_seeded_homogeneous_slopes(gpu):
    """The learned-predictor configuration at a reduced size: seeded homogeneous slopes (height,
    face angle, c, phi drawn per seed in the configuration's ranges, h = 2 m), the network trained on
    six seeds' own SSR histories, evaluated on two held-out seeds against the plain (secant) method:
    the factor of safety within 1e-3 of the plain method's on every held-out seed; the Newton and
    wall-time ratios are printed."""
    from slope_stability_petsc4py_amd.learned_slopes import collect, evaluate, slope_case, summary
    print("case: learned predictor on seeded homogeneous slopes", flush=True)
    case_of = lambda seed: slope_case(seed, h=2.0)  # noqa: E731
    samples, runs = collect(case_of, range(6), log=lambda *a: None)
    assert len(samples) >= 30 and all(r["steps"] >= 5 for r in runs)
    net = NodeMLP(seed=0)
    losses = train(net, samples, epochs=30, lr=2e-3)
    assert losses[-1] < losses[0]
    evals = evaluate(net, case_of, [1000, 1001], log=lambda *a: None)
    sm = summary(evals)
    print("held-out:", [(e["seed"], e["newton_plain"], e["newton_learned"], round(e["fos_abs_diff"], 6))
                        for e in evals], "summary", sm, flush=True)
    assert all(e["fos_abs_diff"] < 1e-3 for e in evals)
