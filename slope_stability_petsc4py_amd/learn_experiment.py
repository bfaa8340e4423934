"""Learned-predictor experiment: seeded 3D homogeneous slopes (height, face angle, c, phi drawn per
seed), their own SSR histories as training data, evaluation on held-out seeds against the plain
secant continuation (Newton iterations, wall time, FoS difference).

``python -m slope_stability_petsc4py_amd.learn_experiment --train 12 --test 4``
prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import cases, mesh
from .learn import LearnedPredictor, NodeMLP, node_info, samples_from_history, train
from .materials import MaterialSpec


def slope_case(seed: int, h: float = 2.0, nz: int = 6) -> cases.Case:
    rng = np.random.default_rng(seed)
    height = float(rng.uniform(8.0, 15.0))
    beta = float(rng.uniform(30.0, 60.0))
    c0 = float(rng.uniform(5.0, 20.0))
    phi = float(rng.uniform(25.0, 40.0))
    run = height / np.tan(np.deg2rad(beta))
    m = mesh.slope_3d(h=h, x1=10.0, x2=run, x3=10.0, y1=6.0, y2=height, z_len=12.0, nz=nz)
    mat = MaterialSpec(c0=c0, phi=phi, psi=0.0, young=40000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0)
    p = cases.SSRParams(lambda_init=0.7, d_lambda_init=0.1, d_lambda_min=1e-4, d_lambda_diff_scaled_min=1e-3,
                        step_max=40, omega_max_stop=1e12, it_newt_max=60)
    c = cases.Case(f"slope_seed{seed}", m, [mat], p)
    c.extra = {"height": height, "beta": beta, "c0": c0, "phi": phi}
    return c


def hetero_variant(seed: int, spread: float = 0.2, dphi: float = 3.0) -> cases.Case:
    """The 3D heterogeneous L1 slope (P2) with its four layer materials perturbed per seed:
    c0 scaled by U(1 - spread, 1 + spread), phi shifted by U(-dphi, dphi) degrees.  Seed < 0 gives
    the unperturbed reference materials."""
    base = cases.hetero3d_L1()
    if seed < 0:
        return base
    rng = np.random.default_rng(seed)
    mats = []
    for m in cases.HETERO_L1_MATERIALS:
        mats.append(MaterialSpec(c0=m.c0 * float(rng.uniform(1 - spread, 1 + spread)),
                                 phi=m.phi + float(rng.uniform(-dphi, dphi)), psi=m.psi, young=m.young,
                                 poisson=m.poisson, gamma_sat=m.gamma_sat, gamma_unsat=m.gamma_unsat))
    return cases.Case(f"hetero3d_L1_seed{seed}", base.mesh, mats, base.params)


def run_case(case, predictor=None, precision="fp64", store=False):
    s = cases.make_solver(case, precision=precision)
    pred = predictor(case) if callable(predictor) else None
    res = cases.run_ssr(s, predictor=pred, store_step_u=store)
    return s, res


def hetero_main(a, log):
    """Train on perturbed-material SSR histories of the heterogeneous L1 slope, evaluate on the
    reference materials: plain FP64 (secant), FP32 (secant), FP32 + learned predictor."""
    t0 = time.perf_counter()
    samples, train_info = [], []
    for seed in range(0 if a.load else a.train):
        case = hetero_variant(seed, a.spread, a.dphi)
        s, res = run_case(case, precision="fp32", store=True)
        info = node_info(case.mesh, case.materials)
        samples += samples_from_history(info, res.step_U, res.omega_hist, res.lambda_hist)
        train_info.append({"seed": seed, "fos": res.fos, "newton": res.newton_iterations,
                           "steps": len(res.lambda_hist), "wall": res.wall_time})
        log("train run", train_info[-1])
        del s, res
        torch.cuda.empty_cache()
    t_data = time.perf_counter() - t0
    t1 = time.perf_counter()
    if a.load:
        net = NodeMLP.load(a.load)
        losses = []
    else:
        net = NodeMLP(seed=0)
        info0 = node_info(hetero_variant(-1).mesh, cases.HETERO_L1_MATERIALS)
        losses = train(net, samples, epochs=a.epochs, lr=2e-3, log=lambda ep, l: log("epoch", ep, "loss", l),
                       info=info0 if a.train_smoothing else None, smoothing=a.train_smoothing)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    if a.save:
        net.save(a.save)
    ref = hetero_variant(-1)
    # one short warm-up run so every timed run below starts from loaded kernels and libraries
    warm = cases.make_solver(ref, precision="fp32")
    ref.params.step_max = 4
    cases.run_ssr(warm)
    ref.params.step_max = cases.SSRParams().step_max
    del warm
    evals = {}
    runs = [("plain_fp64", "fp64", None), ("fp32", "fp32", None)]
    runs += [(f"learned_fp32_s{k}", "fp32", k) for k in a.smoothing]
    for name, prec, smooth in runs:
        case = hetero_variant(-1)
        info = node_info(case.mesh, case.materials)
        s = cases.make_solver(case, precision=prec)
        pred = LearnedPredictor(net, info, smoothing=smooth) if smooth is not None else None
        res = cases.run_ssr(s, predictor=pred)
        used = sum(1 for r in res.step_info if r["predictor"] == "learned")
        gates = [(round(r["secant_first_residual"], 3), round(r["learned_first_residual"], 3))
                 for r in res.step_info if "secant_first_residual" in r]
        evals[name] = {"fos": res.fos, "wall": res.wall_time, "newton": res.newton_iterations,
                       "linear": res.linear_iterations, "attempts": res.attempts, "learned_steps_used": used,
                       "steps": len(res.lambda_hist), "gate_secant_vs_learned": gates}
        log("eval", name, evals[name])
        del s, res
    best = min((k for k in evals if k.startswith("learned")), key=lambda k: evals[k]["wall"], default=None)
    return {"case": "hetero3d_L1", "train_seeds": a.train, "samples": len(samples), "epochs": a.epochs,
            "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None,
            "data_gen_s": t_data, "train_s": t_train, "train_runs": train_info, "eval": evals,
            "best_learned": best,
            "speedup_learned_fp32_vs_plain_fp64": evals["plain_fp64"]["wall"] / evals[best]["wall"] if best else None,
            "newton_ratio_learned_vs_fp32": evals[best]["newton"] / max(1, evals["fp32"]["newton"]) if best else None}


def footing_variant(seed: int, h: float, spread: float = 0.2, dphi: float = 3.0) -> cases.Case:
    """The layered strip-footing case with its four layer materials perturbed per seed (c0 scaled
    by U(1 - spread, 1 + spread), phi shifted by U(-dphi, dphi) degrees); seed < 0: unperturbed."""
    base = cases.footing3d_hetero(h=h)
    if seed < 0:
        return base
    rng = np.random.default_rng(seed)
    mats = [MaterialSpec(c0=m.c0 * float(rng.uniform(1 - spread, 1 + spread)),
                         phi=m.phi + float(rng.uniform(-dphi, dphi)), psi=m.psi, young=m.young, poisson=m.poisson,
                         gamma_sat=m.gamma_sat, gamma_unsat=m.gamma_unsat) for m in base.materials]
    return cases.Case(f"footing3d_seed{seed}", base.mesh, mats, base.params, footing=base.footing,
                      gravity=base.gravity)


def footing_main(a, log):
    """Limit load of the layered strip footing: train on perturbed-material LL histories, evaluate
    plain FP64, FP32 and FP32 + learned predictor on the reference materials (time to the same
    omega, load factor there)."""
    t0 = time.perf_counter()
    samples, train_info = [], []
    for seed in range(a.train):
        case = footing_variant(seed, a.h, a.spread, a.dphi)
        res = cases.run_ll(cases.make_solver(case, precision="fp32"), store_step_u=True)
        info = node_info(case.mesh, case.materials)
        samples += samples_from_history(info, res.step_U, res.omega_hist, res.t_hist)
        train_info.append({"seed": seed, "t": res.limit_load, "newton": res.newton_iterations,
                           "steps": len(res.t_hist), "wall": res.wall_time})
        log("train run", train_info[-1])
        del res
        torch.cuda.empty_cache()
    t_data = time.perf_counter() - t0
    t1 = time.perf_counter()
    net = NodeMLP(seed=0)
    losses = train(net, samples, epochs=a.epochs, lr=2e-3, log=lambda ep, l: log("epoch", ep, "loss", l))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    warm = footing_variant(-1, a.h)
    warm.params.step_max = 4
    cases.run_ll(cases.make_solver(warm, precision="fp32"))
    evals = {}
    eh = a.eval_h or a.h
    for name, prec, learned in (("plain_fp64", "fp64", False), ("fp32", "fp32", False),
                                ("learned_fp32", "fp32", True)):
        case = footing_variant(-1, eh)
        if a.eval_omega:
            case.params.omega_max_stop = a.eval_omega
        pred = LearnedPredictor(net, node_info(case.mesh, case.materials), smoothing=2) if learned else None
        res = cases.run_ll(cases.make_solver(case, precision=prec), predictor=pred,
                           progress=lambda ev: log("step", name, ev.get("t"), ev.get("newton_it")))
        evals[name] = {"t": res.limit_load, "omega": float(res.omega_hist[-1]), "stop": res.stop_reason,
                       "wall": res.wall_time, "newton": res.newton_iterations, "linear": res.linear_iterations,
                       "learned_steps_used": sum(1 for r in res.step_info if r["predictor"] == "learned"),
                       "steps": len(res.t_hist)}
        log("eval", name, evals[name])
    return {"case": f"footing3d_hetero train h={a.h}, eval h={eh}", "train_seeds": a.train, "samples": len(samples),
            "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None,
            "data_gen_s": t_data, "train_s": t_train, "train_runs": train_info, "eval": evals,
            "speedup_learned_fp32_vs_plain_fp64": evals["plain_fp64"]["wall"] / evals["learned_fp32"]["wall"]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="homo", choices=["homo", "hetero3d_L1", "footing3d"])
    ap.add_argument("--save", default=None, help="write the trained weights to this .npz")
    ap.add_argument("--load", default=None, help="skip training: evaluate these weights")
    ap.add_argument("--train-smoothing", type=int, default=0,
                    help="train on the prediction smoothed by this many passes (hetero3d_L1)")
    ap.add_argument("--smoothing", type=int, nargs="+", default=[0, 3, 8],
                    help="smoothing passes of the learned correction to evaluate")
    ap.add_argument("--train", type=int, default=12)
    ap.add_argument("--test", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--spread", type=float, default=0.2, help="hetero3d_L1 training: relative spread of c0")
    ap.add_argument("--dphi", type=float, default=3.0, help="hetero3d_L1 training: spread of phi (degrees)")
    ap.add_argument("--h", type=float, default=2.0)
    ap.add_argument("--eval-h", type=float, default=None, help="footing3d: evaluate on this mesh size (default --h)")
    ap.add_argument("--eval-omega", type=float, default=None, help="footing3d: evaluation runs stop at this omega")
    ap.add_argument("--nz", type=int, default=6)
    a = ap.parse_args(argv)
    log = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    if a.case == "hetero3d_L1":
        print(json.dumps(hetero_main(a, log)), flush=True)
        return
    if a.case == "footing3d":
        print(json.dumps(footing_main(a, log)), flush=True)
        return
    t0 = time.perf_counter()
    samples = []
    train_info = []
    for seed in range(a.train):
        case = slope_case(seed, a.h, a.nz)
        s, res = run_case(case, store=True)
        info = node_info(case.mesh, case.materials)
        samples += samples_from_history(info, res.step_U, res.omega_hist, res.lambda_hist)
        train_info.append({"seed": seed, "fos": res.fos, "newton": res.newton_iterations, "steps": len(res.lambda_hist)})
        log("train run", train_info[-1])
        del s, res
        torch.cuda.empty_cache()
    t_data = time.perf_counter() - t0
    t1 = time.perf_counter()
    if a.load:
        net = NodeMLP.load(a.load)
        losses = []
    else:
        net = NodeMLP(seed=0)
        info0 = node_info(hetero_variant(-1).mesh, cases.HETERO_L1_MATERIALS)
        losses = train(net, samples, epochs=a.epochs, lr=2e-3, log=lambda ep, l: log("epoch", ep, "loss", l),
                       info=info0 if a.train_smoothing else None, smoothing=a.train_smoothing)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    evals = []
    for seed in range(1000, 1000 + a.test):
        case = slope_case(seed, a.h, a.nz)
        _, base = run_case(case)
        info = node_info(case.mesh, case.materials)
        _, lrn = run_case(case, predictor=lambda c: LearnedPredictor(net, info))
        used = sum(1 for r in lrn.step_info if r["predictor"] == "learned")
        evals.append({"seed": seed, "fos_plain": base.fos, "fos_learned": lrn.fos,
                      "fos_abs_diff": abs(base.fos - lrn.fos), "newton_plain": base.newton_iterations,
                      "newton_learned": lrn.newton_iterations, "wall_plain": base.wall_time,
                      "wall_learned": lrn.wall_time, "learned_steps_used": used,
                      "attempts_plain": base.attempts, "attempts_learned": lrn.attempts, **case.extra})
        log("eval", evals[-1])
    out = {"train_seeds": a.train, "test_seeds": a.test, "samples": len(samples), "epochs": a.epochs,
           "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None,
           "data_gen_s": t_data, "train_s": t_train, "train_runs": train_info, "eval": evals,
           "newton_ratio": float(np.sum([e["newton_learned"] for e in evals]) /
                                 max(1, np.sum([e["newton_plain"] for e in evals]))),
           "max_fos_diff": float(max(e["fos_abs_diff"] for e in evals)) if evals else None}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
