"""The learned continuation predictor on seeded slopes: data, training and evaluation.

``--case homo`` (the learned-predictor configuration): 3D homogeneous slopes drawn per seed --
height U(8, 15) m, face angle U(30, 60) deg, c U(5, 20) kPa, phi U(25, 40) deg, E = 40 MPa,
nu = 0.3, gamma = 20 kN/m3 -- meshed with P2 tetrahedra (h = 0.9 m: 35k-54k tets, ~40k on average).
Each training seed runs the plain SSR continuation and keeps its accepted states; the per-node MLP
(4 layers x 128, bf16 matrix-core GEMMs with FP32 accumulation, trained on the device) learns the
next increment from them; the held-out seeds (from 1000 on) run with the plain secant predictor and
with the learned one, and the Newton iterations, wall times and FoS differences are reported.
With ``--out DIR`` every training run is also written in the reference's run formats
(DIR/seed_<k>/petsc_run.npz with step_U, run_info.json), and ``--from-runs DIR`` trains on such
runs instead of running them again.

``--case hetero3d_L1``: the reference's 3D heterogeneous P2 slope with its four layer materials
perturbed per seed (c0 x U(0.8, 1.2), phi + U(-3, 3) deg); ``--save`` writes the weights the
benchmark's learned form loads (slope_stability_petsc4py_amd/data/predictor_hetero3d_L1.npz).

``python -m slope_stability_petsc4py_amd.learned_slopes --case homo --train 200 --test 20``
prints one JSON line (progress on stderr).
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch

from . import cases, mesh
from .learn import LearnedPredictor, NodeMLP, node_info, samples_from_history, train
from .materials import MaterialSpec

HOMO_H = 0.9


def slope_case(seed: int, h: float = HOMO_H, nz: int | None = None) -> cases.Case:
    """Seeded homogeneous slope: 6 m of foundation below the toe, 10 m benches at crest and toe,
    12 m out of plane."""
    rng = np.random.default_rng(seed)
    height = float(rng.uniform(8.0, 15.0))
    beta = float(rng.uniform(30.0, 60.0))
    c0 = float(rng.uniform(5.0, 20.0))
    phi = float(rng.uniform(25.0, 40.0))
    run = height / np.tan(np.deg2rad(beta))
    m = mesh.slope_3d(h=h, x1=10.0, x2=run, x3=10.0, y1=6.0, y2=height, z_len=12.0, nz=nz)
    mat = MaterialSpec(c0=c0, phi=phi, psi=0.0, young=40000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0)
    p = cases.SSRParams(lambda_init=0.7, d_lambda_init=0.1, d_lambda_min=1e-3, d_lambda_diff_scaled_min=1e-3,
                        step_max=40, omega_max_stop=1e12, it_newt_max=60)
    c = cases.Case(f"slope_seed{seed}", m, [mat], p)
    c.extra = {"height": height, "beta": beta, "c0": c0, "phi": phi}
    return c


def hetero_variant(seed: int, spread: float = 0.2, dphi: float = 3.0) -> cases.Case:
    """The heterogeneous L1 slope with its layer materials perturbed per seed; seed < 0 gives the
    reference materials."""
    base = cases.hetero3d_L1()
    if seed < 0:
        return base
    rng = np.random.default_rng(seed)
    mats = [MaterialSpec(c0=m.c0 * float(rng.uniform(1 - spread, 1 + spread)),
                         phi=m.phi + float(rng.uniform(-dphi, dphi)), psi=m.psi, young=m.young, poisson=m.poisson,
                         gamma_sat=m.gamma_sat, gamma_unsat=m.gamma_unsat) for m in cases.HETERO_L1_MATERIALS]
    return cases.Case(f"hetero3d_L1_seed{seed}", base.mesh, mats, base.params)


def collect(case_of, seeds, *, precision: str = "fp64", out_dir=None, log=print):
    """Run the plain SSR continuation of every seed with its accepted states kept; returns the
    training samples and one record per run."""
    samples, runs = [], []
    for seed in seeds:
        case = case_of(seed)
        t0 = time.perf_counter()
        s = cases.make_solver(case, precision=precision)
        res = cases.run_ssr(s, store_step_u=True)
        info = node_info(case.mesh, case.materials)
        samples += samples_from_history(info, res.step_U, res.omega_hist, res.lambda_hist)
        rec = {"seed": seed, "n_elem": int(case.mesh.n_elem), "fos": res.fos, "newton": res.newton_iterations,
               "steps": len(res.lambda_hist), "wall": res.wall_time, "total": time.perf_counter() - t0}
        if out_dir is not None:
            from .runio import write_run
            write_run(Path(out_dir) / f"seed_{seed}", case, s, res, precision=precision, extra=case.extra)
        runs.append(rec)
        log("train run", json.dumps(rec))
        del s, res
        torch.cuda.empty_cache()
    return samples, runs


def samples_from_runs(case_of, run_root, seeds):
    """Training samples from runs written by `collect(out_dir=...)` (no continuation is run); seeds
    without a run are skipped."""
    from .runio import read_run, step_states
    samples = []
    for seed in seeds:
        if not (Path(run_root) / f"seed_{seed}" / "petsc_run.npz").exists():
            print(f"seed {seed}: no run in {run_root}, skipped", file=sys.stderr, flush=True)
            continue
        run = read_run(Path(run_root) / f"seed_{seed}")
        case = case_of(seed)
        info = node_info(case.mesh, case.materials)
        samples += samples_from_history(info, step_states(run), run["omega_hist"], run["lambda_hist"])
    return samples


def evaluate(net, case_of, seeds, *, precision: str = "fp64", log=print):
    """Each held-out seed with the secant predictor (plain method) and with the learned one."""
    evals = []
    for seed in seeds:
        case = case_of(seed)
        base = cases.run_ssr(cases.make_solver(case, precision=precision))
        pred = LearnedPredictor(net, node_info(case.mesh, case.materials))
        lrn = cases.run_ssr(cases.make_solver(case, precision=precision), predictor=pred)
        rec = {"seed": seed, "n_elem": int(case.mesh.n_elem), "fos_plain": base.fos, "fos_learned": lrn.fos,
               "fos_abs_diff": abs(base.fos - lrn.fos), "newton_plain": base.newton_iterations,
               "newton_learned": lrn.newton_iterations, "wall_plain": base.wall_time, "wall_learned": lrn.wall_time,
               "learned_steps_used": sum(1 for r in lrn.step_info if r["predictor"] == "learned"),
               "attempts_plain": base.attempts, "attempts_learned": lrn.attempts, **case.extra}
        evals.append(rec)
        log("eval", json.dumps(rec))
        torch.cuda.empty_cache()
    return evals


def summary(evals) -> dict:
    npl = sum(e["newton_plain"] for e in evals)
    nle = sum(e["newton_learned"] for e in evals)
    wpl = sum(e["wall_plain"] for e in evals)
    wle = sum(e["wall_learned"] for e in evals)
    return {"newton_ratio": nle / max(npl, 1), "wall_ratio": wle / max(wpl, 1e-30),
            "max_fos_diff": max((e["fos_abs_diff"] for e in evals), default=None),
            "seeds_with_learned_steps": sum(1 for e in evals if e["learned_steps_used"] > 0)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="homo", choices=["homo", "hetero3d_L1"])
    ap.add_argument("--train", type=int, default=200)
    ap.add_argument("--seed-start", type=int, default=0,
                    help="first training seed (a shard of the training set: seeds seed-start .. seed-start+train-1)")
    ap.add_argument("--seeds", default=None, help="explicit training seeds, comma separated (instead of the range)")
    ap.add_argument("--init", default=None,
                    help="continue training from these weights (training over shards, one call per shard)")
    ap.add_argument("--test", type=int, default=20)
    ap.add_argument("--test-start", type=int, default=1000, help="first held-out seed")
    ap.add_argument("--h", type=float, default=HOMO_H)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--out", default=None, help="write every training run (petsc_run.npz, run_info.json) here")
    ap.add_argument("--from-runs", default=None, help="train on runs written earlier with --out")
    ap.add_argument("--save", default=None, help="write the trained weights (.npz)")
    ap.add_argument("--load", default=None, help="evaluate these weights instead of training")
    a = ap.parse_args(argv)
    log = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    if a.case == "homo":
        case_of = lambda seed: slope_case(seed, a.h)  # noqa: E731
        test_seeds = range(a.test_start, a.test_start + a.test)
    else:
        case_of = hetero_variant
        test_seeds = [-1]
    t0 = time.perf_counter()
    runs = []
    train_seeds = ([int(x) for x in a.seeds.split(",") if x] if a.seeds else
                   list(range(a.seed_start, a.seed_start + a.train)))
    if a.load:
        samples = []
    elif a.from_runs:
        samples = samples_from_runs(case_of, a.from_runs, train_seeds)
    else:
        samples, runs = collect(case_of, train_seeds, precision=a.precision, out_dir=a.out, log=log)
    t_data = time.perf_counter() - t0
    t1 = time.perf_counter()
    if a.load:
        net, losses = NodeMLP.load(a.load), []
    else:
        net = NodeMLP.load(a.init) if a.init else NodeMLP(seed=0)
        losses = train(net, samples, epochs=a.epochs, lr=a.lr, log=lambda ep, l: log("epoch", ep, "loss", l))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    if a.save:
        net.save(a.save)
    evals = evaluate(net, case_of, test_seeds, precision=a.precision, log=log) if a.test > 0 else []
    out = {"case": a.case, "h": a.h, "precision": a.precision, "train_seeds": train_seeds, "test_seeds": len(evals),
           "samples": len(samples),
           "epochs": a.epochs, "loss_first": losses[0] if losses else None,
           "loss_last": losses[-1] if losses else None, "data_gen_s": t_data, "train_s": t_train,
           "train_runs": runs, "eval": evals, **(summary(evals) if evals else {})}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
