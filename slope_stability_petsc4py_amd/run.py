"""Command-line runner: ``python -m slope_stability_petsc4py_amd.run --case homo2d --precision fp64``.

Runs one SSR (or LL) continuation on the device and prints one JSON line with the factor of
safety (or limit load), the histories, wall time and iteration counts; with ``--out DIR`` it also
writes the reference's run artefacts (petsc_run.npz, run_info.json; ``--store-steps`` adds the
accepted states step_U).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import cases


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="homo2d", choices=["homo2d", "hetero3d_L1", "hetero3d_L2_layers",
                                                         "slope3d_homo", "slope3d_hetero",
                                                         "homo2d_ll", "homo3d_ll_L1", "hetero3d_ll_L1",
                                                         "footing2d_prandtl", "footing3d_hetero"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--pc", default="pmg", choices=["pmg", "cheb"])
    ap.add_argument("--analysis", default="ssr", choices=["ssr", "ll"])
    ap.add_argument("--elem-type", default=None, choices=["P1", "P2", "P4"],
                    help="element order of the Gmsh-mesh cases (default P2; P4 for hetero3d_L2_layers); "
                         "hetero3d_L1 with P4 is the reference's 616,322-dof P4 problem")
    ap.add_argument("--h", type=float, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--step-max", type=int, default=None)
    ap.add_argument("--omega-max", type=float, default=None, help="override omega_max_stop")
    ap.add_argument("--coarse-refresh", type=int, default=0,
                    help="rebuild the multigrid coarse operator every N preconditioner setups (0: when a solve "
                         "needs more than --refresh-iterations iterations)")
    ap.add_argument("--refresh-iterations", type=int, default=20,
                    help="FGMRES iterations of one solve above which the coarse operator is rebuilt")
    ap.add_argument("--coarse-precision", default=None, choices=["fp32", "fp64"],
                    help="coarse factorisation precision (default: the preconditioner precision)")
    ap.add_argument("--coarse", default="auto", choices=["auto", "dense", "aggregate"],
                    help="p-multigrid coarse level: dense P1, or sparse P1 above a dense aggregate level "
                         "(auto: aggregate above 50000 free P1 dofs)")
    ap.add_argument("--out", default=None, help="write petsc_run.npz and run_info.json into this directory")
    ap.add_argument("--store-steps", action="store_true", help="keep the accepted states (step_U) in --out")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    if a.elem_type is None and a.case != "hetero3d_L2_layers":
        a.elem_type = "P2"
    if a.case.endswith("_ll") or a.case.endswith("_ll_L1") or a.case.startswith("footing"):
        a.analysis = "ll"
    if a.case == "homo2d":
        case = cases.homo2d(h=a.h or 1.0)
    elif a.case == "homo2d_ll":
        case = cases.homo2d_ll(h=a.h or 0.5)
    elif a.case == "homo3d_ll_L1":
        case = cases.homo3d_ll_L1(a.elem_type)
    elif a.case == "hetero3d_ll_L1":
        case = cases.hetero3d_ll_L1(a.elem_type)
    elif a.case == "hetero3d_L1":
        case = cases.hetero3d_L1(a.elem_type)
    elif a.case == "hetero3d_L2_layers":
        case = cases.hetero3d_L2_layers(a.elem_type or "P4")
    elif a.case == "footing2d_prandtl":
        case = cases.footing2d_prandtl(h=a.h or 0.25)
    elif a.case == "footing3d_hetero":
        case = cases.footing3d_hetero(h=a.h or 2.0)
    else:
        case = cases.slope3d(h=a.h or 2.0, nz=a.nz or 10, hetero=a.case.endswith("hetero"))
    if a.step_max:
        case.params.step_max = a.step_max
    if a.omega_max:
        case.params.omega_max_stop = a.omega_max
    s = cases.make_solver(case, precision=a.precision, pc=a.pc, coarse_refresh=a.coarse_refresh,
                          coarse_precision=a.coarse_precision, refresh_iterations=a.refresh_iterations,
                          coarse=a.coarse)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0

    def prog(ev):
        if a.verbose:
            print(json.dumps({k: (float(v) if isinstance(v, (np.floating, float)) else v) for k, v in ev.items()}),
                  file=sys.stderr, flush=True)

    pc = s.linear.pc
    inner = pc.coarse if hasattr(getattr(pc, "coarse", None), "n_d") else pc   # P4: its P2 multigrid
    coarse_kind = getattr(inner, "coarse", None)
    n_coarse = getattr(inner, "n_d", None)
    if a.analysis == "ssr":
        res = cases.run_ssr(s, progress=prog, store_step_u=a.store_steps)
        out = {"case": case.name, "precision": a.precision, "pc": a.pc, "fos": res.fos,
               "lambda_hist": res.lambda_hist.tolist(), "omega_hist": res.omega_hist.tolist(),
               "Umax_hist": res.Umax_hist.tolist(), "stop_reason": res.stop_reason,
               "wall_time_s": res.wall_time, "setup_s": t_setup, "newton_iterations": res.newton_iterations,
               "linear_iterations": res.linear_iterations, "attempts": res.attempts,
               "n_dof": int(s.op.pattern.n_rows), "n_elem": int(case.mesh.n_elem),
               "coarse_refresh": a.coarse_refresh, "linear_solver": s.linear.iteration_collector.summary(),
               "pc_timings": getattr(s.linear.pc, "timings", None),
               "coarse": coarse_kind, "n_coarse_dense": n_coarse, "timings": s.op.get_total_time()}
    else:
        res = cases.run_ll(s, progress=prog, store_step_u=a.store_steps)
        out = {"case": case.name, "precision": a.precision, "limit_load": res.limit_load,
               "t_hist": res.t_hist.tolist(), "omega_hist": res.omega_hist.tolist(),
               "Umax_hist": res.Umax_hist.tolist(), "wall_time_s": res.wall_time, "setup_s": t_setup,
               "stop_reason": res.stop_reason, "newton_iterations": res.newton_iterations,
               "linear_iterations": res.linear_iterations, "n_dof": int(s.op.pattern.n_rows),
               "n_elem": int(case.mesh.n_elem), "linear_solver": s.linear.iteration_collector.summary(),
               "pc_timings": getattr(s.linear.pc, "timings", None), "extra": case.extra,
               "coarse": coarse_kind, "n_coarse_dense": n_coarse}
    if a.out:
        from .runio import write_run
        write_run(a.out, case, s, res, analysis=a.analysis, precision=a.precision, setup_s=t_setup)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
