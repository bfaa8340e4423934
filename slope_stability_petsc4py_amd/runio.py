"""Run artefacts in the reference's formats: ``petsc_run.npz`` and ``run_info.json``.

The reference writes them at the end of a capture run (cli/run_3D_hetero_SSR_capture.py:1492-1506,
the run payload built just above it): the final displacement ``U`` (dim, n_nodes), the histories
``lambda_hist`` (also as ``load_factor_hist``), ``omega_hist``, ``Umax_hist``, the accepted states
``step_U`` when they were stored (continuation/indirect.py:2669-2671), ``init_meta`` and
``stats_*`` fields; ``run_info.json`` holds ``run_info``, ``params``, ``mesh``, ``timings`` and
``predictor_diagnostics``.  The same keys are written here (the reference's BDDC / owned-pattern
sections describe its MPI set-up and are left out), so a history captured on the device can be
read back offline, e.g. as training data of the learned predictor (`read_run`).
"""

from __future__ import annotations

import json
import os
import platform
import time
from pathlib import Path

import numpy as np
import torch


def _field(U: torch.Tensor | np.ndarray, dim: int) -> np.ndarray:
    """Flat dof vector (dof = dim*node + comp) -> (dim, n_nodes)."""
    u = U.detach().cpu().numpy() if isinstance(U, torch.Tensor) else np.asarray(U)
    return np.ascontiguousarray(u.reshape(-1, dim).T)


def write_run(out_dir, case, solver, res, *, analysis: str = "ssr", precision: str = "fp64",
              setup_s: float | None = None, extra: dict | None = None) -> Path:
    """Write petsc_run.npz and run_info.json of one continuation run into `out_dir`."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    dim = case.mesh.dim
    hist = np.asarray(res.lambda_hist if analysis == "ssr" else res.t_hist, dtype=np.float64)
    arrays = {
        "U": _field(res.U, dim),
        "lambda_hist": hist,
        "load_factor_hist": hist,
        "omega_hist": np.asarray(res.omega_hist, dtype=np.float64),
        "Umax_hist": np.asarray(res.Umax_hist, dtype=np.float64),
        "init_meta": np.array([1]),
        "stats_newton_iterations": np.array([res.newton_iterations]),
        "stats_linear_iterations": np.array([res.linear_iterations]),
        "stats_total_wall_time": np.array([res.wall_time]),
    }
    if res.step_U:
        arrays["step_U"] = np.stack([_field(u, dim) for u in res.step_U])
    # written under a temporary name and renamed, so an interrupted writer leaves no partial file
    tmp = out / "petsc_run.npz.part"
    with open(tmp, "wb") as fh:
        np.savez_compressed(fh, **arrays)
    os.replace(tmp, out / "petsc_run.npz")
    p = case.params
    learned = sum(1 for r in res.step_info if r.get("predictor") == "learned")
    gated = sum(1 for r in res.step_info if r.get("predictor") == "learned_gated")
    payload = {
        "run_info": {
            "timestamp": time.strftime("%Y-%m-%dT%H:%M:%S"),
            "python_version": platform.python_version(),
            "runtime_seconds": float(res.wall_time + (setup_s or 0.0)),
            "mpi_size": 1,
            "mesh_nodes": int(case.mesh.n_nodes),
            "mesh_elements": int(case.mesh.n_elem),
            "unknowns": int(case.mesh.q_mask.sum()),
            "analysis": analysis,
            "solver_type": f"DFGMRES + p-multigrid ({precision} preconditioner), one MI355X",
            "step_count": int(hist.size),
        },
        "params": {
            "analysis": analysis,
            **{k: getattr(p, k) for k in ("lambda_init", "d_lambda_init", "d_lambda_min", "d_lambda_diff_scaled_min",
                                          "step_max", "omega_max_stop", "it_newt_max", "it_damp_max", "tol", "r_min")},
            "elem_type": case.mesh.elem_type,
            "davis_type": case.davis,
            "material_rows": [[m.c0, m.phi, m.psi, m.young, m.poisson, m.gamma_sat, m.gamma_unsat]
                              for m in case.materials],
        },
        "mesh": {"mesh_file": case.name, "coord_shape": list(case.mesh.coord.shape),
                 "elem_shape": list(case.mesh.elem.shape), "surf_shape": list(case.mesh.surf.shape)},
        "timings": {
            "constitutive": solver.op.get_total_time(),
            "linear": solver.linear.iteration_collector.summary(),
            "continuation_total_wall_time": float(res.wall_time),
        },
        "predictor_diagnostics": {"learned_steps": learned, "learned_gated_steps": gated,
                                  "attempts": len(res.step_info)},
    }
    if extra:
        payload["extra"] = extra
    (out / "run_info.json").write_text(json.dumps(payload, indent=2, default=float))
    return out


def read_run(run_dir) -> dict:
    """Histories and accepted states of a run directory (petsc_run.npz, with or without step_U)."""
    d = np.load(Path(run_dir) / "petsc_run.npz", allow_pickle=False)
    out = {k: d[k] for k in d.files}
    info = Path(run_dir) / "run_info.json"
    if info.exists():
        out["run_info"] = json.loads(info.read_text())
    return out


def step_states(run: dict, device=None) -> list[torch.Tensor]:
    """The stored accepted states of a run as flat device dof vectors (dof = dim*node + comp)."""
    dev = device or torch.device("cuda", 0)
    return [torch.from_numpy(np.ascontiguousarray(s.T).reshape(-1)).to(dev) for s in run["step_U"]]
