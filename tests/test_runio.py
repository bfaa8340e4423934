"""Run artefacts in the reference's formats (petsc_run.npz, run_info.json): written from a
continuation result and read back, histories and accepted states unchanged (host tensors here;
the same code runs on device tensors)."""

import json
from types import SimpleNamespace

import numpy as np
import torch

from slope_stability_petsc4py_amd import cases
from slope_stability_petsc4py_amd.runio import read_run, step_states, write_run


def test_write_read_round_trip(tmp_path):
    case = cases.homo2d(h=5.0)
    m = case.mesh
    n = 2 * m.n_nodes
    rng = np.random.default_rng(0)
    steps = [torch.from_numpy(rng.normal(size=n)) for _ in range(4)]
    res = SimpleNamespace(U=steps[-1], lambda_hist=np.array([0.9, 1.0, 1.1, 1.15]),
                          omega_hist=np.array([1.0, 2.0, 3.0, 4.0]), Umax_hist=np.array([0.1, 0.2, 0.3, 0.4]),
                          newton_iterations=17, linear_iterations=99, wall_time=1.5, step_U=steps,
                          step_info=[{"predictor": "secant"}, {"predictor": "learned"},
                                     {"predictor": "learned_gated"}])
    solver = SimpleNamespace(op=SimpleNamespace(get_total_time=lambda: {"stress": 0.1}),
                             linear=SimpleNamespace(iteration_collector=SimpleNamespace(summary=lambda: {"its": 99})))
    out = write_run(tmp_path / "run", case, solver, res, extra={"seed": 3})
    run = read_run(out)
    d = np.load(out / "petsc_run.npz")
    assert {"U", "lambda_hist", "load_factor_hist", "omega_hist", "Umax_hist", "step_U", "init_meta",
            "stats_newton_iterations", "stats_linear_iterations", "stats_total_wall_time"} <= set(d.files)
    assert run["U"].shape == (2, m.n_nodes)
    np.testing.assert_array_equal(run["U"][:, 5], steps[-1].numpy()[10:12])
    np.testing.assert_array_equal(run["lambda_hist"], res.lambda_hist)
    np.testing.assert_array_equal(run["load_factor_hist"], res.lambda_hist)
    back = step_states(run, device=torch.device("cpu"))
    assert len(back) == 4 and all(torch.equal(a, b) for a, b in zip(back, steps))
    info = json.loads((out / "run_info.json").read_text())
    assert info["run_info"]["step_count"] == 4 and info["run_info"]["mesh_nodes"] == m.n_nodes
    assert info["predictor_diagnostics"] == {"learned_steps": 1, "learned_gated_steps": 1, "attempts": 3}
    assert info["params"]["lambda_init"] == case.params.lambda_init and info["extra"] == {"seed": 3}
