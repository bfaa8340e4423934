"""Benchmark of the hot path on one MI355X.

Default workload (the kernel benchmark configuration): a 3D homogeneous P2 slope (10 m high, 45
deg face, 40 m out of plane, ~117k 10-node tetrahedra, 11-point rule, ~1.29M points, ~0.5M dofs).
Inputs resident in HBM: branch-stratified strains (70% elastic, 30% plastic over smooth face,
left/right edge and apex) at every point, the geometry (dphi, w) and the fixed CSR pattern.

One step = the Newton-iteration hot path on those inputs:
  return mapping (E -> S, DS)  +  internal force (S -> F)  +  tangent assembly into the fixed
  pattern with the regularised operator K_r = (1-r)K + r K_el and its FP32 copy.
`value` = algorithmic HBM bytes of the step / step time (GB/s).  `roofline` is the slowest kernel
of the step (the tangent) against the 8 TB/s HBM peak, timed with HIP events on its launch stream,
with the HBM bytes per launch read by rocprofv3 (profiles/traffic.json).  `cpu_baseline` is the
reference's own C kernels (built from its sources into oracle/_ref, OpenMP on the host cores) on
the same arrays.  `p4` is the matrix-core leg: the heterogeneous P4 slope (33,935 tets, 1.14M dofs),
return mapping + FP64 MFMA element matrices + node gather, with its own `roofline` (MFMA).  `ssr` is
the end-to-end leg of the metric: the reference's 3D heterogeneous P2 SSR to the factor of safety,
plain FP64 against the FP32 and learned + FP32 forms; `ssr.ssr2d` is the reference's 2D
homogeneous SSR on the GPU beside its CPU restatement (oracle/ssr2d.py, one host core).

Prints one JSON line.  Rank/world handling: one GPU only (--gpus > 1 exits with a message).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SSR wall time to FoS (learned+FP32 vs plain GPU) at 1e-3; assembly HBM GB/s"
HBM_PEAK_GBS = 8000.0


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(h: float, nz: int, n_unique: int, seed: int):
    import torch

    from slope_stability_petsc4py_amd import mesh, workloads
    from slope_stability_petsc4py_amd.constitutive import ConstitutiveOperator

    t0 = time.time()
    m = mesh.slope_3d(h=h, nz=nz, x1=15.0, x2=10.0, x3=15.0, y1=10.0, y2=10.0, z_len=40.0)
    op = ConstitutiveOperator.from_mesh(m, [workloads.HOMO], Davis_type="B")
    _log(f"mesh+pattern {time.time() - t0:.1f}s: {m.n_elem} tets, {op.n_int} points, {op.pattern.n_rows} dofs, "
         f"nnz {op.pattern.host.nnz}")
    E_u, br_u = workloads.stratified_strains(n_unique, seed=seed)
    reps = -(-op.n_int // n_unique)
    E = np.ascontiguousarray(np.tile(E_u, (reps, 1))[: op.n_int])
    br = np.tile(br_u, reps)[: op.n_int]
    op.reduction(1.0)
    op._E.copy_(torch.from_numpy(E))
    torch.cuda.synchronize()
    return m, op, E, br


def algorithmic_bytes(op, fold: bool = True, branch: bool = False) -> dict:
    """Per-launch algorithmic HBM bytes of each kernel of the step (`branch`: the return mapping
    writes its int8 branch id per point and the tangent reads it)."""
    n = op.n_int
    pat = op.pattern
    nnz = pat.host.nnz
    n_p = pat.n_p
    nb = 1 if branch else 0
    return {
        # E 48 + (c_bar, sin_phi, shear, lame) 32 in; S 48 + DS 288 (+ branch id) out
        "constitutive": n * (48 + 32 + 48 + 288 + nb),
        # S 48 + dphi 3*n_p*8 + w 8 per point in; fe out; then fe in + F out
        "force": n * (48 + 24 * n_p + 8) + pat.n_elem * pat.nld * 8 * 2 + pat.n_rows * 8,
        # DS 288 + dphi 3*n_p*8 + w 8 (+ shear, lame 16 when K_r is folded into the point tangents,
        # else the assembled K_el nnz*8) in; K_r nnz*8 + FP32 copy nnz*4 out
        "tangent": n * (288 + 24 * n_p + 8 + nb) + nnz * (8 + 4) + (n * 16 if fold else nnz * 8),
        # dphi + u gather (3*n_p*8 per element, counted once) + E out
        "strain": n * (24 * n_p + 48) + pat.n_elem * pat.nld * 8,
        # CSR FP64: values 8 + index 4 per nnz, x 8 (once) + y 8 + indptr 4 per row
        "spmv": nnz * 12 + pat.n_rows * 20,
    }


def moved_bytes(op, fold: bool = True, branch: bool = False, sym: bool = False) -> dict:
    """The same kernels' input and output bytes as this build lays them out.  They differ from the
    algorithmic bytes where the mesh takes the constant-Jacobian form (P2 tetrahedra with straight
    edges): the strain, force and tangent kernels then read 10 doubles per element (inverse
    Jacobian and |det J|) in place of the per-point dphi (3 n_p doubles) and w, and (`sym`) the
    return mapping writes, and the tangent reads, the symmetric DS as 21 values per point, not 36."""
    ab = algorithmic_bytes(op, fold, branch)
    pat = op.pattern
    if not pat.affine:
        return ab
    n = op.n_int
    per_point = n * (24 * pat.n_p + 8)            # dphi + w
    geo = pat.n_elem * 80
    mb = dict(ab)
    for k in ("force", "tangent"):
        mb[k] = ab[k] - per_point + geo
    mb["strain"] = ab["strain"] - n * 24 * pat.n_p + geo
    if sym:
        mb["constitutive"] -= n * 15 * 8
        mb["tangent"] -= n * 15 * 8
    return mb


def run_gpu(args):
    import torch

    from slope_stability_petsc4py_amd import _lib
    from slope_stability_petsc4py_amd.device import ptr, stream_ptr

    m, op, E_host, br = build_workload(args.h, args.nz, args.unique, args.seed)
    pat = op.pattern
    pat.tangent_method = args.tangent
    pat.geometry_form = args.geometry
    dev = op._E.device
    # P2 tetrahedra in the constant-Jacobian form: the return mapping writes the symmetric tangent
    # packed (21 values per point) and the tangent reads it so
    sym = bool(op._sym and pat.affine and args.tangent == "nodes")
    S = op._S
    DS = op._DSp if sym else (op._DS if op._DS is not None else
                              torch.empty((op.n_int, 36), dtype=torch.float64, device=dev))
    K = op.pattern.new_csr(with32=True)
    fe, F = op._fe, op._F
    st = stream_ptr()
    r = 1e-4

    # branch ids of the return mapping go with DS into the tangent (elastic points in closed form)
    br = op._branch if args.tangent == "nodes" else None

    def k_const():
        if sym:
            _lib.call("ssx_constitutive_3d_sym", ptr(op._E), ptr(op.c_bar), ptr(op.sin_phi), ptr(op.shear),
                      ptr(op.lame), op.n_int, ptr(S), ptr(DS), ptr(br), st)
        else:
            _lib.call("ssx_constitutive_3d", ptr(op._E), ptr(op.c_bar), ptr(op.sin_phi), ptr(op.shear),
                      ptr(op.lame), op.n_int, ptr(S), ptr(DS), None, ptr(br), st)

    def k_force():
        pat.force(S, fe, F)

    mat = (op.shear, op.lame) if (pat.ns is not None and args.tangent == "nodes") else None
    el = None if mat is not None else op.elastic_values

    def k_tangent():
        pat.tangent(DS, K.values, K.values32, r=r, elastic=el, mat=mat, branch=br if mat is not None else None,
                    sym=sym)

    u = torch.from_numpy(np.random.default_rng(1).normal(0, 1e-3, pat.n_rows)).to(dev)
    E_tmp = torch.empty_like(op._E)
    y = torch.empty(pat.n_rows, dtype=torch.float64, device=dev)

    def k_strain():
        pat.strain(u, E_tmp)

    def k_spmv():
        K.matvec(u, y)

    step = [("constitutive", k_const), ("force", k_force), ("tangent", k_tangent)]
    extra = [("strain", k_strain), ("spmv", k_spmv)]
    for _ in range(args.warmup):
        for _, f in step + extra:
            f()
    torch.cuda.synchronize()

    # timed region: exactly K steps
    ev = {name: [] for name, _ in step + extra}
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record()
    for _ in range(args.steps):
        for name, f in step:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            f()
            b.record()
            ev[name].append((a, b))
    t_end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    step_ms = t_start.elapsed_time(t_end) / args.steps
    # side kernels (not part of the step), timed separately
    for _ in range(args.steps):
        for name, f in extra:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            f()
            b.record()
            ev[name].append((a, b))
    torch.cuda.synchronize()
    if args.copy_probe:
        # counter calibration: a plain device copy of a known size (256 MiB read + 256 MiB written)
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        # second and third calibrations, other access kinds: a read-only reduction of the same
        # 256 MiB, and a gather of 128-byte rows in random order (whole cache lines, as the node
        # gathers read them; 128 MiB read, 128 MiB written)
        for _ in range(3):
            torch.sum(src, dtype=torch.float64)
        rows = src.view(-1, 16)
        idx = torch.randperm(rows.shape[0], device=dev)[: rows.shape[0] // 2]
        for _ in range(3):
            rows.index_select(0, idx)
        torch.cuda.synchronize()
        del src, dst, rows, idx
    ab = algorithmic_bytes(op, fold=mat is not None, branch=br is not None and mat is not None)
    mb = moved_bytes(op, fold=mat is not None, branch=br is not None and mat is not None, sym=sym)
    kern = {}
    for name, lst in ev.items():
        ms = float(np.mean([a.elapsed_time(b) for a, b in lst]))
        kern[name] = {"avg_us": round(ms * 1e3, 2), "alg_bytes": int(ab[name]),
                      "GB_s": round(ab[name] / (ms * 1e-3) / 1e9, 1)}
        if mb[name] != ab[name]:
            kern[name]["moved_bytes"] = int(mb[name])
            kern[name]["moved_GB_s"] = round(mb[name] / (ms * 1e-3) / 1e9, 1)
    step_bytes = ab["constitutive"] + ab["force"] + ab["tangent"]
    value = step_bytes / (step_ms * 1e-3) / 1e9
    moved = (mb["constitutive"] + mb["force"] + mb["tangent"]) / (step_ms * 1e-3) / 1e9
    dom = max(("constitutive", "force", "tangent"), key=lambda k: kern[k]["avg_us"])
    # branch histogram of the inputs as the device classified them
    br_dev = torch.empty(op.n_int, dtype=torch.int8, device=dev)
    _lib.call("ssx_constitutive_3d", ptr(op._E), ptr(op.c_bar), ptr(op.sin_phi), ptr(op.shear), ptr(op.lame),
              op.n_int, ptr(S), None, None, ptr(br_dev), st)
    hist = torch.bincount(br_dev.to(torch.int64), minlength=5).cpu().numpy()
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(dom)
        except Exception:
            traffic = None
    return {
        "m": m, "op": op, "E": E_host, "step_ms": step_ms, "value": value, "kernels": kern, "dominant": dom,
        "wall_s": wall, "hist": hist.tolist(), "traffic": traffic, "ab": ab, "moved_GB_s": moved,
        "geometry": "affine" if pat.affine else "points", "ds_layout": "sym21" if sym else "full36",
    }


def run_probes() -> dict:
    """What this GPU attains, beside the published peaks: a device copy and a STREAM triad
    (a = b + s c) over 1 GiB arrays, and the FP64 vector-FMA and FP64 matrix-core rates
    (`ssx_probe_fp64`)."""
    import ctypes

    import torch

    from slope_stability_petsc4py_amd import _lib
    from slope_stability_petsc4py_amd.device import ptr, stream_ptr

    dev = torch.device("cuda", 0)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) * 1e-3 / reps

    n = 1 << 27                      # 1 GiB of FP64 per array
    x = torch.ones(n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    z = torch.full_like(x, 2.0)
    t_copy = timed(lambda: y.copy_(x), 10)
    t_triad = timed(lambda: torch.add(x, z, alpha=3.0, out=y), 10)
    del x, y, z
    props = torch.cuda.get_device_properties(dev)
    blocks = props.multi_processor_count * 8
    out = torch.empty(blocks * 256, dtype=torch.float64, device=dev)
    flops = ctypes.c_double(0.0)
    rates = {}
    for mode, name, iters in ((0, "fp64_fma", 20000), (1, "fp64_mfma", 4000)):
        t = timed(lambda: _lib.call("ssx_probe_fp64", mode, blocks, iters, ptr(out), ctypes.addressof(flops),
                                    stream_ptr()), 3)
        rates[name] = flops.value / t / 1e12
    return {"hbm_copy_GB_s": round(2 * 8 * n / t_copy / 1e9, 1), "hbm_triad_GB_s": round(3 * 8 * n / t_triad / 1e9, 1),
            "fp64_fma_TFLOP_s": round(rates["fp64_fma"], 2), "fp64_mfma_TFLOP_s": round(rates["fp64_mfma"], 2),
            "published": {"hbm_GB_s": HBM_PEAK_GBS, "fp64_vector_TFLOP_s": 78.6, "fp64_matrix_TFLOP_s": 78.6}}


def cpu_baseline(op, E_host, ab, budget_s: float):
    """The reference's own C kernels (compiled from its sources into oracle/_ref) on the host cores:
    return mapping + force rows + tangent rows over a bounded sample of the same workload."""
    import oracle

    if oracle.ref() is None:
        return None
    geom = op.geom
    pat = op.pattern.host
    n_e_total = geom.n_elem
    # the whole workload (same strains, geometry and row-slot metadata) with the host's OpenMP threads;
    # repeated until the time budget is spent (at most 5 repetitions)
    n = geom.n_int
    cb = np.full(n, 4 * np.sqrt(3))
    sp = np.full(n, 1 / np.sqrt(3))
    from slope_stability_petsc4py_amd import workloads
    G = np.full(n, workloads.HOMO.shear)
    L = np.full(n, workloads.HOMO.lame)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    t0 = time.perf_counter()
    reps = 0
    t_const = t_force = t_tan = 0.0
    while True:
        a = time.perf_counter()
        S, DS = oracle.ref_constitutive_3d(E_host, cb, sp, G, L)
        b = time.perf_counter()
        oracle.ref_force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q)
        c = time.perf_counter()
        oracle.ref_tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
        d = time.perf_counter()
        t_const += b - a
        t_force += c - b
        t_tan += d - c
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 5:
            break
    per = (t_const + t_force + t_tan) / reps
    # the tangent rows once more on one thread, for comparison with the reference's single-thread
    # micro-benchmark figure (the OpenMP runtime the reference kernels were built against)
    t_tan1 = None
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        a = time.perf_counter()
        oracle.ref_tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
        t_tan1 = time.perf_counter() - a
        gomp.omp_set_num_threads(threads)
    except OSError:
        pass
    step_bytes = ab["constitutive"] + ab["force"] + ab["tangent"]
    return {
        "value": round(step_bytes / per / 1e9, 2),
        "unit": "GB/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"full workload ({n_e_total} tets, {n} points), {reps} reps of return mapping + force rows + "
                  f"tangent rows (reference C, -O3 -fopenmp); s/step {per:.3f} "
                  f"(const {t_const / reps:.3f}, force {t_force / reps:.3f}, tangent {t_tan / reps:.3f})",
        "s_per_step": round(per, 4),
        "tangent_rows_1_thread_s": None if t_tan1 is None else round(t_tan1, 3),
    }


def run_ssr_leg(log_progress=True) -> dict:
    """End-to-end leg of the metric: the reference's 3D heterogeneous P2 SSR (SSR_hetero_ada_L1,
    four layers, 82,815 dofs) to the factor of safety, plain FP64 GPU form (FP64 preconditioner,
    secant predictor) against the FP32-preconditioner form with the learned predictor (weights
    trained on the GPU by the project's kernels, `slope_stability_petsc4py_amd/data/`).  Wall time =
    continuation from the first Newton solve to the last accepted step, inputs on the device.  The
    faster forms also report their displacement and stress (FP64 return mapping at their final U and
    lambda) relative to the plain FP64 form's at the same omega, and the FoS difference."""
    import torch

    from slope_stability_petsc4py_amd import cases
    from slope_stability_petsc4py_amd.learn import LearnedPredictor, NodeMLP, node_info

    fos_ref = float(np.load(ROOT / "tests" / "golden" / "ssr_histories.npz")["hetero3d_lambda_hist"][-1])
    weights = ROOT / "slope_stability_petsc4py_amd" / "data" / "predictor_hetero3d_L1.npz"
    case = cases.hetero3d_L1()
    # warm-up: a short run loads every kernel and library once
    case.params.step_max = 4
    cases.run_ssr(cases.make_solver(case, precision="fp32"))
    case.params.step_max = cases.SSRParams().step_max
    out = {"workload": "3D heterogeneous P2 SSR (SSR_hetero_ada_L1, 18,419 tets, 82,815 dofs)", "fos_reference": fos_ref}
    net = NodeMLP.load(weights) if weights.exists() else None
    forms = [("plain_fp64", "fp64", False), ("fp32", "fp32", False)]
    if net is not None:
        forms.append(("learned_fp32", "fp32", True))
    finals = {}
    op64 = None
    for name, prec, learned in forms:
        s = cases.make_solver(case, precision=prec)
        pred = LearnedPredictor(net, node_info(case.mesh, case.materials)) if learned else None
        torch.cuda.synchronize()
        res = cases.run_ssr(s, predictor=pred)
        out[name] = {"wall_s": round(res.wall_time, 3), "fos": res.fos, "fos_abs_err": abs(res.fos - fos_ref),
                     "newton": res.newton_iterations, "linear": res.linear_iterations,
                     "ms_per_linear_iteration": round(1e3 * res.wall_time / max(res.linear_iterations, 1), 3),
                     "learned_steps": sum(1 for r in res.step_info if r["predictor"] == "learned")}
        finals[name] = (res.U.clone(), float(res.fos), float(res.omega_hist[-1]))
        if name == "plain_fp64":
            op64 = s.op
        if log_progress:
            _log(f"ssr {name}: {out[name]}")
        del s, res
    # errors of the faster forms against the plain FP64 form at the end of the branch (same omega):
    # displacement, and stress evaluated by the FP64 return mapping at each form's (U, lambda)
    U64, lam64, om64 = finals["plain_fp64"]
    op64.reduction(lam64)
    S64 = op64.constitutive_problem_stress(U64).clone()
    for name in [n for n in finals if n != "plain_fp64"]:
        U, lam, om = finals[name]
        if abs(om - om64) > 1e-9 * abs(om64):
            continue
        op64.reduction(lam)
        S = op64.constitutive_problem_stress(U)
        out[name]["u_rel_err_vs_fp64"] = float((torch.linalg.vector_norm(U - U64) / torch.linalg.vector_norm(U64)).item())
        out[name]["stress_rel_err_vs_fp64"] = float((torch.linalg.vector_norm(S - S64) /
                                                     torch.linalg.vector_norm(S64)).item())
        out[name]["fos_diff_vs_fp64"] = abs(lam - lam64)
    for name in [n for n in ("fp32", "learned_fp32") if n in out]:
        out[name]["speedup_vs_plain_fp64"] = round(out["plain_fp64"]["wall_s"] / out[name]["wall_s"], 3)
    # the learned form counts as the fast form only when the predictor's guesses were used; with
    # every guess gated off it ran the FP32 form's path (plus the gate's residual evaluations)
    learned_used = "learned_fp32" in out and out["learned_fp32"]["learned_steps"] > 0
    best = "learned_fp32" if learned_used else "fp32"
    out["speedup_vs_plain_fp64"] = out[best]["speedup_vs_plain_fp64"]
    out["fast_form"] = best
    return out


def run_ssr2d_leg(log_progress=True) -> dict:
    """The reference's 2D homogeneous SSR (10 m slope, P2 triangles, h = 1: 1300 elements, 7-point
    rule, 5,442 dofs) to the factor of safety: the plain FP64 method on the GPU against the CPU
    restatement of the same method (oracle/ssr2d.py: vectorised NumPy return mapping, scipy sparse
    B^T D B assembly, SuperLU direct solves; one host core), both timed from the first Newton solve
    to the last accepted step."""
    import torch

    import oracle.ssr2d as ssr2d
    from slope_stability_petsc4py_amd import cases, fem

    fos_ref = float(np.load(ROOT / "tests" / "golden" / "ssr_histories.npz")["homo2d_lambda_hist"][-1])
    case = cases.homo2d()
    p = case.params
    cases.run_ssr(cases.make_solver(case))          # warm-up (kernels, libraries)
    s = cases.make_solver(case)
    torch.cuda.synchronize()
    res = cases.run_ssr(s)
    gpu = {"wall_s": round(res.wall_time, 3), "fos": res.fos, "fos_abs_err": abs(res.fos - fos_ref),
           "newton": res.newton_iterations, "linear": res.linear_iterations}
    if log_progress:
        _log(f"ssr2d gpu: {gpu}")
    m = case.mesh
    geom = fem.geometry(m.coord, m.elem, m.elem_type)
    prob = ssr2d.Problem2D(m, geom, case.materials[0])
    f_int = np.zeros((2, geom.n_int))
    f_int[1] = -case.materials[0].gamma_sat
    f = fem.volume_load(geom, m.elem, m.n_nodes, f_int)
    f[~m.q_mask.astype(bool)] = 0.0
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):       # one host core: the BLAS behind NumPy/scipy limited to one thread
        r = ssr2d.ssr(prob, f.T.reshape(-1).copy(), lambda_init=p.lambda_init, d_lambda_init=p.d_lambda_init,
                      d_lambda_min=p.d_lambda_min, d_lambda_diff_scaled_min=p.d_lambda_diff_scaled_min,
                      step_max=p.step_max, omega_max_stop=p.omega_max_stop, it_newt_max=p.it_newt_max,
                      it_damp_max=p.it_damp_max, tol=p.tol, r_min=p.r_min)
    cpu = {"value": round(r["wall"], 3), "unit": "s to FoS", "cores": 1, "kind": "port",
           "fos": r["fos"], "fos_abs_err": abs(r["fos"] - fos_ref), "newton": r["newton"],
           "sample": "the whole 2D SSR run (oracle/ssr2d.py, NumPy/scipy, direct solves, BLAS limited to one "
                     "thread)"}
    if log_progress:
        _log(f"ssr2d cpu: {cpu}")
    return {"workload": "2D homogeneous P2 SSR (h = 1: 1300 triangles, 9100 points, 5442 dofs)",
            "fos_reference": fos_ref, "plain_fp64_gpu": gpu, "cpu_baseline": cpu,
            "gpu_over_cpu_speedup": round(r["wall"] / res.wall_time, 3)}


def run_p4(args):
    """P4 assembly on the matrix cores: 3D heterogeneous slope (SSR_hetero_ada_L2, 33,935 tets elevated
    to P4, 24-point rule, four layer materials), return mapping + element matrices (FP64 MFMA) +
    row gather into the fixed pattern with K_r and FP32 copy."""
    import torch

    from slope_stability_petsc4py_amd import _lib, cases, workloads
    from slope_stability_petsc4py_amd import mesh as M
    from slope_stability_petsc4py_amd.constitutive import ConstitutiveOperator
    from slope_stability_petsc4py_amd.device import ptr, stream_ptr

    t0 = time.time()
    m = cases.gmsh_fixture_mesh("ssr_hetero_L2_tet4.npz", "P4")
    op = ConstitutiveOperator.from_mesh(m, workloads.HETERO_LAYERS, Davis_type="B")
    pat = op.pattern
    _log(f"P4 mesh+pattern {time.time() - t0:.1f}s: {m.n_elem} tets, {op.n_int} points, {pat.n_rows} dofs, "
         f"nnz {pat.host.nnz}")
    E_u, _ = workloads.stratified_strains(min(args.unique, op.n_int), seed=args.seed)
    reps = -(-op.n_int // E_u.shape[0])
    op._E.copy_(torch.from_numpy(np.ascontiguousarray(np.tile(E_u, (reps, 1))[: op.n_int])))
    op.reduction(1.0)
    K = pat.new_csr(with32=True)
    st = stream_ptr()

    def k_const():
        _lib.call("ssx_constitutive_3d", ptr(op._E), ptr(op.c_bar), ptr(op.sin_phi), ptr(op.shear), ptr(op.lame),
                  op.n_int, ptr(op._S), ptr(op._DS), None, None, st)

    r = 1e-4

    def k_ke():
        # K_r folded into the point tangents: (1 - r) DS + r D_el
        pat.element_matrices_p4(op._DS, r, (op.shear, op.lame))

    def k_gather():
        pat.gather_nodes(pat._ke, K.values, K.values32)

    step = [("constitutive", k_const), ("ke_mfma", k_ke), ("gather", k_gather)]
    for _ in range(args.warmup):
        for _, f in step:
            f()
    torch.cuda.synchronize()
    ev = {n: [] for n, _ in step}
    a0, b0 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a0.record()
    for _ in range(args.steps):
        for name, f in step:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            f()
            b.record()
            ev[name].append((a, b))
    b0.record()
    torch.cuda.synchronize()
    step_ms = a0.elapsed_time(b0) / args.steps
    ms = {n: float(np.mean([a.elapsed_time(b) for a, b in lst])) for n, lst in ev.items()}
    n_e = m.n_elem
    flops_full = 2.0 * 105 * 105 * 144 * n_e
    flops_exec = 28 * 36 * (16 * 16 * 4 * 2) * n_e
    nnz = pat.host.nnz
    ke_bytes = n_e * 105 * 105 * 8
    out = {
        "metric": METRIC + " (P4 MFMA assembly)", "value": round(flops_full / (ms["ke_mfma"] * 1e-3) / 1e12, 2),
        "unit": "TFLOP/s (effective, full B^T D B)", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D hetero P4 slope (SSR_hetero_ada_L2 elevated to P4)", "n_elem": n_e,
                   "n_points": int(op.n_int), "n_dof": int(pat.n_rows), "nnz": int(nnz)},
        "roofline": {"kernel": "ke_mfma", "bound": "mfma", "achieved": round(flops_exec / (ms["ke_mfma"] * 1e-3) / 1e12, 2),
                     "peak": 78.6, "unit": "TFLOP/s", "frac": round(flops_exec / (ms["ke_mfma"] * 1e-3) / 1e12 / 78.6, 4),
                     "traffic": None},
        "kernels": {
            "constitutive": {"avg_us": round(ms["constitutive"] * 1e3, 1),
                             "GB_s": round(op.n_int * 416 / (ms["constitutive"] * 1e-3) / 1e9, 1)},
            "ke_mfma": {"avg_us": round(ms["ke_mfma"] * 1e3, 1), "executed_TFLOP_s":
                        round(flops_exec / (ms["ke_mfma"] * 1e-3) / 1e12, 2),
                        "effective_TFLOP_s": round(flops_full / (ms["ke_mfma"] * 1e-3) / 1e12, 2)},
            # element matrices read once + node-slot offsets (2 B per local column) + K_r FP64/FP32 out
            "gather": {"avg_us": round(ms["gather"] * 1e3, 1),
                       "GB_s": round((ke_bytes + int(pat.ns.elem.size) * 105 * 2 + nnz * 12) / (ms["gather"] * 1e-3)
                                     / 1e9, 1)},
        },
        "cpu_baseline": None,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="kernel3d", choices=["kernel3d", "p4"],
                    help="kernel3d: 3D homogeneous P2 slope (default); p4: heterogeneous P4 slope on MFMA")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--h", type=float, default=1.0)
    ap.add_argument("--nz", type=int, default=30)
    ap.add_argument("--unique", type=int, default=131072)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 256 MiB device copy (calibrates HBM byte counters under rocprofv3 --pmc)")
    ap.add_argument("--no-ssr", action="store_true", help="skip the end-to-end SSR leg")
    ap.add_argument("--no-p4", action="store_true", help="skip the P4 matrix-core leg")
    ap.add_argument("--no-probes", action="store_true", help="skip the copy / triad / FP64 rate probes")
    ap.add_argument("--tangent", default="nodes", choices=["nodes", "rows"],
                    help="P2 tangent path: element matrices + node gather (default) or the fused row kernel")
    ap.add_argument("--geometry", default="affine", choices=["affine", "points"],
                    help="P2 gradients formed from per-element factors (default, straight-sided meshes) or read "
                         "per point (dphi, w)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the path runs on a single GPU, and sharding it over several "
              "GPUs is a design, not an implementation.", file=sys.stderr)
        sys.exit(2)
    if args.config == "p4":
        print(json.dumps(run_p4(args)), flush=True)
        return
    probes = None if args.no_probes else run_probes()
    if probes:
        _log(f"probes: {probes}")
    res = run_gpu(args)
    op = res["op"]
    kern = res["kernels"]
    dom = res["dominant"]
    dk = kern[dom]
    achieved = dk["alg_bytes"] / (dk["avg_us"] * 1e-6) / 1e9
    cpu = None
    if not args.no_cpu:
        _log("cpu baseline (reference C on host cores)...")
        cpu = cpu_baseline(op, res["E"], res["ab"], args.cpu_budget)
    p4 = None
    if not args.no_p4:
        _log("P4 leg (matrix-core element matrices)...")
        r4 = run_p4(args)
        p4 = {k: r4[k] for k in ("value", "unit", "ms_per_step", "config", "roofline", "kernels")}
        if probes:
            # the same achieved rate against what this GPU sustains in the pure-MFMA probe
            p4["roofline"]["frac_of_probe"] = round(p4["roofline"]["achieved"] / probes["fp64_mfma_TFLOP_s"], 4)
    ssr = None
    if not args.no_ssr:
        _log("end-to-end SSR leg (3D heterogeneous P2)...")
        ssr = run_ssr_leg()
        _log("end-to-end 2D SSR leg, GPU and its CPU baseline...")
        ssr["ssr2d"] = run_ssr2d_leg()
    cpu_model = ""
    try:
        cpu_model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        cpu_model = next((l.split(":", 1)[1].strip() for l in cpu_model.splitlines() if "Model name" in l), "")
    except Exception:
        pass
    out = {
        "metric": METRIC,
        "value": round(res["value"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["step_ms"], 4),
        "moved_GB_s": round(res["moved_GB_s"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D homogeneous P2 slope kernel bench: return mapping + force + K_r tangent assembly "
                        "(FP64, FP32 shadow)",
            "n_elem": int(op.pattern.n_elem), "n_points": int(op.n_int), "n_dof": int(op.pattern.n_rows),
            "nnz": int(op.pattern.host.nnz), "branch_hist_el_smooth_left_right_apex": res["hist"],
            "parallelism": "single GPU", "tangent_path": args.tangent, "geometry": res["geometry"],
                       "ds_layout": res["ds_layout"],
        },
        "roofline": {
            "kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": res["traffic"],
            "moved_GB_s": kern[dom].get("moved_GB_s", round(achieved, 1)),
            "traffic_source": ("copied from profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of "
                               "this bench, copy-calibrated; profiles/kernel3d_p2_117k_hbm_traffic.md), not "
                               "measured in this run") if res["traffic"] is not None else None,
        },
        "cpu_baseline": cpu,
        "probes": probes,
        "p4": p4,
        "ssr": ssr,
        "kernels": kern,
        "host": {"cpu_model": cpu_model, "nproc": os.cpu_count(), "python": platform.python_version()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
