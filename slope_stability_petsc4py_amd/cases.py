"""Problem set-ups of the benchmark configurations and the SSR / LL runners on the device.

* ``homo2d``: the reference's 2D homogeneous SSR capture (benchmarks/run_2D_homo_SSR_capture:
  h = 1, x1 = x3 = 15, x2 = y1 = y2 = 10, c = 6 kPa, phi = 45, psi = 0, E = 40 MPa, nu = 0.3,
  gamma = 20, Davis B; continuation lambda_init 0.9, d_lambda_init 0.1, d_lambda_min 1e-5,
  d_lambda_diff_scaled_min 1e-3, omega_max 7e7, Newton tol 1e-4, r_min 1e-4, it_newt_max 50).
* ``hetero3d_L1``: the reference's 3D heterogeneous SSR capture on SSR_hetero_ada_L1 (P2,
  materials of meshes/3d_hetero_ssr/definition.py, Dirichlet labels x:[1,2] y:[5] z:[3,4];
  lambda_init 1.0, d_lambda_init 0.1, d_lambda_min 1e-3, omega_max 1.2e7, it_newt_max 200).
* ``slope3d``: structured 3D slope (homogeneous or four horizontal layers).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from . import mesh as M
from .constitutive import ConstitutiveOperator
from .continuation import ll_indirect_continuation, ssr_indirect_continuation
from .device import to_dev
from .linear import ChebyshevJacobi, DeflatedFGMRES, PMultigrid, PMultigridP4
from .loads import strip_footing_load
from .materials import MaterialSpec
from .workloads import HETERO_LAYERS, HOMO

# the reference's Gmsh tet4 meshes as arrays (written by tests/golden/make_golden.py)
DATA = Path(__file__).resolve().parent / "data" / "meshes"

HETERO_L1_MATERIALS = [
    MaterialSpec(15.0, 30.0, 0.0, 10000.0, 0.33, 19.0, 19.0),
    MaterialSpec(15.0, 38.0, 0.0, 50000.0, 0.30, 22.0, 22.0),
    MaterialSpec(10.0, 35.0, 0.0, 50000.0, 0.30, 21.0, 21.0),
    MaterialSpec(18.0, 32.0, 0.0, 20000.0, 0.33, 20.0, 20.0),
]


@dataclass
class SSRParams:
    lambda_init: float = 1.0
    d_lambda_init: float = 0.1
    d_lambda_min: float = 1e-3
    d_lambda_diff_scaled_min: float = 1e-3
    step_max: int = 100
    omega_max_stop: float = 1.2e7
    it_newt_max: int = 200
    it_damp_max: int = 10
    tol: float = 1e-4
    r_min: float = 1e-4
    linear_tol: float = 1e-1
    linear_maxit: int = 100


@dataclass
class Case:
    name: str
    mesh: M.SlopeMesh
    materials: list
    params: SSRParams
    davis: str = "B"
    extra: dict = field(default_factory=dict)
    # limit load under a strip-footing pressure: (x_min, x_max, q) of the loaded top-surface strip;
    # the continuation load f is then the footing load, and gravity (if kept) a constant load
    footing: tuple | None = None
    gravity: bool = True


def homo2d(h: float = 1.0, x1: float = 15.0, x2: float = 10.0, x3: float = 15.0) -> Case:
    m = M.slope_2d_p2(h=h, x1=x1, x2=x2, x3=x3, y1=10.0, y2=10.0)
    p = SSRParams(lambda_init=0.9, d_lambda_init=0.1, d_lambda_min=1e-5, d_lambda_diff_scaled_min=1e-3,
                  step_max=100, omega_max_stop=7e7, it_newt_max=50)
    return Case("homo2d", m, [HOMO], p)


def gmsh_fixture_mesh(fixture: str, elem_type: str = "P2") -> M.SlopeMesh:
    """A committed Gmsh-family mesh (data/meshes/*_tet4.npz): logical material and boundary ids
    from the file's physical names, Dirichlet labels x: [1, 2], y: [5], z: [3, 4] (the families'
    definition.py)."""
    d = np.load(DATA / fixture, allow_pickle=False)
    return M.tet_mesh_from_arrays(d["coord"], d["tet4"].astype(np.int64), d["tet_label"].astype(np.int64),
                                  d["tri3"].astype(np.int64), d["tri_label"].astype(np.int64),
                                  {"x": [1, 2], "y": [5], "z": [3, 4]}, elem_type)


def hetero3d_L1(elem_type: str = "P2") -> Case:
    """elem_type "P4" gives the reference's P4 L1 problem (616,322 free dofs)."""
    return Case("hetero3d_L1", gmsh_fixture_mesh("ssr_hetero_L1_tet4.npz", elem_type), HETERO_L1_MATERIALS,
                SSRParams())


def hetero3d_L2_layers(elem_type: str = "P4") -> Case:
    """The heterogeneous P4 configuration: the reference's SSR_hetero_ada_L2 mesh (33,935 tets;
    elevated to P4: 1.14M dofs) with the four horizontal layer materials HETERO_LAYERS on its four
    material labels, SSR parameters of the reference's heterogeneous capture."""
    return Case("hetero3d_L2_layers", gmsh_fixture_mesh("ssr_hetero_L2_tet4.npz", elem_type), list(HETERO_LAYERS),
                SSRParams())


HOMO_LL_MATERIAL = MaterialSpec(6.0, 45.0, 0.0, 40000.0, 0.30, 20.0, 20.0)


def _ll_params(d_omega_ini_scale: float, omega_max: float, it_newt_max: int) -> SSRParams:
    p = SSRParams(step_max=100, omega_max_stop=omega_max, it_newt_max=it_newt_max)
    p.d_omega_ini_scale = d_omega_ini_scale
    p.d_t_min = 1e-3
    return p


def _gmsh_case(name: str, fixture: str, materials, elem_type: str, params) -> Case:
    return Case(name, gmsh_fixture_mesh(fixture, elem_type), materials, params)


def homo2d_ll(h: float = 0.5) -> Case:
    """The reference's 2D homogeneous LL (benchmarks/slope_stability_2D_homo_LL/case.toml: h = 0.5,
    x1 = x3 = 15, d_omega_ini_scale 0.2, omega_max 2000, it_newt_max 100)."""
    m = M.slope_2d_p2(h=h, x1=15.0, x2=10.0, x3=15.0, y1=10.0, y2=10.0)
    return Case("homo2d_ll", m, [HOMO], _ll_params(0.2, 2000.0, 100))


def homo3d_ll_L1(elem_type: str = "P2") -> Case:
    """The reference's 3D homogeneous LL on LL_homo_ada_L1 (benchmarks/slope_stability_3D_homo_LL:
    d_omega_ini_scale 0.2, omega_max 4900, it_newt_max 200)."""
    return _gmsh_case("homo3d_ll_L1", "ll_homo_L1_tet4.npz", [HOMO_LL_MATERIAL], elem_type,
                      _ll_params(0.2, 4900.0, 200))


def hetero3d_ll_L1(elem_type: str = "P2") -> Case:
    """The reference's 3D heterogeneous LL on LL_hetero_ada_L1 (benchmarks/slope_stability_3D_hetero_LL:
    the four layer materials of meshes/3d_hetero_ll/definition.py, d_omega_ini_scale 0.5,
    omega_max 8e7, it_newt_max 200)."""
    return _gmsh_case("hetero3d_ll_L1", "ll_hetero_L1_tet4.npz", HETERO_L1_MATERIALS, elem_type,
                      _ll_params(0.5, 8.0e7, 200))


def prandtl_nc(phi_deg: float) -> float:
    """Prandtl's bearing-capacity factor N_c = (N_q - 1) cot(phi), N_q = e^(pi tan phi) tan^2(45 + phi/2):
    the limit pressure of a strip footing on a weightless associated Mohr-Coulomb half-space is c N_c."""
    ph = np.deg2rad(phi_deg)
    if ph == 0.0:
        return 2.0 + np.pi
    nq = np.exp(np.pi * np.tan(ph)) * np.tan(np.pi / 4 + ph / 2) ** 2
    return float((nq - 1.0) / np.tan(ph))


def footing2d_prandtl(h: float = 0.25, phi: float = 10.0, c: float = 10.0) -> Case:
    """Strip footing (half width 1 m at the symmetry plane x = 0) on a weightless, associated
    (psi = phi) Mohr-Coulomb layer, plane strain, P2 triangles; the footing pressure is q = c, so
    the limit load factor approaches N_c(phi) (Prandtl) as the mesh is refined.  The 16 x 9 m
    domain holds the Prandtl mechanism (a few metres wide and deep) well inside."""
    m = M.slope_2d_p2(h=h, x1=12.0, x2=1.0, x3=3.0, y1=8.0, y2=1.0)
    mat = MaterialSpec(c, phi, phi, 40000.0, 0.30, 0.0, 0.0)
    p = _ll_params(1.0, 1e12, 100)
    p.step_max = 400
    return Case("footing2d_prandtl", m, [mat], p, footing=(0.0, 1.0, c), gravity=False,
                extra={"Nc": prandtl_nc(phi)})


def footing3d_hetero(h: float = 2.0, z_len: float = 20.0, width: float = 4.0, q: float = 100.0) -> Case:
    """3D heterogeneous slope (the four horizontal layers of the P4 configuration, psi = 0) under
    its own weight, with a strip footing of the given width at the symmetry plane x = 0 of the crest
    across the whole z extent, pressure q (kPa) scaled by the load factor t (gravity held constant),
    followed to omega = f^T U = 3e4.  P2 tetrahedra."""
    m = M.slope_3d(h=h, z_len=z_len, layers=(15.0, 10.0, 5.0))
    # the load-displacement curve keeps rising slowly past the knee (6.86 at omega = 3e4, 6.94 at
    # 2.4e5 on h = 1.5), so the branch is followed to a fixed omega, where forms are compared
    p = _ll_params(0.5, 3e4, 100)
    p.step_max = 400
    return Case("footing3d_hetero", m, list(HETERO_LAYERS), p, footing=(0.0, width, q), gravity=True)


def slope3d(h: float = 2.0, nz: int = 10, hetero: bool = False) -> Case:
    layers = (15.0, 10.0, 5.0) if hetero else None
    m = M.slope_3d(h=h, nz=nz, layers=layers)
    mats = HETERO_LAYERS if hetero else [HOMO]
    p = SSRParams(lambda_init=1.0 if hetero else 0.9, omega_max_stop=1e9)
    return Case("slope3d_hetero" if hetero else "slope3d_homo", m, mats, p)


@dataclass
class Solver:
    case: Case
    op: ConstitutiveOperator
    f: torch.Tensor
    free_mask: torch.Tensor
    linear: DeflatedFGMRES
    f_const: torch.Tensor | None = None


def make_solver(case: Case, *, precision: str = "fp64", pc: str = "pmg", coarse_refresh: int = 0,
                coarse_precision: str | None = None, refresh_iterations: int = 20,
                refresh_ratio: float = 1.5, coarse: str = "auto") -> Solver:
    """precision 'fp64': FP64 preconditioner; 'fp32': FP32 tangent shadow + FP32 preconditioner
    (the outer FGMRES, residuals and return mapping stay FP64).  `coarse_refresh`: the p-multigrid
    coarse operator is rebuilt every that many preconditioner setups (1 = every Newton iteration),
    or, with 0 (default), when a solve needs more than `refresh_iterations` FGMRES iterations and
    more than `refresh_ratio` times the iterations of the first solve after the last rebuild.
    `coarse`: 'dense' (dense P1 level), 'aggregate' (sparse P1 level above a dense aggregate level)
    or 'auto' (aggregate when the free P1 dofs exceed 50000)."""
    use32 = precision == "fp32"
    op = ConstitutiveOperator.from_mesh(case.mesh, case.materials, Davis_type=case.davis, fp32_shadow=use32)
    f_field = op.gravity_load(case.mesh)
    f = to_dev(f_field.T.reshape(-1).copy())
    f_const = None
    if case.footing is not None:
        x0, x1, q = case.footing
        F_foot, _ = strip_footing_load(case.mesh, (x0, x1), q)
        f_const = f if case.gravity else None
        f = to_dev(F_foot.T.reshape(-1).copy())
    free_mask = to_dev(case.mesh.q_mask.T.reshape(-1).astype(np.float64))
    if pc == "pmg" and case.mesh.elem_type == "P4":
        prec = PMultigridP4(case.mesh, op, use32=use32, coarse_refresh=coarse_refresh,
                            refresh_iterations=refresh_iterations, refresh_ratio=refresh_ratio, coarse=coarse)
    elif pc == "pmg":
        prec = PMultigrid(case.mesh.elem, case.mesh.q_mask, use32=use32, coarse_refresh=coarse_refresh,
                          coarse_precision=coarse_precision, refresh_iterations=refresh_iterations,
                          refresh_ratio=refresh_ratio, coarse=coarse, coord=case.mesh.coord)
    else:
        prec = ChebyshevJacobi(use32=use32)
    lin = DeflatedFGMRES(prec, tolerance=case.params.linear_tol, max_iterations=case.params.linear_maxit)
    return Solver(case, op, f, free_mask, lin, f_const)


def run_ssr(s: Solver, *, predictor=None, store_step_u: bool = False, progress=None):
    p = s.case.params
    return ssr_indirect_continuation(p.lambda_init, p.d_lambda_init, p.d_lambda_min, p.d_lambda_diff_scaled_min,
                                     p.step_max, p.omega_max_stop, p.it_newt_max, p.it_damp_max, p.tol, p.r_min,
                                     s.f, s.op, s.linear, free_mask=s.free_mask, dim=s.case.mesh.dim,
                                     predictor=predictor, store_step_u=store_step_u, progress=progress)


def run_ll(s: Solver, *, d_omega_ini_scale: float | None = None, d_t_min: float | None = None, progress=None,
           predictor=None, store_step_u: bool = False):
    p = s.case.params
    if d_omega_ini_scale is None:
        d_omega_ini_scale = getattr(p, "d_omega_ini_scale", 0.2)
    if d_t_min is None:
        d_t_min = getattr(p, "d_t_min", 1e-3)
    return ll_indirect_continuation(d_omega_ini_scale, d_t_min, p.step_max, p.omega_max_stop, p.it_newt_max,
                                    p.it_damp_max, p.tol, p.r_min, s.f, s.op, s.linear, free_mask=s.free_mask,
                                    dim=s.case.mesh.dim, progress=progress, f_const=s.f_const, predictor=predictor,
                                    store_step_u=store_step_u)
