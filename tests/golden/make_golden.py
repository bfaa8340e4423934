"""Write the golden fixtures under tests/golden/ (run in a container that has the reference tree).

* constitutive_3d.npz -- branch-stratified strains over several materials and the S / DS the
  reference's own C return mapping (cython/constitutive_3D_kernel.h, compiled from source by
  `make -C oracle ref`) returns for them.
* assembly_3d.npz     -- a small P2 tetrahedral slope: geometry, pattern, a displacement, random
  symmetric DS and stress, and the strain / force / tangent-rows values of the reference's C
  kernels (cython/assemble_tangent_values_3d.c) on them.
* ssr_histories.npz   -- data fields of the reference's committed SSR runs (lambda / omega / Umax
  histories of benchmarks/run_2D_homo_SSR_capture and benchmarks/run_3D_hetero_SSR_capture).
* ll_histories.npz    -- the same fields of its committed limit-load runs (2D homo, 3D homo, 3D hetero).
* *_tet4.npz          -- the reference's Gmsh tet4 meshes (SSR hetero L1/L2, LL homo/hetero L1) as arrays,
  written to slope_stability_petsc4py_amd/data/meshes/ (the package's cases load them):
  raw physical tags and the logical material / boundary ids their $PhysicalNames give
  ("material_<k>", "boundary_<k>", io.py:165-189), which the Dirichlet labels refer to.

Usage: python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
MESH_OUT = ROOT / "slope_stability_petsc4py_amd" / "data" / "meshes"
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from slope_stability_petsc4py_amd import fem, mesh, pattern, workloads  # noqa: E402

OUT = Path(__file__).resolve().parent
REF = Path("/root/reference")


def constitutive() -> None:
    rng = np.random.default_rng(20240611)
    Es, mats = [], []
    for c0, phi_deg, young, nu, lam in [(6.0, 45.0, 40000.0, 0.3, 1.0), (15.0, 30.0, 50000.0, 0.3, 1.3),
                                        (10.0, 38.0, 30000.0, 0.3, 1.7), (18.0, 35.0, 60000.0, 0.3, 0.9)]:
        cb, sp = oracle.davis_reduction(np.array([c0]), np.array([np.deg2rad(phi_deg)]), np.array([0.0]), lam, "B")
        G = young / (2 * (1 + nu))
        K = young / (3 * (1 - 2 * nu))
        L = K - 2 * G / 3
        E, _b = workloads.stratified_strains(150, seed=int(rng.integers(1 << 30)), c_bar=float(cb[0]),
                                             sin_phi=float(sp[0]), shear=G, lame=L,
                                             mix=(0.3, 0.25, 0.15, 0.15, 0.15))
        Es.append(E)
        mats.append(np.tile([cb[0], sp[0], G, L], (E.shape[0], 1)))
    E = np.concatenate(Es)
    M = np.concatenate(mats)
    S, DS = oracle.ref_constitutive_3d(E, M[:, 0], M[:, 1], M[:, 2], M[:, 3])
    np.savez_compressed(OUT / "constitutive_3d.npz", E=E, c_bar=M[:, 0], sin_phi=M[:, 1], shear=M[:, 2],
                        lame=M[:, 3], S=S, DS=DS)


def assembly() -> None:
    m = mesh.slope_3d(h=5.0, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=10.0, nz=2)
    geom = fem.geometry(m.coord, m.elem, m.elem_type)
    pat = pattern.build_owned_pattern(m.elem, m.q_mask)
    rng = np.random.default_rng(7)
    u = rng.normal(0.0, 1e-3, 3 * m.n_nodes)
    A = rng.normal(size=(geom.n_int, 6, 6))
    DS = (A + A.transpose(0, 2, 1)).reshape(geom.n_int, 36)
    S = rng.normal(size=(geom.n_int, 6))
    E_ref = oracle.ref_strain_3d(geom.dphi, u, pat.elem_dofs, geom.n_q)
    F_ref = oracle.ref_force_3d_rows(geom.dphi, S, geom.weight, pat, geom.n_q)
    K_ref = oracle.ref_tangent_3d_rows(geom.dphi, DS, geom.weight, pat, geom.n_q)
    np.savez_compressed(
        OUT / "assembly_3d.npz", coord=m.coord, elem=m.elem, q_mask=m.q_mask, dphi=geom.dphi, weight=geom.weight,
        u=u, DS=DS, S=S, E=E_ref, F=F_ref, K=K_ref, indptr=pat.indptr, indices=pat.indices,
        row_slot_ptr=pat.row_slot_ptr, slot_elem=pat.slot_elem, slot_lrow=pat.slot_lrow, slot_pos=pat.slot_pos,
        constrained_diag_positions=pat.constrained_diag_positions, n_q=geom.n_q)


def histories() -> None:
    out = {}
    for tag, case in (("homo2d", "run_2D_homo_SSR_capture"), ("hetero3d", "run_3D_hetero_SSR_capture")):
        d = np.load(REF / "benchmarks" / case / "artifacts" / "simulation" / "data" / "petsc_run.npz",
                    allow_pickle=False)
        for k in ("lambda_hist", "omega_hist", "Umax_hist"):
            out[f"{tag}_{k}"] = np.asarray(d[k])
    np.savez_compressed(OUT / "ssr_histories.npz", **out)


def ll_histories() -> None:
    """Load-factor (lambda_hist = t), omega and Umax histories of the reference's committed limit-load
    runs (benchmarks/slope_stability_{2D_homo,3D_homo,3D_hetero}_LL, PETSc, 8 ranks)."""
    out = {}
    for tag, case in (("homo2d", "slope_stability_2D_homo_LL"), ("homo3d", "slope_stability_3D_homo_LL"),
                      ("hetero3d", "slope_stability_3D_hetero_LL")):
        d = np.load(REF / "benchmarks" / case / "artifacts" / "simulation" / "data" / "petsc_run.npz",
                    allow_pickle=False)
        for k in ("lambda_hist", "omega_hist", "Umax_hist"):
            out[f"{tag}_{k}"] = np.asarray(d[k])
    np.savez_compressed(OUT / "ll_histories.npz", **out)


def meshes() -> None:
    """The reference's Gmsh tet4 meshes as compact arrays (node coordinates, tet4 with physical
    tags, boundary triangles with physical tags), read by this project's Gmsh 2.2 reader."""
    for name, rel in (("ssr_hetero_L1_tet4", "3d_hetero_ssr/SSR_hetero_ada_L1.msh"),
                      ("ssr_hetero_L2_tet4", "3d_hetero_ssr/SSR_hetero_ada_L2.msh"),
                      ("ll_homo_L1_tet4", "3d_homo_ll/LL_homo_ada_L1.msh"),
                      ("ll_hetero_L1_tet4", "3d_hetero_ll/LL_hetero_ada_L1.msh")):
        path = REF / "meshes" / rel
        coord, tet4, tet_phys, tri3, tri_phys = mesh.read_gmsh22(path)
        names = mesh.read_gmsh_physical_names(path)
        np.savez_compressed(MESH_OUT / f"{name}.npz", coord=coord, tet4=tet4.astype(np.int32),
                            tet_phys=tet_phys.astype(np.int8), tri3=tri3.astype(np.int32),
                            tri_phys=tri_phys.astype(np.int8),
                            tet_label=mesh.logical_ids(tet_phys, names, 3, "material").astype(np.int8),
                            tri_label=mesh.logical_ids(tri_phys, names, 2, "boundary").astype(np.int8))


if __name__ == "__main__":
    if oracle.ref() is None:
        raise SystemExit("build the reference kernels first: make -C oracle ref")
    constitutive()
    assembly()
    histories()
    ll_histories()
    meshes()
    print("golden fixtures written to", OUT)
