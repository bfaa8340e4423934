"""The constitutive/assembly operator on the device: the drop-in for the reference's
``ConstitutiveOperator`` (constitutive/problem.py:1193-2210).

Method names, call order and semantics follow the reference so that Newton and the continuation
call it unchanged: ``reduction(lam)`` sets the reduced strength; ``constitutive_problem_stress(U)``
/ ``constitutive_problem_stress_tangent(U)`` evaluate S (and DS) at every point; ``build_F*``
assemble the internal force (constrained rows zeroed, fem/distributed_tangent.py:932);
``build_F_K_tangent*`` / ``build_K_regularized(r)`` / ``build_F_K_regularized*`` return the tangent
or K_r = (1-r) K + r K_el on the fixed pattern as a device CSR handle refreshed in place (the
reference caches and updates its PETSc AIJ the same way, constitutive/problem.py:1572-1644).

Every evaluation runs the ssx HIP kernels; there is no host arithmetic on the path and no CPU
fallback.  U may be a NumPy (dim, n_nodes) array, a torch (dim, n_nodes) tensor or a flat device
vector of dofs (dof = dim*node + comp); results come back in the same kind.
"""

from __future__ import annotations

from time import perf_counter

import numpy as np
import torch

from . import _lib
from .device import DeviceCSR, DevicePattern, device, ptr, stream_ptr, to_dev
from .fem import affine_geometry, geometry, volume_load
from .materials import expand_materials
from .pattern import build_owned_pattern

_DAVIS = {"A": 0, "B": 1, "C": 2}


def reduction(c0, phi, psi, lam: float, Davis_type: str = "A"):
    """Reduced strength (c_bar, sin_phi) per point on the device (constitutive/reduction.py:14-78)."""
    typ = str(Davis_type).upper()
    if typ not in _DAVIS:
        raise ValueError("Incorrect choice of the Davis approach")
    if lam <= 0:
        raise ValueError("Reduction parameter lambda must be positive")
    c0_t, phi_t, psi_t = (to_dev(a, torch.float64).reshape(-1) for a in (c0, phi, psi))
    if not (c0_t.shape == phi_t.shape == psi_t.shape):
        raise ValueError("c0, phi and psi must have the same shape")
    n = c0_t.numel()
    cb = torch.empty(n, dtype=torch.float64, device=c0_t.device)
    sp = torch.empty_like(cb)
    _lib.call("ssx_reduction", ptr(c0_t), ptr(phi_t), ptr(psi_t), n, float(lam), _DAVIS[typ], ptr(cb), ptr(sp),
              stream_ptr())
    return cb, sp


def constitutive_problem_3d(E, c_bar, sin_phi, shear, lame, *, return_tangent: bool = False, fp32_tangent=False,
                            return_branch: bool = False):
    """Batch return mapping on point-major device arrays: E (n, 6) -> S (n, 6) [, DS (n, 36)]."""
    E = to_dev(E, torch.float64)
    n = E.shape[0]
    dev = E.device
    S = torch.empty((n, 6), dtype=torch.float64, device=dev)
    DS = torch.empty((n, 36), dtype=torch.float64, device=dev) if return_tangent else None
    DS32 = torch.empty((n, 36), dtype=torch.float32, device=dev) if fp32_tangent else None
    br = torch.empty(n, dtype=torch.int8, device=dev) if return_branch else None
    args = [to_dev(a, torch.float64) for a in (c_bar, sin_phi, shear, lame)]
    _lib.call("ssx_constitutive_3d", ptr(E), *(ptr(a) for a in args), n, ptr(S), ptr(DS), ptr(DS32), ptr(br),
              stream_ptr())
    out = [S]
    if return_tangent:
        out.append(DS)
    if fp32_tangent:
        out.append(DS32)
    if return_branch:
        out.append(br)
    return out[0] if len(out) == 1 else tuple(out)


def constitutive_problem_2d(E, c_bar, sin_phi, shear, bulk, lame, *, return_tangent: bool = False,
                            return_branch: bool = False):
    """Plane-strain batch return mapping: E (n, 3) -> S (n, 4) [, DS (n, 9)]."""
    E = to_dev(E, torch.float64)
    n = E.shape[0]
    dev = E.device
    S = torch.empty((n, 4), dtype=torch.float64, device=dev)
    DS = torch.empty((n, 9), dtype=torch.float64, device=dev) if return_tangent else None
    br = torch.empty(n, dtype=torch.int8, device=dev) if return_branch else None
    args = [to_dev(a, torch.float64) for a in (c_bar, sin_phi, shear, bulk, lame)]
    _lib.call("ssx_constitutive_2d", ptr(E), *(ptr(a) for a in args), n, ptr(S), ptr(DS), None, ptr(br),
              stream_ptr())
    out = [S]
    if return_tangent:
        out.append(DS)
    if return_branch:
        out.append(br)
    return out[0] if len(out) == 1 else tuple(out)


class ConstitutiveOperator:
    """Device constitutive operator on the fixed owned pattern (one device owns every row)."""

    def __init__(self, pattern: DevicePattern, c0, phi, psi, shear, bulk, lame, *, Davis_type: str = "B",
                 q_mask: np.ndarray | None = None, fp32_shadow: bool = True):
        self.pattern = pattern
        self.owned_tangent_pattern = pattern
        self.dim = pattern.dim
        self.n_strain = pattern.n_strain
        self.n_int = pattern.n_int
        self.Davis_type = str(Davis_type).upper()
        self.c0, self.phi, self.psi = (to_dev(a, torch.float64) for a in (c0, phi, psi))
        self.shear, self.bulk, self.lame = (to_dev(a, torch.float64) for a in (shear, bulk, lame))
        self.WEIGHT = pattern.weight
        self.q_mask = q_mask
        self.fp32_shadow = bool(fp32_shadow)
        dev = pattern.indptr.device
        ns = self.n_strain
        self._E = torch.empty((self.n_int, ns), dtype=torch.float64, device=dev)
        self._S = torch.empty((self.n_int, ns + (1 if self.dim == 2 else 0)), dtype=torch.float64, device=dev)
        # branch ids of the return mapping that wrote the tangent (P2 tetrahedra: elastic points then
        # enter the tangent in closed form)
        self._branch = (torch.empty(self.n_int, dtype=torch.int8, device=dev)
                        if self.dim == 3 and pattern.n_p == 10 and pattern.ns is not None else None)
        # P2 tetrahedra in the constant-Jacobian form keep the (symmetric) tangent packed, 21 values
        # per point (_DSp); the full (n_int, 36) array is unpacked only when DS is asked for
        self._sym = self._branch is not None and pattern.affine
        self._DSp = torch.empty((self.n_int, 21), dtype=torch.float64, device=dev) if self._sym else None
        self._DS = None if self._sym else torch.empty((self.n_int, ns * ns), dtype=torch.float64, device=dev)
        self._ds_unpacked = False
        self._fe = torch.empty(pattern.n_elem * pattern.nld, dtype=torch.float64, device=dev)
        self._F = torch.empty(pattern.n_rows, dtype=torch.float64, device=dev)
        self._ds_ready = False
        self._s_ready = False
        self.c_bar = None
        self.sin_phi = None
        self._owned_tangent_mat: DeviceCSR | None = None
        self._owned_regularized_mat: DeviceCSR | None = None
        self.elastic_values = self._assemble_elastic_values()
        self._reset_timers()

    # ---------------------------------------------------------------- construction helpers
    @classmethod
    def from_mesh(cls, mesh, materials, *, Davis_type: str = "B", fp32_shadow: bool = True):
        geom = geometry(mesh.coord, mesh.elem, mesh.elem_type)
        pat = build_owned_pattern(mesh.elem, mesh.q_mask)
        aff = affine_geometry(mesh.coord, mesh.elem, mesh.elem_type) if mesh.elem_type.upper() == "P2" else None
        dpat = DevicePattern(pat, geom, coord=mesh.coord, affine=aff)
        c0, phi, psi, shear, bulk, lame, gamma = expand_materials(mesh.material, geom.n_q, materials)
        op = cls(dpat, c0, phi, psi, shear, bulk, lame, Davis_type=Davis_type, q_mask=mesh.q_mask,
                 fp32_shadow=fp32_shadow)
        op.gamma = gamma
        op.geom = geom
        return op

    def gravity_load(self, mesh) -> np.ndarray:
        """Volume load f_V with gravity along -y, constrained dofs zeroed (fem/distributed_elastic.py:126-129)."""
        f_int = np.zeros((self.dim, self.n_int))
        f_int[1] = -self.gamma
        f = volume_load(self.geom, mesh.elem, mesh.n_nodes, f_int)
        f[~np.asarray(mesh.q_mask, dtype=bool)] = 0.0
        return f

    def _assemble_elastic_values(self) -> torch.Tensor:
        """K_el on the pattern: the tangent kernel fed the elastic D (lame*vol + 2G*ident)."""
        n = self.n_int
        dev = self._E.device
        zeros = torch.zeros((n, self.n_strain), dtype=torch.float64, device=dev)
        huge = torch.full((n,), 1e300, dtype=torch.float64, device=dev)
        zero = torch.zeros(n, dtype=torch.float64, device=dev)
        if self.dim == 3:
            _S, DS = constitutive_problem_3d(zeros, huge, zero, self.shear, self.lame, return_tangent=True)
        else:
            _S, DS = constitutive_problem_2d(zeros, huge, zero, self.shear, self.bulk, self.lame, return_tangent=True)
        vals = torch.empty(self.pattern.host.nnz, dtype=torch.float64, device=dev)
        self.pattern.tangent(DS, vals)
        return vals

    def _reset_timers(self):
        self.time_reduction = []
        self.time_stress = []
        self.time_stress_tangent = []
        self.time_build_F = []
        self.time_build_F_K_tangent = []
        self.time_build_tangent_local = []
        self.time_local_strain = []
        self.time_local_constitutive = []
        self.time_local_force_assembly = []

    # ---------------------------------------------------------------- U / F conversions
    def _flat_u(self, U) -> tuple[torch.Tensor, str]:
        if isinstance(U, np.ndarray):
            U = np.asarray(U, dtype=np.float64)
            if U.shape[0] != self.dim:
                raise ValueError("U shape first dimension must match dim")
            return to_dev(U.reshape(-1, order="F")), "numpy"
        if isinstance(U, torch.Tensor):
            if U.dim() == 1:
                return U.to(torch.float64).contiguous(), "flat"
            if U.shape[0] != self.dim:
                raise ValueError("U shape first dimension must match dim")
            return U.to(torch.float64).t().contiguous().reshape(-1), "torch"
        raise TypeError("U must be a numpy array or a torch tensor")

    def _field_out(self, flat: torch.Tensor, kind: str):
        if kind == "flat":
            return flat
        field = flat.view(-1, self.dim).t()
        return field.cpu().numpy() if kind == "numpy" else field

    # ---------------------------------------------------------------- reference API
    def reduction(self, lam: float):
        t0 = perf_counter()
        self.c_bar, self.sin_phi = reduction(self.c0, self.phi, self.psi, lam, self.Davis_type)
        self.time_reduction.append(perf_counter() - t0)

    def _evaluate(self, U, tangent: bool):
        if self.c_bar is None or self.sin_phi is None:
            raise ValueError("Material reduction not set. Call reduction(lambda) first.")
        u, kind = self._flat_u(U)
        self._kind = kind
        t0 = perf_counter()
        self.pattern.strain(u, self._E)
        t1 = perf_counter()
        DS = self._DS if tangent else None
        br = self._branch if tangent else None
        if self.dim == 3 and self._sym:
            _lib.call("ssx_constitutive_3d_sym", ptr(self._E), ptr(self.c_bar), ptr(self.sin_phi), ptr(self.shear),
                      ptr(self.lame), self.n_int, ptr(self._S), ptr(self._DSp if tangent else None), ptr(br),
                      stream_ptr())
            self._ds_unpacked = False
        elif self.dim == 3:
            _lib.call("ssx_constitutive_3d", ptr(self._E), ptr(self.c_bar), ptr(self.sin_phi), ptr(self.shear),
                      ptr(self.lame), self.n_int, ptr(self._S), ptr(DS), None, ptr(br), stream_ptr())
        else:
            _lib.call("ssx_constitutive_2d", ptr(self._E), ptr(self.c_bar), ptr(self.sin_phi), ptr(self.shear),
                      ptr(self.bulk), ptr(self.lame), self.n_int, ptr(self._S), ptr(DS), None, None, stream_ptr())
        self.time_local_strain.append(t1 - t0)
        self.time_local_constitutive.append(perf_counter() - t1)
        self._s_ready = True
        if tangent:
            self._ds_ready = True

    @property
    def S(self):
        """Stress as (n_strain(+1 in 2D), n_int), the reference's orientation (a transposed view)."""
        return self._S.t() if self._s_ready else None

    @property
    def DS(self):
        """Consistent tangent as (n_strain^2, n_int), the reference's orientation (a transposed view)."""
        if not self._ds_ready:
            return None
        if self._sym and not self._ds_unpacked:
            if self._DS is None:
                self._DS = torch.empty((self.n_int, 36), dtype=torch.float64, device=self._DSp.device)
            _lib.call("ssx_ds_unpack_sym", self.n_int, ptr(self._DSp), ptr(self._DS), stream_ptr())
            self._ds_unpacked = True
        return self._DS.t()

    def _tangent_input(self):
        """(DS as the tangent kernels take it, packed?)"""
        return (self._DSp, True) if self._sym else (self._DS, False)

    @property
    def E(self):
        return self._E.t()

    def constitutive_problem_stress(self, U):
        t0 = perf_counter()
        self._evaluate(U, tangent=False)
        self.time_stress.append(perf_counter() - t0)
        return self.S

    def constitutive_problem_stress_tangent(self, U):
        t0 = perf_counter()
        self._evaluate(U, tangent=True)
        self.time_stress_tangent.append(perf_counter() - t0)
        return self.S

    def _force(self) -> torch.Tensor:
        if not self._s_ready:
            raise ValueError("Stress tensor not computed")
        t0 = perf_counter()
        S = self._S if self.dim == 3 else self._S[:, :3].contiguous()
        self.pattern.force(S, self._fe, self._F)
        dt = perf_counter() - t0
        self.time_local_force_assembly.append(dt)
        self.time_build_F.append(dt)
        return self._F

    def build_F(self):
        return self._field_out(self._force(), getattr(self, "_kind", "torch"))

    def build_F_local(self):
        F = self._force()
        return F.cpu().numpy() if getattr(self, "_kind", "") == "numpy" else F

    def build_F_free(self):
        F = self._force().index_select(0, self.pattern.free_idx)
        return F.cpu().numpy() if getattr(self, "_kind", "") == "numpy" else F

    def build_F_free_local(self):
        """Internal force at the owned free rows (constitutive/problem.py:1997-2010).  One device
        owns every row, so the owned free rows are all free rows, in free-dof order."""
        return self.build_F_free()

    # flat full-length device vectors (constrained dofs zero) for the device Newton solvers
    def build_F_vec(self) -> torch.Tensor:
        return self._force()

    def build_F_all_vec(self, lam, U) -> torch.Tensor:
        self.reduction(lam)
        self.constitutive_problem_stress(U)
        return self._force()

    def build_F_reduced_vec(self, U) -> torch.Tensor:
        self.constitutive_problem_stress(U)
        return self._force()

    def _tangent_matrix(self) -> DeviceCSR:
        if not self._ds_ready:
            raise ValueError("Tangent DS not computed")
        if self._owned_tangent_mat is None:
            self._owned_tangent_mat = self.pattern.new_csr(with32=False)
        t0 = perf_counter()
        if self._branch is not None:
            # K = K_r at r = 0, with the elastic points in closed form
            DS, sym = self._tangent_input()
            self.pattern.tangent(DS, self._owned_tangent_mat.values, r=0.0, mat=(self.shear, self.lame),
                                 branch=self._branch, sym=sym)
        else:
            self.pattern.tangent(self._DS, self._owned_tangent_mat.values)
        self.time_build_tangent_local.append(perf_counter() - t0)
        return self._owned_tangent_mat

    def build_F_K_tangent(self):
        t0 = perf_counter()
        F = self.build_F()
        K = self._tangent_matrix()
        self.time_build_F_K_tangent.append(perf_counter() - t0)
        return F, K

    def build_K_regularized(self, r: float):
        if not self._ds_ready:
            raise ValueError("Tangent DS not computed")
        if self._owned_regularized_mat is None:
            self._owned_regularized_mat = self.pattern.new_csr(with32=self.fp32_shadow)
        K = self._owned_regularized_mat
        self.last_regularization = float(r)
        t0 = perf_counter()
        if self.pattern.ns is not None:
            # K_r folded into the point tangents: (1-r) DS + r D_el (same operator, no K_el read)
            DS, sym = self._tangent_input()
            self.pattern.tangent(DS, K.values, K.values32, r=float(r), mat=(self.shear, self.lame),
                                 branch=self._branch, sym=sym)
        else:
            self.pattern.tangent(self._DS, K.values, K.values32, r=float(r), elastic=self.elastic_values)
        self.time_build_tangent_local.append(perf_counter() - t0)
        return K

    # composite calls (constitutive/problem.py:2030-2146)
    def build_F_all(self, lam, U):
        self.reduction(lam)
        self.constitutive_problem_stress(U)
        return self.build_F()

    def build_F_all_local(self, lam, U):
        self.reduction(lam)
        self.constitutive_problem_stress(U)
        return self.build_F_local()

    def build_F_all_free(self, lam, U):
        self.reduction(lam)
        self.constitutive_problem_stress(U)
        return self.build_F_free()

    def build_F_all_free_local(self, lam, U):
        """reduction(lam), stress at U, force at the owned free rows (constitutive/problem.py:
        2045-2048); probed by the Newton damping, nonlinear/damping.py:181-182."""
        self.reduction(lam)
        self.constitutive_problem_stress(U)
        return self.build_F_free_local()

    def build_F_K_tangent_all(self, lam, U):
        self.reduction(lam)
        self.constitutive_problem_stress_tangent(U)
        return self.build_F_K_tangent()

    def build_F_K_tangent_all_free(self, lam, U):
        self.reduction(lam)
        self.constitutive_problem_stress_tangent(U)
        return self.build_F_free(), self._tangent_matrix()

    def build_F_reduced(self, U):
        self.constitutive_problem_stress(U)
        return self.build_F()

    def build_F_reduced_local(self, U):
        self.constitutive_problem_stress(U)
        return self.build_F_local()

    def build_F_reduced_free(self, U):
        self.constitutive_problem_stress(U)
        return self.build_F_free()

    def build_F_reduced_free_local(self, U):
        """Stress at U and force at the owned free rows (constitutive/problem.py:2077-2079);
        probed by the line search, nonlinear/damping.py:105-106."""
        self.constitutive_problem_stress(U)
        return self.build_F_free_local()

    def build_F_K_tangent_reduced(self, U):
        self.constitutive_problem_stress_tangent(U)
        return self.build_F_K_tangent()

    def build_F_K_tangent_reduced_free(self, U):
        self.constitutive_problem_stress_tangent(U)
        return self.build_F_free(), self._tangent_matrix()

    def build_F_K_regularized_all(self, lam, U, r):
        self.reduction(lam)
        self.constitutive_problem_stress_tangent(U)
        t0 = perf_counter()
        out = self.build_F(), self.build_K_regularized(r)
        self.time_build_F_K_tangent.append(perf_counter() - t0)
        return out

    def build_F_K_regularized_all_free(self, lam, U, r):
        self.reduction(lam)
        self.constitutive_problem_stress_tangent(U)
        t0 = perf_counter()
        out = self.build_F_free(), self.build_K_regularized(r)
        self.time_build_F_K_tangent.append(perf_counter() - t0)
        return out

    def build_F_K_regularized_reduced(self, U, r):
        self.constitutive_problem_stress_tangent(U)
        t0 = perf_counter()
        out = self.build_F(), self.build_K_regularized(r)
        self.time_build_F_K_tangent.append(perf_counter() - t0)
        return out

    def build_F_K_regularized_reduced_free(self, U, r):
        self.constitutive_problem_stress_tangent(U)
        t0 = perf_counter()
        out = self.build_F_free(), self.build_K_regularized(r)
        self.time_build_F_K_tangent.append(perf_counter() - t0)
        return out

    def elastic_matrix(self) -> DeviceCSR:
        K = self.pattern.new_csr(with32=self.fp32_shadow)
        K.values.copy_(self.elastic_values)
        if K.values32 is not None:
            K.values32.copy_(self.elastic_values.to(torch.float32))
        return K

    def get_total_time(self):
        return {
            "reduction": sum(self.time_reduction),
            "stress": sum(self.time_stress),
            "stress_tangent": sum(self.time_stress_tangent),
            "build_F": sum(self.time_build_F),
            "build_F_K_tangent": sum(self.time_build_F_K_tangent),
            "build_tangent_local": sum(self.time_build_tangent_local),
            "local_strain": sum(self.time_local_strain),
            "local_constitutive": sum(self.time_local_constitutive),
            "local_constitutive_comm": 0.0,
            "local_force_assembly": sum(self.time_local_force_assembly),
            "local_force_gather": 0.0,
            "potential": 0.0,
        }

    def release_petsc_caches(self) -> None:
        """Drop the cached device matrices (the reference releases its PETSc AIJ handles)."""
        self._owned_tangent_mat = None
        self._owned_regularized_mat = None

    def set_owned_tangent_pattern(self, pattern: DevicePattern, **_kwargs) -> None:
        self.pattern = pattern
        self.owned_tangent_pattern = pattern
        self.release_petsc_caches()
