"""Device buffers: torch-ROCm tensors used purely as HBM storage for the ssx kernels, the device
copy of the fixed pattern, and the device CSR matrix handle returned in place of a PETSc AIJ.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .fem import AffineGeometry, Geometry
from .pattern import OwnedPattern


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.SSXError("no GPU visible: the ssx path runs only on a HIP device (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


def to_dev(a, dtype=None) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        t = a
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        return t.to(device()).contiguous()
    arr = np.ascontiguousarray(a)
    t = torch.from_numpy(arr)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device())


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else int(t.data_ptr())


def spmv_lanes(indptr: np.ndarray) -> int:
    """Lanes per row for the CSR kernels.  Rows of 48 or more entries on average (3D P2/P4
    operators) use the aligned-quadruple rows: the power of two nearest mean / 12, between 4 and 64
    (measured best on MI355X: 8-16 lanes for the 77-entry P2 rows and 16 for the 212-entry P4
    rows).  Shorter rows: the smallest power of two >= mean / 4 (scalar rows below 4 lanes)."""
    n = max(int(indptr.size) - 1, 1)
    mean = float(indptr[-1]) / n
    if mean >= 48.0:
        return int(min(64, max(4, 2 ** round(np.log2(mean / 12.0)))))
    lanes = 1
    while lanes < 64 and lanes * 4 < mean:
        lanes *= 2
    return lanes


@dataclass
class DeviceCSR:
    """Square CSR matrix on the device; ``values`` (FP64) is refreshed in place, ``values32`` is the
    FP32 shadow fed to the preconditioner."""

    n_rows: int
    indptr: torch.Tensor     # int32
    indices: torch.Tensor    # int32
    values: torch.Tensor     # float64
    values32: torch.Tensor | None
    lanes: int

    @property
    def nnz(self) -> int:
        return int(self.indices.numel())

    @property
    def shape(self) -> tuple[int, int]:
        return (self.n_rows, self.n_rows)

    def matvec(self, x: torch.Tensor, y: torch.Tensor | None = None, alpha: float = 1.0, beta: float = 0.0):
        if y is None:
            y = torch.empty(self.n_rows, dtype=torch.float64, device=x.device)
        _lib.call("ssx_spmv_f64", self.n_rows, ptr(self.indptr), ptr(self.indices), ptr(self.values), ptr(x), ptr(y),
                  float(alpha), float(beta), self.lanes, stream_ptr())
        return y

    def matvec32(self, x: torch.Tensor, y: torch.Tensor | None = None, alpha: float = 1.0, beta: float = 0.0):
        if self.values32 is None:
            raise ValueError("matrix has no FP32 shadow")
        if y is None:
            y = torch.empty(self.n_rows, dtype=torch.float32, device=x.device)
        _lib.call("ssx_spmv_f32", self.n_rows, ptr(self.indptr), ptr(self.indices), ptr(self.values32), ptr(x),
                  ptr(y), float(alpha), float(beta), self.lanes, stream_ptr())
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values.cpu().numpy(), self.indices.cpu().numpy(), self.indptr.cpu().numpy()),
                             shape=self.shape)


class DevicePattern:
    """The fixed pattern, row slots and per-point geometry resident in HBM."""

    def __init__(self, pat: OwnedPattern, geom: Geometry, coord: np.ndarray | None = None,
                 affine: AffineGeometry | None = None):
        self.host = pat
        self.geom = geom
        self.dim = pat.dim
        self.n_p = pat.n_p
        self.n_q = geom.n_q
        self.n_elem = geom.n_elem
        self.n_int = geom.n_int
        self.n_rows = pat.n_rows
        self.nld = pat.dim * pat.n_p
        self.n_strain = 6 if pat.dim == 3 else 3
        self.indptr = to_dev(pat.indptr)
        self.indices = to_dev(pat.indices)
        self.row_free = to_dev(pat.row_free)
        self.row_slot_ptr = to_dev(pat.row_slot_ptr)
        self.slot_elem = to_dev(pat.slot_elem)
        self.slot_lrow = to_dev(pat.slot_lrow)
        self._slot_pos = None      # uploaded on first use (row-slot tangent path only)
        self.elem_dofs = to_dev(pat.elem_dofs)
        self.free_idx = to_dev(pat.free_idx)
        self.dphi = to_dev(geom.dphi)
        self.weight = to_dev(geom.weight)
        # P2 tetrahedra with a constant Jacobian (11-point rule): the strain, force and tangent
        # kernels form the gradients from 10 numbers per element instead of reading dphi and w
        self.gm = self.gref = None
        if affine is not None and self.dim == 3 and self.n_p == 10 and self.n_q == 11:
            if affine.gm.shape != (self.n_elem, 10) or affine.ref.shape != (3 * self.n_q * self.n_p + self.n_q,):
                raise ValueError("affine factors do not match the mesh")
            self.gm = to_dev(affine.gm)
            self.gref = to_dev(affine.ref)
        self.max_block_nnz = pat.max_block_nnz(64)
        self.lanes = spmv_lanes(pat.indptr)
        self.ns = None
        if self.dim == 3 and self.n_p in (4, 10, 35):
            from .pattern import node_slots
            ns = node_slots(pat)
            self.ns = ns
            self.ns_ptr = to_dev(ns.ptr)
            self.ns_elem = to_dev(ns.elem)
            self.ns_lnode = to_dev(ns.lnode)
            self.ns_off = to_dev(ns.off)
            # one gather launch per row-length class when rows run long (beyond 640 entries the LDS
            # of the longest row, 15 KB, would hold every workgroup to a few per CU; P4 vertex rows
            # reach 2,457 entries against 105 for interior nodes); shorter meshes keep one launch
            groups = ns.buckets() if ns.max_len > 640 else []
            self.ns_groups = [(to_dev(ids), ml) for ids, ml in groups] if len(groups) > 1 else []
            # launch order of the block gather (P2): Morton order of the nodes over the XCDs
            self.gather_ids = None
            if coord is not None and self.n_p == 10 and self.n_q == 11:
                from .pattern import gather_order
                self.gather_ids = to_dev(gather_order(coord))
        self._ke = None
        self._kb = None
        self.force_gather = "nodes"     # "nodes" (node slots, 3D) or "rows" (row slots)

    @property
    def slot_pos(self) -> torch.Tensor:
        if self._slot_pos is None:
            self._slot_pos = to_dev(self.host.slot_pos)
        return self._slot_pos

    def new_csr(self, with32: bool = True) -> DeviceCSR:
        dev = self.indptr.device
        return DeviceCSR(self.n_rows, self.indptr, self.indices,
                         torch.zeros(self.host.nnz, dtype=torch.float64, device=dev),
                         torch.zeros(self.host.nnz, dtype=torch.float32, device=dev) if with32 else None,
                         self.lanes)

    # -- kernels -----------------------------------------------------------------------------
    # "affine" (per-element factors, when the mesh has them) or "points" (per-point dphi and w)
    geometry_form = "affine"

    @property
    def affine(self) -> bool:
        return self.gm is not None and self.geometry_form == "affine"

    def strain(self, u: torch.Tensor, E: torch.Tensor) -> None:
        if self.affine:
            _lib.call("ssx_strain_affine_3d", self.n_p, ptr(self.gm), ptr(self.gref), ptr(self.elem_dofs), ptr(u),
                      self.n_elem, self.n_q, ptr(E), stream_ptr())
            return
        _lib.call("ssx_strain", self.dim, self.n_p, ptr(self.dphi), ptr(self.elem_dofs), ptr(u), self.n_elem,
                  self.n_q, ptr(E), stream_ptr())

    def force(self, S: torch.Tensor, fe: torch.Tensor, F: torch.Tensor) -> None:
        if self.affine:
            _lib.call("ssx_force_elem_affine_3d", self.n_p, ptr(self.gm), ptr(self.gref), ptr(S), self.n_elem,
                      self.n_q, ptr(fe), stream_ptr())
        else:
            _lib.call("ssx_force_elem", self.dim, self.n_p, ptr(self.dphi), ptr(S), ptr(self.weight), self.n_elem,
                      self.n_q, ptr(fe), stream_ptr())
        if self.ns is not None and self.force_gather == "nodes":
            _lib.call("ssx_force_gather_nodes", self.dim, self.nld, ptr(fe), ptr(self.ns_ptr), ptr(self.ns_elem),
                      ptr(self.ns_lnode), ptr(self.row_free), self.host.n_nodes, ptr(F), stream_ptr())
        else:
            _lib.call("ssx_force_gather", ptr(fe), ptr(self.row_slot_ptr), ptr(self.slot_elem), ptr(self.slot_lrow),
                      self.n_rows, self.nld, ptr(F), stream_ptr())

    def _ke_buffer(self, dev) -> torch.Tensor:
        if self._ke is None:
            self._ke = torch.empty((self.n_elem, self.nld, self.nld), dtype=torch.float64, device=dev)
        return self._ke

    def element_matrices_p4(self, DS: torch.Tensor, r: float = 0.0, mat=None) -> torch.Tensor:
        """(n_elem, 105, 105) element tangents of a P4 mesh on the FP64 matrix cores; with
        mat = (shear, lame) per point, the regularised (1-r) K + r K_el element matrices."""
        ke = self._ke_buffer(DS.device)
        sh, la = (None, None) if mat is None else mat
        _lib.call("ssx_element_matrices_p4", ptr(self.dphi), ptr(DS), ptr(self.weight), self.n_elem, float(r),
                  ptr(sh), ptr(la), ptr(ke), stream_ptr())
        return ke

    def element_matrices_3d(self, DS: torch.Tensor, r: float = 0.0, mat=None) -> torch.Tensor:
        """(n_elem, nld, nld) element tangents of a P1/P2 tetrahedral mesh (one wavefront per element)."""
        ke = self._ke_buffer(DS.device)
        sh, la = (None, None) if mat is None else mat
        _lib.call("ssx_element_matrices_3d", self.n_p, ptr(self.dphi), ptr(DS), ptr(self.weight), self.n_elem,
                  self.n_q, float(r), ptr(sh), ptr(la), ptr(ke), stream_ptr())
        return ke

    def element_blocks_3d(self, DS: torch.Tensor, r: float = 0.0, mat=None,
                          branch: torch.Tensor | None = None, sym: bool = False) -> torch.Tensor:
        """(n_elem, 55, 3, 3) upper node-pair blocks of the P2 element tangents; `branch` = the return
        mapping's per-point branch ids for this DS (elastic points then take the closed form); `sym`:
        DS is the packed (n_int, 21) symmetric tangent (constant-Jacobian meshes only)."""
        if sym and not self.affine:
            raise ValueError("a packed symmetric DS needs the constant-Jacobian form of the mesh")
        if self._kb is None:
            self._kb = torch.empty((self.n_elem, 55, 3, 3), dtype=torch.float64, device=DS.device)
        sh, la = (None, None) if mat is None else mat
        if self.affine:
            _lib.call("ssx_element_blocks_affine_3d", self.n_p, ptr(self.gm), ptr(self.gref), ptr(DS), int(sym),
                      self.n_elem, self.n_q, float(r), ptr(sh), ptr(la), ptr(branch), ptr(self._kb), stream_ptr())
        else:
            _lib.call("ssx_element_blocks_3d", self.n_p, ptr(self.dphi), ptr(DS), ptr(self.weight), self.n_elem,
                      self.n_q, float(r), ptr(sh), ptr(la), ptr(branch), ptr(self._kb), stream_ptr())
        return self._kb

    def gather_blocks(self, kb: torch.Tensor, out: torch.Tensor, out32: torch.Tensor | None = None, r: float = 0.0,
                      elastic: torch.Tensor | None = None) -> None:
        _lib.call("ssx_gather_node_blocks", self.n_p, ptr(kb), ptr(self.ns_ptr), ptr(self.ns_elem),
                  ptr(self.ns_lnode), ptr(self.ns_off), ptr(self.indptr), ptr(self.row_free), ptr(self.gather_ids),
                  self.host.n_nodes, self.ns.max_len, float(r), ptr(elastic), ptr(out), ptr(out32), stream_ptr())

    def gather_nodes(self, ke: torch.Tensor, out: torch.Tensor, out32: torch.Tensor | None = None, r: float = 0.0,
                     elastic: torch.Tensor | None = None) -> None:
        groups = self.ns_groups or [(None, self.ns.max_len)]
        for ids, max_len in groups:
            n = self.host.n_nodes if ids is None else ids.numel()
            _lib.call("ssx_gather_node_rows", self.dim, self.nld, ptr(ke), ptr(self.ns_ptr), ptr(self.ns_elem),
                      ptr(self.ns_lnode), ptr(self.ns_off), ptr(self.indptr), ptr(self.row_free), ptr(ids), n,
                      max_len, float(r), ptr(elastic), ptr(out), ptr(out32), stream_ptr())

    # 3D tetrahedra: "nodes" (element matrices + node gather, the default; P2 meshes take the
    # symmetric node-pair blocks instead of full element matrices) or "rows" (fused row kernel,
    # the 2D path)
    tangent_method = "nodes"

    def tangent(self, DS: torch.Tensor, out: torch.Tensor, out32: torch.Tensor | None = None, r: float = 0.0,
                elastic: torch.Tensor | None = None, mat=None, branch: torch.Tensor | None = None,
                sym: bool = False) -> None:
        """Tangent values on the pattern.  K_r = (1-r) K + r K_el either from the assembled elastic
        values (`elastic`) or folded into the point tangents (`mat` = (shear, lame) per point).
        `branch` (P2 tetrahedra, with `mat`): the return mapping's branch ids for this DS.  `sym`:
        DS is packed symmetric (n_int, 21) (P2 tetrahedra in the constant-Jacobian form)."""
        if sym and not (self.affine and self.ns is not None and self.tangent_method == "nodes"):
            raise ValueError("a packed symmetric DS needs the P2 block path in the constant-Jacobian form")
        if mat is not None and self.ns is None:
            raise ValueError("the K_r fold needs the node-gather path (3D tetrahedra)")
        if branch is not None and mat is None:
            raise ValueError("branch ids need the material (mat = (shear, lame))")
        if self.ns is not None and (self.n_p == 35 or self.tangent_method == "nodes"):
            if self.n_p == 10 and self.n_q == 11:
                kb = self.element_blocks_3d(DS, r, mat, branch, sym)
                self.gather_blocks(kb, out, out32, r, None if mat is not None else elastic)
                return
            if self.n_p == 35:
                ke = self.element_matrices_p4(DS, r, mat)
            else:
                ke = self.element_matrices_3d(DS, r, mat)
            self.gather_nodes(ke, out, out32, r, None if mat is not None else elastic)
            return
        if mat is not None:
            raise ValueError("the K_r fold needs the node-gather path")
        _lib.call("ssx_tangent_rows", self.dim, self.n_p, ptr(self.dphi), ptr(DS), ptr(self.weight),
                  ptr(self.row_slot_ptr), ptr(self.slot_elem), ptr(self.slot_lrow), ptr(self.slot_pos),
                  ptr(self.indptr), ptr(self.row_free), self.n_rows, self.n_elem, self.n_q, float(r),
                  ptr(elastic), ptr(out), ptr(out32), self.max_block_nnz, stream_ptr())
