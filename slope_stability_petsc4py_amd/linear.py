"""Device linear solvers: deflated flexible GMRES (FP64) with an FP32 or FP64 p-multigrid or
Chebyshev-Jacobi preconditioner.

Algorithm (linear/deflated_fgmres.py:174-295, the MATLAB-exact variant): coarse initial guess
x0 = W (W^T b) from an A-orthonormal deflation basis W; each Arnoldi step applies the
preconditioner, projects z <- z - W ((AW)^T z), multiplies by A and orthogonalises against the
Krylov basis; the small least-squares problem on the Hessenberg matrix runs on the host; stop at
relative residual < rtol (default 1e-1) or maxit (100).  The deflation basis follows
linear/solver.py:2749-2852 (A-orthogonalised newest-first with the 1e-3 threshold of
linear/orthogonalize.py:49-170, capped at 48 vectors, linear/solver.py:2288-2304).

Differences, by design for the device: Gram-Schmidt is classical with one re-orthogonalisation
pass (two batched dot launches per step instead of j+1 dependent ones); vectors are full-length
(constrained dofs carry zeros and unit diagonal rows, so the free block is solved exactly as the
reference's free-dof system).

Preconditioners
  * ``PMultigrid``: V-cycle P2 -> P1 (vertex dofs), Chebyshev(Jacobi) smoothing in FP32
    (values32) or FP64, Galerkin coarse operator P^T A P formed on the device, inverted densely
    by the library's blocked Cholesky, triangular inverse and W^T W in the preconditioner's
    precision (a coarse solve is one symmetric matrix-vector product).  Replaces
    Hypre BoomerAMG / the manual p-MG shell (linear/solver.py:96-779).  Above a size limit the P1
    level stays sparse (Galerkin product on the P1 pattern, Chebyshev smoothing) and the dense
    level below it holds rigid-body modes on node aggregates.
  * ``ChebyshevJacobi``: polynomial preconditioner alone.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from . import _lib
from .device import DeviceCSR, ptr, spmv_lanes, stream_ptr


def _st():
    return stream_ptr()


# ---------------------------------------------------------------------------------------------
# small device vector helpers (all through the ssx kernels)
# ---------------------------------------------------------------------------------------------

class _Scratch:
    def __init__(self, dev):
        self.part = torch.empty(1024, dtype=torch.float64, device=dev)
        self.cols = torch.empty(64 * 160, dtype=torch.float64, device=dev)   # dot_cols partials
        self.out = torch.empty(128, dtype=torch.float64, device=dev)

    def cols_for(self, k: int) -> torch.Tensor:
        """Partials scratch of ssx_dot_cols for k columns (64 doubles per column)."""
        if self.cols.numel() < 64 * k:
            self.cols = torch.empty(64 * k, dtype=torch.float64, device=self.cols.device)
        return self.cols


def dot(x: torch.Tensor, y: torch.Tensor, scr: _Scratch) -> float:
    _lib.call("ssx_dot", x.numel(), ptr(x), ptr(y), ptr(scr.part), ptr(scr.out), _st())
    return float(scr.out[0].item())


def axpby(a: float, x: torch.Tensor, b: float, y: torch.Tensor) -> None:
    if y.dtype == torch.float64:
        _lib.call("ssx_axpby_f64", y.numel(), float(a), ptr(x), float(b), ptr(y), _st())
    else:
        _lib.call("ssx_axpby_f32", y.numel(), float(a), ptr(x), float(b), ptr(y), _st())


def spd_inverse_work(n: int) -> int:
    """Scratch elements (of the computation precision) of `ssx_spd_inverse_tiles` for n x n."""
    import ctypes
    out = ctypes.c_int64(0)
    _lib.call("ssx_spd_inverse_sizes", int(n), ctypes.addressof(out))
    return int(out.value)


def sym_tile_sizes(n: int) -> tuple[int, int]:
    """Element counts of the symmetric tile store of an n x n matrix and of its apply's partials."""
    import ctypes
    out = (ctypes.c_int64 * 2)()
    _lib.call("ssx_sym_tiles_sizes", int(n), ctypes.addressof(out), ctypes.addressof(out) + 8)
    return int(out[0]), int(out[1])


def spmv(A: DeviceCSR, x: torch.Tensor, y: torch.Tensor, alpha=1.0, beta=0.0, use32: bool = False) -> None:
    if use32:
        _lib.call("ssx_spmv_f32", A.n_rows, ptr(A.indptr), ptr(A.indices), ptr(A.values32), ptr(x), ptr(y),
                  float(alpha), float(beta), A.lanes, _st())
    else:
        _lib.call("ssx_spmv_f64", A.n_rows, ptr(A.indptr), ptr(A.indices), ptr(A.values), ptr(x), ptr(y),
                  float(alpha), float(beta), A.lanes, _st())


@dataclass
class IterationCollector:
    """Per-solve statistics (linear/collector.py:14-73)."""

    iterations: list = field(default_factory=list)
    solve_time: list = field(default_factory=list)
    preconditioner_time: list = field(default_factory=list)
    orthogonalization_time: list = field(default_factory=list)
    setup_time: list = field(default_factory=list)

    def store(self, it, t_solve, t_pc, t_orth):
        self.iterations.append(int(it))
        self.solve_time.append(float(t_solve))
        self.preconditioner_time.append(float(t_pc))
        self.orthogonalization_time.append(float(t_orth))

    def get_total_iterations(self):
        return int(sum(self.iterations))

    def get_total_solve_time(self):
        return float(sum(self.solve_time))

    def get_total_preconditioner_time(self):
        return float(sum(self.preconditioner_time))

    def get_total_orthogonalization_time(self):
        return float(sum(self.orthogonalization_time))

    def summary(self) -> dict:
        return {"solves": len(self.iterations), "iterations": self.get_total_iterations(),
                "setups": len(self.setup_time), "setup_s": float(sum(self.setup_time)),
                "solve_s": self.get_total_solve_time(), "preconditioner_s": self.get_total_preconditioner_time(),
                "orthogonalization_s": self.get_total_orthogonalization_time()}


# ---------------------------------------------------------------------------------------------
# Chebyshev smoother on one level
# ---------------------------------------------------------------------------------------------

class _Level:
    """One multigrid level: CSR operator (FP32 or FP64 values), inverse diagonal, spectrum bound."""

    def __init__(self, A: DeviceCSR, use32: bool):
        self.A = A
        self.use32 = use32
        self.dtype = torch.float32 if use32 else torch.float64
        dev = A.indptr.device
        n = A.n_rows
        self.n = n
        self.dinv = torch.empty(n, dtype=self.dtype, device=dev)
        self.r = torch.empty(n, dtype=self.dtype, device=dev)
        self.d = torch.empty(n, dtype=self.dtype, device=dev)
        self.x = torch.empty(n, dtype=self.dtype, device=dev)
        self.xb = torch.empty(n, dtype=self.dtype, device=dev)    # second iterate buffer (fused steps)
        self.lmax = 2.0
        self.fused = True
        self._v = None      # last power-iteration iterate (normalised), the next refresh's start

    def refresh(self, power_its: int = 12, seed: int = 0, warm_its: int = 4):
        A = self.A
        fn = "ssx_inv_diag_f32" if self.use32 else "ssx_inv_diag_f64"
        vals = A.values32 if self.use32 else A.values
        _lib.call(fn, A.n_rows, 0, ptr(A.indptr), ptr(A.indices), ptr(vals), ptr(self.dinv), _st())
        # power iteration on D^-1 A without per-step normalisation (the iterate grows by at most the
        # spectral radius per step, far inside the FP32 range over a dozen steps): the estimate is
        # |w_k| / |w_{k-1}|, the same quotient the normalised iteration forms, from two norms and one
        # host read.  The first refresh starts from a seeded random vector; later ones start from
        # the previous refresh's last iterate (the operator changes little between Newton steps)
        # and take `warm_its` steps
        if power_its < 1:
            self.lmax = 1.0
            return
        if self._v is None:
            g = torch.Generator(device="cpu").manual_seed(seed)
            v = torch.rand(self.n, generator=g, dtype=torch.float64).to(self.dinv.device).to(self.dtype)
            v /= torch.linalg.vector_norm(v)
        else:
            v = self._v
            power_its = max(1, min(power_its, warm_its))
        src, dst = v, self.r
        for _ in range(power_its - 1):
            spmv(A, src, dst, use32=self.use32)
            dst.mul_(self.dinv)
            src, dst = dst, (self.d if dst is self.r else self.r)
        prev = torch.linalg.vector_norm(src)
        spmv(A, src, dst, use32=self.use32)
        dst.mul_(self.dinv)
        nrm = torch.linalg.vector_norm(dst)
        ratio = nrm / torch.clamp(prev, min=torch.finfo(self.dtype).tiny)
        lam = float(ratio.item())
        self.lmax = 1.1 * lam if lam > 0.0 and np.isfinite(lam) else 1.0
        self._v = (dst / nrm) if lam > 0.0 and np.isfinite(lam) else None

    def cheb(self, b: torch.Tensor, x: torch.Tensor, degree: int, lo_frac: float, zero_guess: bool):
        """`degree` Chebyshev(Jacobi) steps on A x = b over [lo_frac*lmax, lmax]."""
        if self.fused:
            self._cheb_fused(b, x, degree, lo_frac, zero_guess)
            return
        beta = self.lmax
        alpha = lo_frac * self.lmax
        theta = 0.5 * (beta + alpha)
        delta = 0.5 * (beta - alpha)
        sigma = theta / delta
        rho = 1.0 / sigma
        upd = "ssx_cheb_update_f32" if self.use32 else "ssx_cheb_update_f64"
        # r = b - A x
        axpby(1.0, b, 0.0, self.r)
        if not zero_guess:
            spmv(self.A, x, self.r, alpha=-1.0, beta=1.0, use32=self.use32)
        else:
            x.zero_()
        self.d.zero_()
        _lib.call(upd, self.n, ptr(self.dinv), ptr(self.r), 1.0 / theta, 0.0, ptr(x), ptr(self.d), _st())
        for _ in range(degree - 1):
            rho_new = 1.0 / (2.0 * sigma - rho)
            axpby(1.0, b, 0.0, self.r)
            spmv(self.A, x, self.r, alpha=-1.0, beta=1.0, use32=self.use32)
            _lib.call(upd, self.n, ptr(self.dinv), ptr(self.r), 2.0 * rho_new / delta, rho_new * rho, ptr(x),
                      ptr(self.d), _st())
            rho = rho_new

    def _cheb_fused(self, b: torch.Tensor, x: torch.Tensor, degree: int, lo_frac: float, zero_guess: bool):
        """The same steps as `cheb`, one launch each (`ssx_cheb_step_*`: residual, smoother and update
        fused), the iterate alternating between x and a second buffer so that the last step writes x."""
        beta = self.lmax
        alpha = lo_frac * self.lmax
        theta = 0.5 * (beta + alpha)
        delta = 0.5 * (beta - alpha)
        sigma = theta / delta
        rho = 1.0 / sigma
        A = self.A
        fn = "ssx_cheb_step_f32" if self.use32 else "ssx_cheb_step_f64"
        vals = A.values32 if self.use32 else A.values
        other = {id(x): self.xb, id(self.xb): x}
        # destination of step k: x when degree - 1 - k is even
        dst = x if (degree - 1) % 2 == 0 else self.xb
        src = None
        if not zero_guess:
            if dst is x:                 # the first step must not read the buffer it writes
                self.xb.copy_(x)
                src = self.xb
            else:
                src = x
        c, a = 1.0 / theta, 0.0
        for k in range(degree):
            if k:
                rho_new = 1.0 / (2.0 * sigma - rho)
                c, a = 2.0 * rho_new / delta, rho_new * rho
                rho = rho_new
            _lib.call(fn, A.n_rows, ptr(A.indptr), ptr(A.indices), ptr(vals), ptr(b), ptr(src), ptr(dst),
                      ptr(self.d), ptr(self.dinv), c, a, 1 if k == 0 else 0, A.lanes, _st())
            src, dst = dst, other[id(dst)]


class ChebyshevJacobi:
    """Polynomial preconditioner: `degree` Chebyshev(Jacobi) steps from a zero guess."""

    def __init__(self, degree: int = 8, use32: bool = True, lo_frac: float = 0.02):
        self.degree = degree
        self.use32 = use32
        self.lo_frac = lo_frac
        self.level = None

    def setup(self, A: DeviceCSR, mesh_info=None):
        if self.level is None or self.level.A is not A:
            self.level = _Level(A, self.use32)
        self.level.refresh()

    def apply(self, r: torch.Tensor, z: torch.Tensor) -> None:
        L = self.level
        if self.use32:
            b = torch.empty(L.n, dtype=torch.float32, device=r.device)
            _lib.call("ssx_cast_f64_f32", L.n, ptr(r), ptr(b), _st())
            L.cheb(b, L.x, self.degree, self.lo_frac, zero_guess=True)
            _lib.call("ssx_cast_f32_f64", L.n, ptr(L.x), ptr(z), _st())
        else:
            L.cheb(r, z, self.degree, self.lo_frac, zero_guess=True)


# ---------------------------------------------------------------------------------------------
# p-multigrid P2 -> P1
# ---------------------------------------------------------------------------------------------

_EDGES = {3: ((0, 1), (1, 2), (0, 2), (1, 3), (2, 3), (0, 3)), 2: ((1, 2), (0, 2), (0, 1))}


def p2_to_p1_prolongation(elem: np.ndarray, q_mask: np.ndarray):
    """Rows of P (fine dof -> coarse dof, weight) for the P2 -> P1 interpolation on free dofs.

    Vertex nodes inject, edge-midpoint nodes average their two vertices (the P1 basis evaluated at
    the P2 nodes, linear/pmg.py:365-436).  Constrained fine dofs get empty rows; constrained coarse
    dofs are left out of P (they carry an identity block in the coarse operator)."""
    elem = np.asarray(elem, dtype=np.int64)
    dim, n_nodes = q_mask.shape
    nv = dim + 1
    verts = np.unique(elem[:nv].reshape(-1))
    cnum = np.full(n_nodes, -1, dtype=np.int64)
    cnum[verts] = np.arange(verts.size)
    parents = {}
    for k, (a, b) in enumerate(_EDGES[dim]):
        mids = elem[nv + k]
        for m, va, vb in zip(mids, elem[a], elem[b]):
            parents[int(m)] = (int(va), int(vb))
    free = q_mask.T.reshape(-1)
    n_c_nodes = verts.size
    rows_ptr = np.zeros(dim * n_nodes + 1, dtype=np.int64)
    idx, w = [], []
    for node in range(n_nodes):
        for c in range(dim):
            f = dim * node + c
            ent = []
            if free[f]:
                if cnum[node] >= 0:
                    ent = [(dim * cnum[node] + c, 1.0)]
                elif node in parents:
                    va, vb = parents[node]
                    ent = [(dim * cnum[va] + c, 0.5), (dim * cnum[vb] + c, 0.5)]
            for (J, wt) in ent:
                if free[dim * verts[J // dim] + c]:
                    idx.append(J)
                    w.append(wt)
            rows_ptr[f + 1] = len(idx)
    c_free = np.zeros(dim * n_c_nodes, dtype=bool)
    c_free[(dim * np.arange(n_c_nodes)[:, None] + np.arange(dim)).reshape(-1)] = free[
        (dim * verts[:, None] + np.arange(dim)).reshape(-1)]
    return rows_ptr, np.asarray(idx, dtype=np.int64), np.asarray(w, dtype=np.float64), dim * n_c_nodes, c_free


_P4_TO_P2_LOCAL = (0, 1, 2, 3, 5, 8, 11, 14, 17, 20)  # P4 local nodes that are the P2 nodes


def p4_to_p2_level(coord, elem35, q_mask):
    """The P2 mesh inside a P4 mesh and the P4 <- P2 interpolation (P2 basis at the P4 nodes,
    linear/pmg.py:365-436 for the same construction between P2 and P1).

    Returns (coord2, elem2, q_mask2, P rows (ptr, idx, w) over fine dofs, n_c dofs)."""
    from .fem import basis_3d, tetra_node_tuples
    elem35 = np.asarray(elem35, dtype=np.int64)
    dim, n_f_nodes = q_mask.shape
    sub = elem35[list(_P4_TO_P2_LOCAL)]
    nodes2 = np.unique(sub.reshape(-1))
    num2 = np.full(n_f_nodes, -1, dtype=np.int64)
    num2[nodes2] = np.arange(nodes2.size)
    elem2 = num2[sub]
    coord2 = coord[:, nodes2]
    q2 = q_mask[:, nodes2]
    tup = np.array(tetra_node_tuples(4), dtype=np.float64)[:, 1:].T / 4.0   # (3, 35) reference coords
    phi = basis_3d("P2", tup)[0]                                             # (10, 35)
    free_f = q_mask.T.reshape(-1)
    free_c = q2.T.reshape(-1)
    seen = np.zeros(n_f_nodes, dtype=bool)
    rows = {}
    for k in range(35):
        nodes_k = elem35[k]
        new = ~seen[nodes_k]
        seen[nodes_k] = True
        idx = np.flatnonzero(new)
        for e in idx:
            wts = phi[:, k]
            nz = np.flatnonzero(np.abs(wts) > 1e-14)
            rows[int(nodes_k[e])] = (elem2[nz, e], wts[nz])
    p_ptr = np.zeros(dim * n_f_nodes + 1, dtype=np.int64)
    pi, pw = [], []
    for node in range(n_f_nodes):
        cn, wts = rows[node]
        for c in range(dim):
            f = dim * node + c
            if free_f[f]:
                for J, wt in zip(cn, wts):
                    if free_c[dim * J + c]:
                        pi.append(dim * int(J) + c)
                        pw.append(float(wt))
            p_ptr[f + 1] = len(pi)
    return coord2, elem2, q2, (p_ptr, np.asarray(pi, dtype=np.int64), np.asarray(pw)), dim * nodes2.size


class _Transfer:
    """Prolongation P (fine x coarse) and restriction P^T on the device, FP64 and FP32 values."""

    def __init__(self, p_ptr, p_idx, p_w, n_f, n_c, dev):
        import scipy.sparse as sp
        P = sp.csr_matrix((p_w, p_idx, p_ptr), shape=(n_f, n_c))
        PT = P.T.tocsr()
        mk = lambda M: (torch.from_numpy(M.indptr.astype(np.int32)).to(dev),  # noqa: E731
                        torch.from_numpy(M.indices.astype(np.int32)).to(dev),
                        torch.from_numpy(M.data.astype(np.float64)).to(dev))
        self.P = DeviceCSR(n_f, *mk(P), None, spmv_lanes(P.indptr))
        self.P.values32 = self.P.values.to(torch.float32)
        self.PT = DeviceCSR(n_c, *mk(PT), None, spmv_lanes(PT.indptr))
        self.PT.values32 = self.PT.values.to(torch.float32)
        self.p_ptr = torch.from_numpy(p_ptr.astype(np.int32)).to(dev)
        self.p_idx = torch.from_numpy(p_idx.astype(np.int32)).to(dev)
        self.p_w = torch.from_numpy(np.asarray(p_w, dtype=np.float64)).to(dev)


def galerkin_weight_bound(p_ptr: np.ndarray, p_idx: np.ndarray, p_w: np.ndarray, n_c: int, max_row_len: int) -> float:
    """Bound on sum |p_iI| |p_jJ| over the products of one coarse entry of P^T A P: the largest
    column sum of |P| times the longest row of A times max |p| (the fixed-point scale of the
    Galerkin kernels rests on it)."""
    aw = np.abs(np.asarray(p_w, dtype=np.float64))
    if aw.size == 0:
        return 1.0
    colsum = np.bincount(np.asarray(p_idx, dtype=np.int64), weights=aw, minlength=int(n_c))
    return float(max(colsum.max(), 1e-300) * max(int(max_row_len), 1) * aw.max())


class PMultigridP4:
    """Three-level V-cycle P4 -> P2 -> P1 for order-4 tetrahedra.

    The P2 level operator is rediscretised from the current tangent: the P2 basis integrated at the
    P4 quadrature points with the same DS and regularisation (K_r at the P2 level); the P1 level is
    the dense Galerkin product of the P2 level, factorised by Cholesky (as `PMultigrid`).  Chebyshev
    (Jacobi) smoothing on the P4 and P2 levels, FP32 or FP64."""

    def __init__(self, mesh, op, *, use32: bool = True, degree: int = 3, lo_frac: float = 0.1,
                 coarse_refresh: int = 0, refresh_iterations: int = 20, refresh_ratio: float = 1.5,
                 coarse: str = "auto"):
        from .device import DevicePattern
        from .fem import geometry, quadrature_3d
        from .pattern import build_owned_pattern
        self.op = op
        self.use32 = use32
        self.degree = degree
        self.lo_frac = lo_frac
        coord2, elem2, q2, prows, n2 = p4_to_p2_level(mesh.coord, mesh.elem, mesh.q_mask)
        geom2 = geometry(coord2, elem2, "P2", quad=quadrature_3d("P4"))
        pat2 = build_owned_pattern(elem2, q2)
        self.pat2 = DevicePattern(pat2, geom2)
        dev = self.pat2.indptr.device
        self.t42 = _Transfer(*prows, mesh.q_mask.size, n2, dev)
        # the P2 level shares the P4 points (same rule), so K_r at that level folds the same
        # per-point shear and lame into its element matrices
        self.A2 = self.pat2.new_csr(with32=True)
        # the P1 level below: dense, or (large meshes, coarse="auto" above 50k free P1 dofs, or
        # coarse="aggregate") sparse with the rigid-body-mode aggregate level, which needs the P2
        # level's node coordinates
        self.coarse = PMultigrid(elem2, q2, use32=use32, degree=degree, lo_frac=lo_frac,
                                 coarse_refresh=coarse_refresh, refresh_iterations=refresh_iterations,
                                 refresh_ratio=refresh_ratio, coarse=coarse, coord=coord2)
        self.note_iterations = self.coarse.note_iterations
        self.timings = self.coarse.timings
        self.level = None
        self.r = 1e-4

    def setup(self, A: DeviceCSR, mesh_info=None):
        if self.level is None or self.level.A is not A:
            self.level = _Level(A, self.use32)
        self.level.refresh()
        r = float(getattr(self.op, "last_regularization", self.r))
        self.pat2.tangent(self.op._DS, self.A2.values, self.A2.values32, r=r, mat=(self.op.shear, self.op.lame))
        self.coarse.setup(self.A2)
        dt = torch.float32 if self.use32 else torch.float64
        dev = A.indptr.device
        self.b2 = torch.empty(self.A2.n_rows, dtype=dt, device=dev)
        self.x2 = torch.empty(self.A2.n_rows, dtype=dt, device=dev)

    def apply(self, r: torch.Tensor, z: torch.Tensor) -> None:
        Lv = self.level
        n = Lv.n
        if self.use32:
            b = torch.empty(n, dtype=torch.float32, device=r.device)
            _lib.call("ssx_cast_f64_f32", n, ptr(r), ptr(b), _st())
        else:
            b = r
        x = Lv.x
        Lv.cheb(b, x, self.degree, self.lo_frac, zero_guess=True)
        res = torch.empty_like(x)
        axpby(1.0, b, 0.0, res)
        spmv(Lv.A, x, res, alpha=-1.0, beta=1.0, use32=self.use32)
        spmv(self.t42.PT, res, self.b2, use32=self.use32)
        self.coarse.cycle(self.b2, self.x2)
        spmv(self.t42.P, self.x2, x, alpha=1.0, beta=1.0, use32=self.use32)
        Lv.cheb(b, x, self.degree, self.lo_frac, zero_guess=False)
        if self.use32:
            _lib.call("ssx_cast_f32_f64", n, ptr(x), ptr(z), _st())
        else:
            z.copy_(x)


def rbm_aggregates(coord: np.ndarray, elem: np.ndarray, free: np.ndarray, radius: int = 2):
    """Tentative prolongator of rigid-body modes on node aggregates (unsmoothed aggregation).

    Aggregates: every node whose graph neighbourhood of the given radius (node adjacency of the
    elements `elem`, vertex numbering) is still unassigned becomes a root and takes that whole
    neighbourhood; the remaining nodes then join the aggregate most of their neighbours belong to,
    sweep by sweep.  Columns: per aggregate, an orthonormal basis (SVD, relative cut 1e-10) of the
    rigid-body modes (dim translations and dim(dim-1)/2 rotations about the aggregate's centroid)
    restricted to the free dofs (dof = dim*node + comp), so constrained dofs get empty rows and a
    partly constrained aggregate keeps only the modes its free dofs can carry.

    Returns (t_ptr, t_idx, t_w, n_t): the CSR rows of T over the dim*n_nodes dofs."""
    import scipy.sparse as sp
    coord = np.asarray(coord, dtype=np.float64)
    elem = np.asarray(elem, dtype=np.int64)
    free = np.asarray(free, dtype=bool)
    dim, n = coord.shape
    nv = elem.shape[0]
    ri = np.repeat(elem.T, nv, axis=1).reshape(-1)
    ci = np.tile(elem.T, (1, nv)).reshape(-1)
    G = sp.csr_matrix((np.ones(ri.size, dtype=np.float32), (ri, ci)), shape=(n, n))
    G.sum_duplicates()
    G.data[:] = 1.0
    H = G
    for _ in range(max(1, int(radius)) - 1):
        H = H @ G
        H.data[:] = 1.0
    agg = np.full(n, -1, dtype=np.int64)
    na = 0
    hp, hi = H.indptr, H.indices
    for i in range(n):
        if agg[i] >= 0:
            continue
        nb = hi[hp[i]:hp[i + 1]]
        if (agg[nb] < 0).all():
            agg[nb] = na
            na += 1
    gp, gi = G.indptr, G.indices
    while (agg < 0).any():
        left = np.flatnonzero(agg < 0)
        new = agg.copy()
        for i in left:
            a = agg[gi[gp[i]:gp[i + 1]]]
            a = a[a >= 0]
            if a.size:
                new[i] = np.bincount(a).argmax()
        if (new == agg).all():           # nodes with no assigned neighbour: aggregates of their own
            new[left] = na + np.arange(left.size)
            na += left.size
        agg = new
    order = np.argsort(agg, kind="stable")
    starts = np.searchsorted(agg[order], np.arange(na + 1))
    n_rot = dim * (dim - 1) // 2
    m = dim + n_rot
    rows, cols, vals = [], [], []
    n_t = 0
    for a in range(na):
        nodes = order[starts[a]:starts[a + 1]]
        d = coord[:, nodes] - coord[:, nodes].mean(axis=1, keepdims=True)
        R = np.zeros((nodes.size, dim, m))
        for c in range(dim):
            R[:, c, c] = 1.0
        if dim == 3:
            R[:, 1, 3], R[:, 2, 3] = -d[2], d[1]      # about x
            R[:, 0, 4], R[:, 2, 4] = d[2], -d[0]      # about y
            R[:, 0, 5], R[:, 1, 5] = -d[1], d[0]      # about z
        else:
            R[:, 0, 2], R[:, 1, 2] = -d[1], d[0]
        dofs = (dim * nodes[:, None] + np.arange(dim)).reshape(-1)
        f = free[dofs]
        if not f.any():
            continue
        U, s, _ = np.linalg.svd(R.reshape(-1, m)[f], full_matrices=False)
        k = int((s > 1e-10 * s[0]).sum()) if s[0] > 0 else 0
        fd = dofs[f]
        rows.append(np.repeat(fd, k))
        cols.append(np.tile(n_t + np.arange(k), fd.size))
        vals.append(U[:, :k].reshape(-1))
        n_t += k
    rows = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
    cols = np.concatenate(cols) if cols else np.zeros(0, dtype=np.int64)
    vals = np.concatenate(vals) if vals else np.zeros(0)
    T = sp.csr_matrix((vals, (rows, cols)), shape=(dim * n, n_t))
    T.sort_indices()
    return T.indptr.astype(np.int64), T.indices.astype(np.int64), T.data.astype(np.float64), n_t


class PMultigrid:
    """Two-level V-cycle: Chebyshev(Jacobi) pre/post smoothing on the P2 operator, Galerkin coarse
    operator on the P1 vertex dofs.

    The coarse problem is solved with an explicit inverse of the Galerkin operator (Cholesky in
    FP64, inverse from it, stored in the level precision), so a coarse solve is one dense
    matrix-vector product: bandwidth-bound, instead of two latency-bound triangular solves.  The
    coarse operator is rebuilt on the first setup and then every `coarse_refresh` setups (the fine
    level always uses the current operator; FGMRES is flexible, so a lagged coarse correction
    only costs iterations, never correctness).

    When the free P1 dofs exceed `coarse_dense_max` (coarse="auto"; the dense matrix grows as its
    square) or with coarse="aggregate", the P1 level is kept sparse instead: the Galerkin product on
    the P1 pattern (`ssx_galerkin_csr`, formed at every setup), Chebyshev smoothing of degree
    `degree1`, and below it the dense level of rigid-body modes on node aggregates
    (`rbm_aggregates`), inverted and refreshed as above."""

    def __init__(self, elem, q_mask, *, use32: bool = True, degree: int = 3, lo_frac: float = 0.1,
                 coarse_refresh: int = 0, coarse_precision: str | None = None, refresh_iterations: int = 20,
                 refresh_ratio: float = 1.5, coarse: str = "auto", coord=None, coarse_dense_max: int = 50000,
                 degree1: int = 12):
        self.use32 = use32
        self.degree = degree
        self.lo_frac = lo_frac
        # coarse_refresh N >= 1: rebuild the coarse operator every N setups; 0: rebuild only when a
        # solve needed more than `refresh_iterations` FGMRES iterations (the lagged coarse
        # correction has become too weak)
        self.coarse_refresh = max(0, int(coarse_refresh))
        self.refresh_iterations = int(refresh_iterations)
        self.refresh_ratio = float(refresh_ratio)
        self._stale = False
        self._fresh_its = None    # iterations of the first solve after the latest coarse rebuild
        # precision of the coarse factorisation: FP64, or FP32 (falls back to FP64 when the FP32
        # Cholesky meets a non-positive pivot); default follows the level precision
        self.coarse_precision = coarse_precision or ("fp32" if use32 else "fp64")
        self.timings = {"level_refresh": 0.0, "galerkin": 0.0, "factor": 0.0, "coarse_builds": 0,
                        "coarse_fp64_fallbacks": 0}
        self._setups = 0
        p_ptr, p_idx, p_w, n_c, c_free = p2_to_p1_prolongation(elem, q_mask)
        n_free_c = int(c_free.sum())
        if coarse not in ("auto", "dense", "aggregate"):
            raise ValueError(f"coarse must be 'auto', 'dense' or 'aggregate', not {coarse!r}")
        if coarse == "auto":
            coarse = "aggregate" if n_free_c > coarse_dense_max else "dense"
        if coarse == "aggregate" and coord is None:
            raise ValueError("the aggregation level needs the node coordinates")
        self.coarse = coarse
        self.degree1 = int(degree1)
        if coarse == "dense":
            # the coarse space is numbered over the free coarse dofs only: no column of P touches a
            # constrained coarse dof, so the dense coarse system (factorised and inverted at each
            # rebuild, applied every Krylov iteration) carries no identity block
            keep = np.flatnonzero(c_free)
            remap = np.full(n_c, -1, dtype=np.int64)
            remap[keep] = np.arange(keep.size)
            p_idx = remap[p_idx]
            if p_idx.size and p_idx.min() < 0:
                raise AssertionError("P2 -> P1 interpolation references a constrained coarse dof")
            self.n_c = int(keep.size)
            self.c_free = np.ones(self.n_c, dtype=bool)
            self.n_d = self.n_c
        else:
            # sparse P1 level (the Galerkin product on the P1 pattern, unit rows on constrained
            # dofs, Chebyshev smoothing) above a dense level of rigid-body modes on aggregates
            from .pattern import build_owned_pattern
            elem = np.asarray(elem, dtype=np.int64)
            dim = q_mask.shape[0]
            verts = np.unique(elem[: dim + 1].reshape(-1))
            cnum = np.full(q_mask.shape[1], -1, dtype=np.int64)
            cnum[verts] = np.arange(verts.size)
            elem_c = cnum[elem[: dim + 1]]
            self.c_pat = build_owned_pattern(elem_c, q_mask[:, verts])
            self.n_c = int(n_c)
            self.c_free = c_free
            self._t_host = rbm_aggregates(np.asarray(coord, dtype=np.float64)[:, verts], elem_c, c_free)
            self.n_d = int(self._t_host[3])
        self._p_host = (p_ptr, p_idx, p_w)
        self.level = None
        self.level1 = None
        self.Ainv = None          # packed symmetric tiles of the coarse inverse
        self._tile_part = None
        self._spd_work = None
        self._spd_info = None
        self._dev_ready = False
        self._wb = {}

    def _to_device(self, dev, n_f):
        p_ptr, p_idx, p_w = self._p_host
        import scipy.sparse as sp
        P = sp.csr_matrix((p_w, p_idx, p_ptr), shape=(n_f, self.n_c))
        PT = P.T.tocsr()
        mk = lambda M: (torch.from_numpy(M.indptr.astype(np.int32)).to(dev),  # noqa: E731
                        torch.from_numpy(M.indices.astype(np.int32)).to(dev),
                        torch.from_numpy(M.data.astype(np.float64)).to(dev))
        self.P = DeviceCSR(n_f, *mk(P), None, spmv_lanes(P.indptr))
        self.P.values32 = self.P.values.to(torch.float32)
        self.PT = DeviceCSR(self.n_c, *mk(PT), None, spmv_lanes(PT.indptr))
        self.PT.values32 = self.PT.values.to(torch.float32)
        self.p_ptr = torch.from_numpy(p_ptr.astype(np.int32)).to(dev)
        self.p_idx = torch.from_numpy(p_idx.astype(np.int32)).to(dev)
        self.p_w = torch.from_numpy(p_w).to(dev)
        dt = torch.float32 if self.use32 else torch.float64
        self._gwork = torch.zeros(1, dtype=torch.int64, device=dev)   # scratch of the Galerkin kernels
        self.rc = torch.empty(self.n_c, dtype=dt, device=dev)
        self.xc = torch.empty(self.n_c, dtype=dt, device=dev)
        if self.coarse == "dense":
            # the dense level is the P1 level: Galerkin product of the fine operator through P
            self.d_ptr, self.d_idx, self.d_w = self.p_ptr, self.p_idx, self.p_w
            self.cfix = torch.from_numpy(np.flatnonzero(~self.c_free)).to(dev)
        else:
            cp = self.c_pat
            self.Kc = DeviceCSR(self.n_c, torch.from_numpy(cp.indptr).to(dev), torch.from_numpy(cp.indices).to(dev),
                                torch.zeros(cp.nnz, dtype=torch.float64, device=dev),
                                torch.zeros(cp.nnz, dtype=torch.float32, device=dev), spmv_lanes(cp.indptr))
            self.kc_cdiag = torch.from_numpy(cp.constrained_diag_positions.astype(np.int64)).to(dev)
            self.kc_max_row = int(np.diff(cp.indptr).max()) if cp.indptr.size > 1 else 0
            t_ptr, t_idx, t_w, n_t = self._t_host
            T = sp.csr_matrix((t_w, t_idx, t_ptr), shape=(self.n_c, n_t))
            self.T = DeviceCSR(self.n_c, *mk(T), None, spmv_lanes(T.indptr))
            self.T.values32 = self.T.values.to(torch.float32)
            TT = T.T.tocsr()
            self.TT = DeviceCSR(n_t, *mk(TT), None, spmv_lanes(TT.indptr))
            self.TT.values32 = self.TT.values.to(torch.float32)
            # the dense level is the aggregate level: Galerkin product of the P1 level through T
            self.d_ptr = torch.from_numpy(t_ptr.astype(np.int32)).to(dev)
            self.d_idx = torch.from_numpy(t_idx.astype(np.int32)).to(dev)
            self.d_w = torch.from_numpy(t_w).to(dev)
            self.cfix = torch.zeros(0, dtype=torch.int64, device=dev)
            self.res1 = torch.empty(self.n_c, dtype=dt, device=dev)
            self.r2 = torch.empty(n_t, dtype=dt, device=dev)
            self.x2 = torch.empty(n_t, dtype=dt, device=dev)
        self._dev_ready = True

    def _w_bound(self, A: DeviceCSR, which: str) -> float:
        """Weight bound of the fixed-point Galerkin sums for operator A through P ("p") or through
        the dense level's map ("d"); cached per pattern."""
        key = (which, int(A.indptr.data_ptr()), A.n_rows)
        if key not in self._wb:
            max_len = int((A.indptr[1:] - A.indptr[:-1]).max().item()) if A.n_rows else 0
            if which == "p":
                p_ptr, p_idx, p_w = self._p_host
                n_c = self.n_c
            elif self.coarse == "dense":
                p_ptr, p_idx, p_w = self._p_host
                n_c = self.n_d
            else:
                p_ptr, p_idx, p_w, n_c = self._t_host
            self._wb[key] = galerkin_weight_bound(p_ptr, p_idx, p_w, n_c, max_len)
        return self._wb[key]

    def setup(self, A: DeviceCSR, mesh_info=None):
        dev = A.indptr.device
        if not self._dev_ready:
            self._to_device(dev, A.n_rows)
        if self.level is None or self.level.A is not A:
            self.level = _Level(A, self.use32)
        t0 = perf_counter()
        self.level.refresh()
        self.timings["level_refresh"] += perf_counter() - t0
        dense_src = A
        due = self.coarse_refresh > 0 and self._setups % self.coarse_refresh == 0
        rebuild = self.Ainv is None or due or self._stale
        if self.coarse == "aggregate":
            # the sparse P1 level follows the current operator at every setup
            t0 = perf_counter()
            Kc = self.Kc
            _lib.call("ssx_galerkin_csr", A.n_rows, ptr(A.indptr), ptr(A.indices), ptr(A.values), A.nnz,
                      ptr(self.p_ptr), ptr(self.p_idx), ptr(self.p_w), ptr(self.PT.indptr), ptr(self.PT.indices),
                      ptr(self.PT.values), self.n_c, ptr(Kc.indptr), ptr(Kc.indices), Kc.nnz, self.kc_max_row,
                      self._w_bound(A, "p"), ptr(Kc.values), ptr(self._gwork), _st())
            Kc.values[self.kc_cdiag] = 1.0
            _lib.call("ssx_cast_f64_f32", Kc.nnz, ptr(Kc.values), ptr(Kc.values32), _st())
            if self.level1 is None:
                self.level1 = _Level(Kc, self.use32)
            self.level1.refresh()
            torch.cuda.synchronize()
            self.timings["p1_level"] = self.timings.get("p1_level", 0.0) + perf_counter() - t0
        if self.coarse == "aggregate":
            dense_src = self.Kc
        if rebuild:
            self._fresh_its = None
            self._coarse(dense_src, dev)
            self._stale = False
        self._setups += 1

    def _spd_inverse(self, Ac: torch.Tensor, f64: bool, dt) -> bool:
        """Ainv tiles <- Ac^-1 (`ssx_spd_inverse_tiles`, computed in FP64 or FP32); False when the
        factorisation met a non-positive pivot."""
        n = self.n_d
        need = spd_inverse_work(n) * (8 if f64 else 4)
        if self._spd_work is None or self._spd_work.numel() < need:
            self._spd_work = None
            self._spd_work = torch.empty(need, dtype=torch.uint8, device=Ac.device)
        if self._spd_info is None:
            self._spd_info = torch.zeros(1, dtype=torch.int32, device=Ac.device)
        _lib.call("ssx_spd_inverse_tiles", n, ptr(Ac), int(f64), ptr(self._spd_work), ptr(self.Ainv),
                  int(dt == torch.float64), ptr(self._spd_info), _st())
        return int(self._spd_info.item()) == 0

    def note_iterations(self, its: int) -> None:
        """Called by the Krylov solver after each solve with its iteration count.  The coarse
        operator is marked stale when a solve needs more than `refresh_iterations` iterations and
        more than `refresh_ratio` times what the first solve after the last rebuild needed (when
        even a fresh coarse operator leaves many iterations, the fine level is what limits them,
        and a rebuild would not help)."""
        if self._fresh_its is None:
            self._fresh_its = its
        if self.coarse_refresh == 0 and its > max(self.refresh_iterations, self.refresh_ratio * self._fresh_its):
            self._stale = True

    def _coarse(self, A: DeviceCSR, dev) -> None:
        # coarse operator (FP64 accumulate), identity on constrained coarse dofs; the Cholesky
        # factorisation reads the lower triangle only
        tm = self.timings
        t0 = perf_counter()
        self.Ainv = None
        Ac = torch.zeros((self.n_d, self.n_d), dtype=torch.float64, device=dev)
        _lib.call("ssx_galerkin_dense", A.n_rows, ptr(A.indptr), ptr(A.indices), ptr(A.values), A.nnz,
                  ptr(self.d_ptr), ptr(self.d_idx), ptr(self.d_w), self.n_d, self._w_bound(A, "d"), ptr(Ac),
                  ptr(self._gwork), _st())
        Ac[self.cfix, self.cfix] = 1.0
        torch.cuda.synchronize()
        t1 = perf_counter()
        tm["galerkin"] += t1 - t0
        # explicit inverse in the symmetric tile store (the coarse solve is then one tile GEMV):
        # blocked Cholesky + triangular inverse + W^T W on the device, in the coarse precision; an
        # FP32 factorisation that meets a non-positive pivot is redone in FP64, and an indefinite
        # operator (strong softening) gets a diagonal shift raised by 100x until it factorises (the
        # coarse level only has to be a preconditioner for the outer FGMRES)
        dt = torch.float32 if self.use32 else torch.float64
        te, pe = sym_tile_sizes(self.n_d)
        if self._tile_part is None or self._tile_part.dtype != dt:
            self._tile_part = torch.empty(pe, dtype=dt, device=dev)
        self.Ainv = torch.empty(te, dtype=dt, device=dev)
        ok = False
        if self.coarse_precision == "fp32":
            ok = self._spd_inverse(Ac, False, dt)
            if not ok:
                tm["coarse_fp64_fallbacks"] += 1
        if not ok:
            ok = self._spd_inverse(Ac, True, dt)
        if not ok:
            dmax = float(Ac.diagonal().abs().max().item()) or 1.0
            for rel_shift in (1e-8, 1e-6, 1e-4, 1e-2, 1.0):
                Ac.diagonal().add_(rel_shift * dmax)
                tm["coarse_shifts"] = tm.get("coarse_shifts", 0) + 1
                if self._spd_inverse(Ac, True, dt):
                    break
            else:
                raise RuntimeError("p-multigrid: coarse operator not positive definite after shifting")
        del Ac
        torch.cuda.synchronize()
        tm["factor"] += perf_counter() - t1
        tm["coarse_builds"] += 1

    def cycle(self, b: torch.Tensor, x: torch.Tensor) -> None:
        """One V-cycle on A x = b in the level precision (x overwritten)."""
        Lv = self.level
        Lv.cheb(b, x, self.degree, self.lo_frac, zero_guess=True)          # pre-smooth
        res = torch.empty_like(x)
        axpby(1.0, b, 0.0, res)
        spmv(Lv.A, x, res, alpha=-1.0, beta=1.0, use32=self.use32)          # res = b - A x
        spmv(self.PT, res, self.rc, use32=self.use32)                          # restrict
        if self.coarse == "dense":
            self._dense_solve(self.rc, self.xc)                                 # coarse solve
        else:
            L1 = self.level1                                                    # P1 level V-cycle
            L1.cheb(self.rc, self.xc, self.degree1, self.lo_frac, zero_guess=True)
            axpby(1.0, self.rc, 0.0, self.res1)
            spmv(L1.A, self.xc, self.res1, alpha=-1.0, beta=1.0, use32=self.use32)
            spmv(self.TT, self.res1, self.r2, use32=self.use32)
            self._dense_solve(self.r2, self.x2)
            spmv(self.T, self.x2, self.xc, alpha=1.0, beta=1.0, use32=self.use32)
            L1.cheb(self.rc, self.xc, self.degree1, self.lo_frac, zero_guess=False)
        spmv(self.P, self.xc, x, alpha=1.0, beta=1.0, use32=self.use32)       # prolong + correct
        Lv.cheb(b, x, self.degree, self.lo_frac, zero_guess=False)          # post-smooth

    def _dense_solve(self, b: torch.Tensor, x: torch.Tensor) -> None:
        _lib.call("ssx_sym_tiles_gemv_f32" if self.use32 else "ssx_sym_tiles_gemv_f64", self.n_d,
                  ptr(self.Ainv), ptr(b), ptr(self._tile_part), ptr(x), _st())

    def apply(self, r: torch.Tensor, z: torch.Tensor) -> None:
        Lv = self.level
        n = Lv.n
        if self.use32:
            b = torch.empty(n, dtype=torch.float32, device=r.device)
            _lib.call("ssx_cast_f64_f32", n, ptr(r), ptr(b), _st())
        else:
            b = r
        x = Lv.x
        self.cycle(b, x)
        if self.use32:
            _lib.call("ssx_cast_f32_f64", n, ptr(x), ptr(z), _st())
        else:
            z.copy_(x)


# ---------------------------------------------------------------------------------------------
# Deflated FGMRES with the reference's solver duck API
# ---------------------------------------------------------------------------------------------

class DeflatedFGMRES:
    def __init__(self, preconditioner, *, tolerance: float = 1e-1, max_iterations: int = 100,
                 tolerance_deflation_basis: float = 1e-3, max_deflation_basis_vectors: int = 48,
                 reorth_passes: int = 1, orthogonalization: str = "cgs2"):
        """``orthogonalization``: "cgs2" (classical Gram-Schmidt plus ``reorth_passes`` passes, two
        batched launches per pass) or "mgs" (modified Gram-Schmidt in the reference's order,
        linear/deflated_fgmres.py:259-263: one dot and one update per basis vector, no
        re-orthogonalisation)."""
        if orthogonalization not in ("cgs2", "mgs"):
            raise ValueError(f"unknown orthogonalization {orthogonalization!r}")
        self.orthogonalization = orthogonalization
        self.pc = preconditioner
        self.tolerance = float(tolerance)
        self.max_iterations = int(max_iterations)
        self.tolerance_deflation_basis = float(tolerance_deflation_basis)
        self.max_basis = int(max_deflation_basis_vectors)
        self.reorth_passes = int(reorth_passes)
        self.iteration_collector = IterationCollector()
        self.deflation_basis: torch.Tensor | None = None   # (k, n) raw vectors, newest last
        self._W = None      # A-orthonormal basis (k, n)
        self._AW = None
        self._A = None
        self._scr = None
        self.last_residual_history = None

    # duck API --------------------------------------------------------------------------------
    def prefers_full_system_operator(self):
        return True

    def supports_a_orthogonalization(self):
        return True

    def supports_dynamic_deflation_basis(self):
        return True

    def preconditioner_requires_explicit_matrix(self):
        return False

    def needs_preconditioning_matrix_refresh(self):
        return False

    def release_iteration_resources(self):
        pass

    def notify_continuation_attempt(self, success: bool = True):
        pass

    def copy(self):
        c = DeflatedFGMRES(self.pc, tolerance=self.tolerance, max_iterations=self.max_iterations,
                           tolerance_deflation_basis=self.tolerance_deflation_basis,
                           max_deflation_basis_vectors=self.max_basis, reorth_passes=self.reorth_passes,
                           orthogonalization=self.orthogonalization)
        c.deflation_basis = self.deflation_basis
        c.iteration_collector = self.iteration_collector
        return c

    def get_deflation_basis_snapshot(self):
        return None if self.deflation_basis is None else self.deflation_basis.clone()

    def restore_deflation_basis(self, snap):
        self.deflation_basis = None if snap is None else snap.clone()
        self._W = None

    def expand_deflation_basis(self, v: torch.Tensor):
        v = v.reshape(1, -1).to(torch.float64)
        if self.deflation_basis is None:
            self.deflation_basis = v.clone()
        else:
            self.deflation_basis = torch.cat([self.deflation_basis, v], dim=0)
        if self.deflation_basis.shape[0] > self.max_basis:
            self.deflation_basis = self.deflation_basis[-self.max_basis:].contiguous()
        self._W = None

    def setup_preconditioner(self, A: DeviceCSR, full_matrix=None, free_indices=None, preconditioning_matrix=None):
        t0 = perf_counter()
        self._A = A
        if self._scr is None:
            self._scr = _Scratch(A.indptr.device)
        self.pc.setup(A)
        if A.values.is_cuda:
            torch.cuda.synchronize()
        self._pc_setup_time = perf_counter() - t0
        self.iteration_collector.setup_time.append(self._pc_setup_time)

    def A_orthogonalize(self, A: DeviceCSR):
        """A-orthonormalise the basis newest-first, dropping columns with |v^T A v| <= eps
        (linear/orthogonalize.py:49-170).

        One C-ABI call (`ssx_a_orthogonalize`): all raw products A b come from one multi-vector
        SpMV, the coefficients, the A-norm test and the scaling stay on the device, a dropped
        column becomes a zero row with sign 0 (so it takes no part in later projections), and A W
        is carried by the same recurrence; the kept rows are compacted after one host read."""
        self._A = A
        B = self.deflation_basis
        if B is None or B.shape[0] == 0:
            self._W = None
            self._AW = None
            return
        dev = B.device
        B = B.contiguous()
        kmax, n = B.shape
        if n != A.n_rows:
            raise ValueError(f"A_orthogonalize: basis length {n} != operator rows {A.n_rows}")
        Wbuf = torch.empty((kmax, n), dtype=torch.float64, device=dev)
        AWbuf = torch.empty((kmax, n), dtype=torch.float64, device=dev)
        sgn = torch.empty(kmax, dtype=torch.float64, device=dev)
        work = torch.empty(3 * kmax + max(64 * kmax, 1024), dtype=torch.float64, device=dev)
        _lib.call("ssx_a_orthogonalize", n, kmax, ptr(B), n, ptr(A.indptr), ptr(A.indices), ptr(A.values), A.lanes,
                  float(self.tolerance_deflation_basis), ptr(Wbuf), ptr(AWbuf), ptr(sgn), ptr(work), _st())
        kept = torch.nonzero(sgn != 0).flatten().cpu().numpy()   # the one host read
        if kept.size == 0:
            self._W = None
            self._AW = None
            return
        order = torch.from_numpy(kept[::-1].copy()).to(dev)     # oldest first
        self._W = Wbuf.index_select(0, order).contiguous()
        self._AW = AWbuf.index_select(0, order).contiguous()

    def solve(self, A: DeviceCSR, b: torch.Tensor, full_rhs=None, local_rhs=None, free_indices=None):
        t0 = perf_counter()
        if self._scr is None:
            self._scr = _Scratch(A.indptr.device)
        scr = self._scr
        dev = b.device
        n = b.numel()
        m = self.max_iterations
        tol = self.tolerance
        t_pc = 0.0
        t_orth = 0.0
        W, AW = self._W, self._AW
        k = 0 if W is None else W.shape[0]
        x = torch.zeros(n, dtype=torch.float64, device=dev)
        if k:
            h = torch.empty(k, dtype=torch.float64, device=dev)
            _lib.call("ssx_dot_cols", n, k, ptr(W), n, ptr(b), ptr(scr.cols_for(k)), ptr(h), _st())
            _lib.call("ssx_gemv_cols", n, k, ptr(W), n, ptr(h), 1.0, ptr(x), _st())
        r = b.clone()
        spmv(A, x, r, alpha=-1.0, beta=1.0)
        bnorm = np.sqrt(max(dot(b, b, scr), 0.0)) or 1.0
        rnorm = np.sqrt(max(dot(r, r, scr), 0.0))
        hist = [rnorm / bnorm]
        if hist[0] < tol or rnorm == 0.0:
            self.iteration_collector.store(0, perf_counter() - t0, 0.0, 0.0)
            self.last_residual_history = np.array(hist)
            return x
        scr.cols_for(max(m + 1, k))
        V = torch.empty((m + 1, n), dtype=torch.float64, device=dev)
        Z = torch.empty((m, n), dtype=torch.float64, device=dev)
        H = np.zeros((m + 1, m))
        g = np.zeros(m + 1)
        g[0] = rnorm
        axpby(1.0 / rnorm, r, 0.0, V[0])
        hbuf = torch.empty(m + 1, dtype=torch.float64, device=dev)
        hbuf2 = torch.empty(m + 1, dtype=torch.float64, device=dev)
        cbuf = torch.empty(max(k, 1), dtype=torch.float64, device=dev)
        iters = 0
        y = np.zeros(0)
        for j in range(m):
            tp = perf_counter()
            self.pc.apply(V[j], Z[j])
            t_pc += perf_counter() - tp
            to = perf_counter()
            if k:
                _lib.call("ssx_dot_cols", n, k, ptr(AW), n, ptr(Z[j]), ptr(scr.cols), ptr(cbuf), _st())
                _lib.call("ssx_gemv_cols", n, k, ptr(W), n, ptr(cbuf), -1.0, ptr(Z[j]), _st())
            u = V[j + 1]
            spmv(A, Z[j], u)
            if self.orthogonalization == "mgs":
                for i in range(j + 1):
                    _lib.call("ssx_dot_cols", n, 1, ptr(V[i]), n, ptr(u), ptr(scr.cols), ptr(hbuf[i:]), _st())
                    _lib.call("ssx_gemv_cols", n, 1, ptr(V[i]), n, ptr(hbuf[i:]), -1.0, ptr(u), _st())
            else:
                _lib.call("ssx_dot_cols", n, j + 1, ptr(V), n, ptr(u), ptr(scr.cols), ptr(hbuf), _st())
                _lib.call("ssx_gemv_cols", n, j + 1, ptr(V), n, ptr(hbuf), -1.0, ptr(u), _st())
                for _ in range(self.reorth_passes):
                    _lib.call("ssx_dot_cols", n, j + 1, ptr(V), n, ptr(u), ptr(scr.cols), ptr(hbuf2), _st())
                    _lib.call("ssx_gemv_cols", n, j + 1, ptr(V), n, ptr(hbuf2), -1.0, ptr(u), _st())
                    hbuf[: j + 1] += hbuf2[: j + 1]
            # |u|^2 lands in hbuf[j+1]: the Hessenberg column and the norm come back in one read
            _lib.call("ssx_dot", n, ptr(u), ptr(u), ptr(scr.part), ptr(hbuf[j + 1:]), _st())
            col = hbuf[: j + 2].cpu().numpy()
            hn = float(np.sqrt(max(col[j + 1], 0.0)))
            H[: j + 1, j] = col[: j + 1]
            H[j + 1, j] = hn
            t_orth += perf_counter() - to
            iters = j + 1
            if hn < 1e-14:
                y, *_ = np.linalg.lstsq(H[: j + 1, : j + 1], g[: j + 1], rcond=None)
                break
            axpby(1.0 / hn, u, 0.0, u)
            y, *_ = np.linalg.lstsq(H[: j + 2, : j + 1], g[: j + 2], rcond=None)
            res = float(np.linalg.norm(g[: j + 2] - H[: j + 2, : j + 1] @ y)) / bnorm
            hist.append(res)
            if not np.isfinite(res):
                break
            if res < tol:
                break
        if iters:
            yt = torch.from_numpy(np.ascontiguousarray(y[:iters])).to(dev)
            _lib.call("ssx_gemv_cols", n, iters, ptr(Z), n, ptr(yt), 1.0, ptr(x), _st())
        self.last_residual_history = np.array(hist)
        self.iteration_collector.store(iters, perf_counter() - t0, t_pc, t_orth)
        note = getattr(self.pc, "note_iterations", None)
        if note is not None:
            note(iters)
        return x
