"""The fixed owned-row CSR pattern and the row-slot metadata of the assembly kernels.

Setup-time host code (vectorised NumPy).  The integer structure is the reference's, entry for
entry, so the device values land at the same CSR positions PETSc would hold:

* structural rule (fem/distributed_tangent.py:440-478): each free row holds the sorted unique
  free columns of every element touching it; a constrained row holds only its diagonal;
* row-slot lists (fem/distributed_tangent.py:369-437): for each owned free row, one slot per
  element touching it, in element order; ``slot_lrow`` is the local dof (uint8), ``slot_pos``
  the CSR position of every local column (-1 for constrained columns);
* constrained diagonal positions (fem/distributed_tangent.py:297-325), int32 overflow guards
  (:122-124, :383-384).

On one device the owned range is the whole mesh, so the overlap elements are all elements in
their original order and overlap-local dofs are global dofs.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

INT32_MAX = int(np.iinfo(np.int32).max)


def _check_int32(name: str, value: int) -> None:
    if int(value) > INT32_MAX:
        raise OverflowError(f"{name}={int(value)} exceeds int32 capacity")


@dataclass
class OwnedPattern:
    dim: int
    n_p: int
    n_nodes: int
    indptr: np.ndarray             # (n_rows+1,) int32
    indices: np.ndarray            # (nnz,) int32
    row_free: np.ndarray           # (n_rows,) uint8
    constrained_diag_positions: np.ndarray  # int32
    row_slot_ptr: np.ndarray       # (n_rows+1,) int32
    slot_elem: np.ndarray          # (n_slots,) int32
    slot_lrow: np.ndarray          # (n_slots,) uint8
    elem_dofs: np.ndarray          # (n_elem, dim*n_p) int32
    free_idx: np.ndarray           # (n_free,) int64 free dofs (Fortran order of (dim, n_nodes))
    stats: dict = field(default_factory=dict)
    # slot_pos (n_slots, dim*n_p) int32, the CSR position of every local column of every slot, is
    # formed on first use (the node-gather path needs it for one row per node only)
    _slot_positions: object = None
    _slot_pos: np.ndarray | None = None

    @property
    def slot_pos(self) -> np.ndarray:
        if self._slot_pos is None:
            n = int(self.slot_elem.size)
            self._slot_pos = np.ascontiguousarray(self.slot_positions(np.arange(n)).astype(np.int32))
        return self._slot_pos

    def slot_positions(self, sel: np.ndarray) -> np.ndarray:
        """slot_pos rows of the slots `sel` (int64, -1 for constrained columns)."""
        return self._slot_positions(np.asarray(sel, dtype=np.int64))

    @property
    def n_rows(self) -> int:
        return int(self.indptr.size - 1)

    @property
    def nnz(self) -> int:
        return int(self.indices.size)

    def max_block_nnz(self, rows_per_block: int = 64) -> int:
        ip = self.indptr.astype(np.int64)
        starts = ip[0:-1:rows_per_block]
        ends = ip[np.minimum(np.arange(rows_per_block, self.n_rows + rows_per_block, rows_per_block), self.n_rows)]
        return int((ends - starts).max()) if starts.size else 0


def element_dofs(elem: np.ndarray, dim: int) -> np.ndarray:
    """(n_elem, dim*n_p) global dofs, dof = dim*node + comp (utils.py:28-36 ordering)."""
    elem = np.asarray(elem, dtype=np.int64)
    n_p, n_e = elem.shape
    return (dim * elem.T[:, :, None] + np.arange(dim)[None, None, :]).reshape(n_e, dim * n_p)


def build_owned_pattern(elem: np.ndarray, q_mask: np.ndarray) -> OwnedPattern:
    elem = np.asarray(elem, dtype=np.int64)
    q_mask = np.asarray(q_mask, dtype=bool)
    dim, n_nodes = q_mask.shape
    n_p, n_e = elem.shape
    nld = dim * n_p
    if nld > 255:
        raise OverflowError(f"n_local_dof={nld} exceeds uint8 slot_lrow capacity")
    n_rows = dim * n_nodes
    free_dof = q_mask.T.reshape(-1)          # free_dof[dim*node + comp]
    free_cnt = q_mask.sum(axis=0).astype(np.int64)  # free comps per node

    # --- node adjacency (sorted unique neighbour nodes incl. self) ---
    ei = np.repeat(elem.T, n_p, axis=1).reshape(-1)   # (n_e*n_p*n_p) row node
    ej = np.tile(elem.T, (1, n_p)).reshape(-1)         # col node
    key = np.unique(ei * n_nodes + ej)                 # sorted by (row node, col node)
    nb_row = key // n_nodes
    nb_col = key - nb_row * n_nodes
    nb_ptr = np.zeros(n_nodes + 1, dtype=np.int64)
    np.add.at(nb_ptr, nb_row + 1, 1)
    nb_ptr = np.cumsum(nb_ptr)
    # prefix of free-column counts within each node's neighbour list
    cnt = free_cnt[nb_col]
    csum = np.cumsum(cnt)
    colstart = csum - cnt - np.repeat(np.concatenate(([0], csum))[nb_ptr[:-1]], np.diff(nb_ptr))
    node_len = np.add.reduceat(cnt, nb_ptr[:-1]) if key.size else np.zeros(n_nodes, dtype=np.int64)
    node_len = np.where(np.diff(nb_ptr) > 0, node_len, 0)

    # --- CSR over dofs ---
    row_len = np.where(free_dof, np.repeat(node_len, dim), 1).astype(np.int64)
    indptr = np.zeros(n_rows + 1, dtype=np.int64)
    indptr[1:] = np.cumsum(row_len)
    nnz = int(indptr[-1])
    _check_int32("local_matrix.nnz", nnz)
    # free columns of every node's neighbour list, in order (node-level template)
    col_node = np.repeat(nb_col, cnt)                          # neighbour node per template entry
    comp_rank = np.arange(col_node.size) - np.repeat(csum - cnt, cnt)  # k-th free comp of that node
    # comp of the k-th free component of node m
    rank = np.cumsum(q_mask, axis=0) - 1                    # (dim, n_nodes): rank of comp d among free comps
    comp_table = np.full((n_nodes, dim), -1, dtype=np.int64)
    for d in range(dim):
        sel = np.flatnonzero(q_mask[d])
        comp_table[sel, rank[d, sel]] = d
    tmpl_cols = dim * col_node + comp_table[col_node, comp_rank]
    tmpl_ptr = np.concatenate(([0], np.cumsum(node_len)))
    indices = np.empty(nnz, dtype=np.int64)
    node_of_row = np.arange(n_rows) // dim
    # free rows copy their node's template; constrained rows get their diagonal
    fr = np.flatnonzero(free_dof)
    seg_len = row_len[fr]
    offset = np.arange(int(seg_len.sum())) - np.repeat(np.cumsum(seg_len) - seg_len, seg_len)
    indices[np.repeat(indptr[fr], seg_len) + offset] = tmpl_cols[np.repeat(tmpl_ptr[node_of_row[fr]], seg_len) + offset]
    cr = np.flatnonzero(~free_dof)
    indices[indptr[cr]] = cr
    constrained_diag = indptr[cr]

    # --- row slots in element order ---
    edofs = element_dofs(elem, dim)                          # (n_e, nld)
    flat_rows = edofs.reshape(-1)
    flat_elem = np.repeat(np.arange(n_e), nld)
    flat_a = np.tile(np.arange(nld), n_e)
    keep = free_dof[flat_rows]
    flat_rows, flat_elem, flat_a = flat_rows[keep], flat_elem[keep], flat_a[keep]
    order = np.argsort(flat_rows, kind="stable")            # stable: element order within a row
    slot_row = flat_rows[order]
    slot_elem = flat_elem[order]
    slot_lrow = flat_a[order]
    n_slots = slot_row.size
    _check_int32("row_slot_count", n_slots)
    row_slot_ptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.add.at(row_slot_ptr, slot_row + 1, 1)
    row_slot_ptr = np.cumsum(row_slot_ptr)

    # --- slot_pos: CSR position of every local column of the slot's element (on demand) ---
    def slot_positions(sel: np.ndarray) -> np.ndarray:
        s_row, s_elem, s_lrow = slot_row[sel], slot_elem[sel], slot_lrow[sel].astype(np.int64)
        node_a = elem[s_lrow // dim, s_elem]                  # row node per slot
        col_nodes = elem[:, s_elem].T                           # (n_sel, n_p)
        pair_idx = np.searchsorted(key, node_a[:, None] * n_nodes + col_nodes)
        start = colstart[pair_idx]                              # offset of node j's first free col in row
        pos = np.full((sel.size, nld), -1, dtype=np.int64)
        base = indptr[s_row][:, None]
        for d in range(dim):
            fr_d = q_mask[d][col_nodes]
            pos[:, d::dim] = np.where(fr_d, base + start + rank[d][col_nodes], -1)
        return pos

    return OwnedPattern(
        dim=dim, n_p=n_p, n_nodes=n_nodes,
        indptr=indptr.astype(np.int32), indices=indices.astype(np.int32),
        row_free=free_dof.astype(np.uint8),
        constrained_diag_positions=constrained_diag.astype(np.int32),
        row_slot_ptr=row_slot_ptr.astype(np.int32), slot_elem=slot_elem.astype(np.int32),
        slot_lrow=slot_lrow.astype(np.uint8),
        elem_dofs=np.ascontiguousarray(edofs.astype(np.int32)),
        free_idx=np.flatnonzero(q_mask.reshape(-1, order="F")).astype(np.int64),
        stats={"n_slots": float(n_slots), "nnz": float(nnz)},
        _slot_positions=slot_positions,
    )


@dataclass
class NodeSlots:
    """Per-node element lists of the node-gather assembly.

    For every node with at least one free component: the elements touching it in element order
    (the row-slot order of its free rows), the node's local index in each, and the CSR position of
    every local column relative to the start of the node's rows (int16; a constrained column holds
    the row length L, which is the spare word after each row in the gathers' LDS copy of the rows,
    so the gathers index it without a test).  All free rows of a node hold the same column list, so
    one offset table serves them all.  Nodes without free components get no slots (their rows are
    unit rows)."""

    ptr: np.ndarray        # (n_nodes+1,) int32
    elem: np.ndarray       # (n_slots,) int32
    lnode: np.ndarray      # (n_slots,) uint8
    off: np.ndarray        # (n_slots, dim*n_p) int16
    max_len: int           # longest free row
    node_len: np.ndarray   # (n_nodes,) int32: length of the node's free rows (0: no free row)

    def buckets(self, base: int = 96) -> list[tuple[np.ndarray, int]]:
        """Nodes grouped by row length, (node ids, longest row of the group): lengths up to `base`,
        then doubling classes.  The gather's LDS per workgroup is sized by the group's longest row,
        so short-row nodes (most nodes of a P4 mesh) are not held to the longest row's LDS."""
        out, lo, hi = [], -1, base
        while lo < self.max_len:
            ids = np.flatnonzero((self.node_len > lo) & (self.node_len <= hi)).astype(np.int32)
            if ids.size:
                out.append((ids, int(self.node_len[ids].max())))
            lo, hi = hi, 2 * hi
        return out


def gather_order(coord: np.ndarray, n_groups: int = 8) -> np.ndarray:
    """Launch order of a node gather (workgroup k gathers node order[k]).

    Nodes are sorted along a Morton (Z-order) curve of their coordinates, so nodes gathered close
    together share elements.  Workgroups are dealt to the GPU's n_groups L2 domains (XCDs) in
    turn (workgroup k to domain k mod n_groups), so the sorted list is cut into n_groups
    contiguous runs and workgroup k takes entry k // n_groups of run k mod n_groups: each domain
    walks one compact region, and its L2 serves the element data its nodes share (on the 117k-tet
    P2 slope the node gather's HBM reads fall from 793 to 440 MiB of FETCH_SIZE).  The gathered
    values do not depend on the order."""
    X = np.asarray(coord, dtype=np.float64)
    n = X.shape[1]
    lo = X.min(axis=1, keepdims=True)
    span = np.maximum(X.max(axis=1, keepdims=True) - lo, 1e-300)
    q = np.minimum(((X - lo) / span * 1024).astype(np.int64), 1023)
    code = np.zeros(n, dtype=np.int64)
    for bit in range(10):
        for d in range(X.shape[0]):
            code |= ((q[d] >> bit) & 1) << (X.shape[0] * bit + d)
    order = np.argsort(code, kind="stable")
    c, rem = divmod(n, n_groups)
    k = np.arange(n, dtype=np.int64)
    x, idx = k % n_groups, k // n_groups
    return order[x * c + np.minimum(x, rem) + idx].astype(np.int32)


def node_slots(pat: OwnedPattern) -> NodeSlots:
    dim, n_nodes = pat.dim, pat.n_nodes
    nld = dim * pat.n_p
    free = pat.row_free.reshape(n_nodes, dim).astype(bool)
    has_free = free.any(axis=1)
    r0 = dim * np.arange(n_nodes, dtype=np.int64) + np.argmax(free, axis=1)   # first free row of each node
    rsp = pat.row_slot_ptr.astype(np.int64)
    ip = pat.indptr.astype(np.int64)
    cnt = np.where(has_free, rsp[r0 + 1] - rsp[r0], 0)
    ptr = np.zeros(n_nodes + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(cnt)
    n = int(ptr[-1])
    _check_int32("node_slot_count", n)
    sel = np.repeat(rsp[r0] - ptr[:-1], cnt) + np.arange(n)                    # row-slot index per node slot
    base = np.repeat(ip[r0], cnt)
    pos = pat.slot_positions(sel)
    row_len = np.diff(ip)
    node_len = np.repeat(row_len[r0], cnt)
    off = np.where(pos >= 0, pos - base[:, None], node_len[:, None])
    max_len = int(row_len[pat.row_free.astype(bool)].max()) if pat.row_free.any() else 1
    if max_len >= 32768:
        raise OverflowError(f"row length {max_len} exceeds the int16 column offsets")
    if n and ((off.max(axis=1) > node_len).any() or (off.min(axis=1) < 0).any()):
        raise AssertionError("node-slot column offset outside its row")
    return NodeSlots(ptr=ptr.astype(np.int32), elem=pat.slot_elem[sel].astype(np.int32),
                     lnode=(pat.slot_lrow[sel] // dim).astype(np.uint8),
                     off=np.ascontiguousarray(off.astype(np.int16).reshape(n, nld)), max_len=max_len,
                     node_len=np.where(has_free, row_len[r0], 0).astype(np.int32))
