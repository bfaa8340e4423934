"""The allocation-free A-orthonormalisation (ssx_a_orthogonalize_noalloc, scratch in a static
device array) against the entry point with a caller workspace (ssx_a_orthogonalize): identical
W, A W and signs on two calls in a row, the second with a smaller operator and basis; kmax above
the static limit is rejected."""

import numpy as np
import pytest
import scipy.sparse as sps
import torch

from slope_stability_petsc4py_amd import _lib
from slope_stability_petsc4py_amd.device import ptr, spmv_lanes, stream_ptr

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    M = sps.random(n, n, density=8.0 / n, random_state=seed, format="csr")
    A = (M + M.T + sps.eye(n) * (2.0 + 16.0)).tocsr()
    A.sort_indices()
    return A, rng


def _run(name, A, B, dev, extra):
    n = A.shape[0]
    kmax = B.shape[0]
    ip = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    ix = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    vv = torch.from_numpy(A.data).to(dev)
    Bd = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    W = torch.full((kmax, n), np.nan, dtype=torch.float64, device=dev)
    AW = torch.full((kmax, n), np.nan, dtype=torch.float64, device=dev)
    sgn = torch.full((kmax,), np.nan, dtype=torch.float64, device=dev)
    _lib.call(name, n, kmax, ptr(Bd), n, ptr(ip), ptr(ix), ptr(vv), spmv_lanes(A.indptr), 1e-3, ptr(W), ptr(AW),
              ptr(sgn), *extra, stream_ptr())
    torch.cuda.synchronize()
    return W.cpu().numpy(), AW.cpu().numpy(), sgn.cpu().numpy()


@pytest.mark.parametrize("calls", [[(3000, 12), (1200, 5)]])
def test_noalloc_matches_workspace_entry_point(gpu, calls):
    print("case: a_orthogonalize_noalloc", flush=True)
    for n, kmax in calls:
        A, rng = _spd(n, seed=n)
        B = rng.normal(size=(kmax, n))
        B[2] = B[0] + 1e-9 * rng.normal(size=n)      # a nearly dependent column: dropped (sign 0)
        work = torch.empty(3 * kmax + max(64 * kmax, 1024), dtype=torch.float64, device=gpu)
        ref = _run("ssx_a_orthogonalize", A, B, gpu, [ptr(work)])
        got = _run("ssx_a_orthogonalize_noalloc", A, B, gpu, [])
        for a, b in zip(got, ref):
            assert np.isfinite(a).all()
            assert np.array_equal(a, b)
        assert (ref[2] == 0).sum() >= 1 and (ref[2] != 0).sum() >= kmax - 2


def test_noalloc_rejects_kmax_beyond_static_scratch(gpu):
    A, rng = _spd(200, seed=1)
    B = rng.normal(size=(65, 200))
    with pytest.raises(ValueError, match="at most 64"):
        _run("ssx_a_orthogonalize_noalloc", A, B, gpu, [])
