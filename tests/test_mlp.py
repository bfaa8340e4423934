"""The MLP oracle's backward is pinned by central finite differences of its float64 forward; the
device network (bf16 MFMA, FP32 accumulation) is checked against the oracle within bf16 limits."""

import numpy as np
import pytest

from oracle import mlp as O


def _net(seed=0, dims=(16, 32, 32, 32, 3)):
    rng = np.random.default_rng(seed)
    W = [rng.normal(size=(k, n)) / np.sqrt(k) for k, n in zip(dims[:-1], dims[1:])]
    b = [rng.normal(size=n) * 0.1 for n in dims[1:]]
    return W, b


def test_oracle_backward_matches_finite_differences():
    W, b = _net()
    rng = np.random.default_rng(1)
    X = rng.normal(size=(20, 16))
    T = rng.normal(size=(20, 3))

    def loss(W_, b_):
        out = O.forward(W_, b_, X, round_bf16=False)[-1]
        return O.mse(out, T)[0]

    acts = O.forward(W, b, X, round_bf16=False)
    _, g = O.mse(acts[-1], T)
    gW, gb = O.backward(W, acts, g, round_bf16=False)
    h = 1e-6
    for li in range(len(W)):
        for (i, j) in [(0, 0), (3, 1), (W[li].shape[0] - 1, W[li].shape[1] - 1)]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[li][i, j] += h
            Wm[li][i, j] -= h
            fd = (loss(Wp, b) - loss(Wm, b)) / (2 * h)
            assert abs(fd - gW[li][i, j]) < 1e-6 * max(1.0, abs(fd)), (li, i, j)
        bp = [x.copy() for x in b]
        bm = [x.copy() for x in b]
        bp[li][0] += h
        bm[li][0] -= h
        fd = (loss(W, bp) - loss(W, bm)) / (2 * h)
        assert abs(fd - gb[li][0]) < 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], dtype=np.float32)
    r = O.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12


def test_adam_first_step_is_sign_times_lr():
    w = np.array([1.0, -2.0])
    g = np.array([0.3, -5.0])
    w2, m, v = O.adam(w, g, np.zeros(2), np.zeros(2), 1, lr=0.01)
    np.testing.assert_allclose(w2, w - 0.01 * np.sign(g), rtol=1e-6)


def test_smooth_nodal_keeps_constants_and_damps_noise():
    """The mesh smoother of the learned correction: a constant field is a fixed point, a random
    field loses most of its point-to-point variation."""
    import torch
    from slope_stability_petsc4py_amd import mesh
    from slope_stability_petsc4py_amd.learn import NodeInfo, smooth_nodal
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=5.0, nz=2)
    elem = torch.from_numpy(np.ascontiguousarray(m.elem.T.astype(np.int64)))
    cnt = np.bincount(m.elem.reshape(-1), minlength=m.n_nodes).astype(np.float32)
    info = NodeInfo(None, None, None, 3, elem, torch.from_numpy((1.0 / np.maximum(cnt, 1))[:, None]))
    c = torch.full((m.n_nodes, 3), 2.5)
    assert torch.allclose(smooth_nodal(info, c, 4), c)
    v = torch.from_numpy(np.random.default_rng(0).normal(size=(m.n_nodes, 3)).astype(np.float32))
    sv = smooth_nodal(info, v, 4)
    assert float(sv.std()) < 0.5 * float(v.std())
    assert smooth_nodal(info, v, 0) is v


def test_smooth_nodal_transpose_is_adjoint():
    """<S v, g> == <v, S^T g> for the mesh smoother and its transpose (used for training through
    the smoother)."""
    import torch
    from slope_stability_petsc4py_amd import mesh
    from slope_stability_petsc4py_amd.learn import NodeInfo, smooth_nodal, smooth_nodal_T
    m = mesh.slope_3d(h=2.5, x1=10.0, x2=10.0, x3=10.0, y1=10.0, y2=10.0, z_len=5.0, nz=2)
    elem = torch.from_numpy(np.ascontiguousarray(m.elem.T.astype(np.int64)))
    cnt = np.bincount(m.elem.reshape(-1), minlength=m.n_nodes).astype(np.float64)
    info = NodeInfo(None, None, None, 3, elem, torch.from_numpy((1.0 / np.maximum(cnt, 1))[:, None]))
    rng = np.random.default_rng(1)
    v = torch.from_numpy(rng.normal(size=(m.n_nodes, 3)))
    g = torch.from_numpy(rng.normal(size=(m.n_nodes, 3)))
    for k in (1, 3):
        lhs = float((smooth_nodal(info, v, k) * g).sum())
        rhs = float((v * smooth_nodal_T(info, g, k)).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
