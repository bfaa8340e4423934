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
