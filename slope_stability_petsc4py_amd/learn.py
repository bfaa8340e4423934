"""Learned continuation predictor: a per-node MLP trained and evaluated with the project's own
HIP kernels (bf16 matrix-core GEMMs with FP32 accumulation, tanh, Adam).

It fills the predictor slot of the SSR continuation (continuation/indirect.py:2703-2830).  The
network predicts a correction to the secant increment, per node:

    dU_pred(node) = sec(node) + loc(node) * smooth(MLP(features(node)))

with sec = a * dU_{k-1} the secant increment, a = (omega_target - omega) / (omega - omega_old),
loc = |sec| + 0.02 * |a| * s the local scale (s = max |dU_{k-1}|) and smooth a few passes of
element/nodal averaging over the mesh.
Features (16 per node): normalised coordinates (3), the node's material strength (c0 scaled,
tan phi) (2), lambda and its last increment (2), dU_{k-1}/s, dU_{k-2}/s and the secant increment
a*dU_{k-1}/s (3 each, zero-padded in 2D).  The last layer starts at zero, so an untrained network
is exactly the secant.  The continuation keeps plain Newton as the arbiter: a guess whose first
residual is not below the secant's is rejected.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import ptr, stream_ptr

N_FEAT = 16


def _bf16(t: torch.Tensor) -> torch.Tensor:
    out = torch.empty(t.shape, dtype=torch.int16, device=t.device)
    _lib.call("ssx_f32_to_bf16", t.numel(), ptr(t), ptr(out), stream_ptr())
    return out


class NodeMLP:
    """Fully connected n_in -> hidden x (layers-1) -> n_out, tanh on hidden layers."""

    def __init__(self, n_in: int = N_FEAT, hidden: int = 128, layers: int = 4, n_out: int = 3, seed: int = 0,
                 device=None, zero_last: bool = True):
        dev = device or torch.device("cuda", 0)
        g = torch.Generator().manual_seed(seed)
        dims = [n_in] + [hidden] * (layers - 1) + [n_out]
        self.dims = dims
        self.W, self.b = [], []
        for li, (k, n) in enumerate(zip(dims[:-1], dims[1:])):
            w = torch.randn(k, n, generator=g, dtype=torch.float32) * (1.0 / np.sqrt(k))
            if zero_last and li == len(dims) - 2:
                w.zero_()
            self.W.append(w.to(dev).contiguous())
            self.b.append(torch.zeros(n, dtype=torch.float32, device=dev))
        self.Wb = [_bf16(w) for w in self.W]
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.gW = [torch.zeros_like(w) for w in self.W]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.step_count = 0
        self.device = dev

    # forward / backward ----------------------------------------------------------------------
    def forward(self, X: torch.Tensor, keep: bool = True) -> torch.Tensor:
        """X (M, n_in) float32 -> (M, n_out) float32."""
        M = X.shape[0]
        xb = _bf16(X.contiguous())
        acts = [(X, xb)]
        L = len(self.W)
        for li in range(L):
            K, N = self.dims[li], self.dims[li + 1]
            Y = torch.empty((M, N), dtype=torch.float32, device=X.device)
            Yb = torch.empty((M, N), dtype=torch.int16, device=X.device)
            _lib.call("ssx_gemm_bf16", M, N, K, ptr(acts[-1][1]), K, 1, ptr(self.Wb[li]), N, 1, ptr(Y), N,
                      ptr(self.b[li]), 1 if li < L - 1 else 0, ptr(Yb), 0.0, stream_ptr())
            acts.append((Y, Yb))
        if keep:
            self._acts = acts
        return acts[-1][0]

    def backward(self, dOut: torch.Tensor) -> None:
        """Gradients of sum(dOut * out) w.r.t. all weights (after a forward with keep=True)."""
        acts = self._acts
        M = dOut.shape[0]
        L = len(self.W)
        dZ = dOut.contiguous()
        dZb = _bf16(dZ)
        for li in range(L - 1, -1, -1):
            K, N = self.dims[li], self.dims[li + 1]
            Xb = acts[li][1]
            # dW = X^T dZ : (K x M) (M x N)
            _lib.call("ssx_gemm_bf16", K, N, M, ptr(Xb), 1, K, ptr(dZb), N, 1, ptr(self.gW[li]), N, None, 0, None,
                      0.0, stream_ptr())
            _lib.call("ssx_colsum_f32", M, N, ptr(dZ), ptr(self.gb[li]), stream_ptr())
            if li == 0:
                break
            # dX = dZ W^T : (M x N) (N x K), then through tanh of the previous layer
            dX = torch.empty((M, K), dtype=torch.float32, device=dOut.device)
            _lib.call("ssx_gemm_bf16", M, K, N, ptr(dZb), N, 1, ptr(self.Wb[li]), 1, N, ptr(dX), K, None, 0, None,
                      0.0, stream_ptr())
            Yprev = acts[li][0]
            dZ = torch.empty_like(dX)
            dZb = torch.empty((M, K), dtype=torch.int16, device=dOut.device)
            _lib.call("ssx_tanh_bwd", M * K, ptr(dX), ptr(Yprev), ptr(dZ), ptr(dZb), stream_ptr())

    def adam(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.step_count += 1
        for li in range(len(self.W)):
            _lib.call("ssx_adam_f32", self.W[li].numel(), ptr(self.W[li]), ptr(self.gW[li]), ptr(self.mW[li]),
                      ptr(self.vW[li]), lr, b1, b2, eps, self.step_count, ptr(self.Wb[li]), stream_ptr())
            _lib.call("ssx_adam_f32", self.b[li].numel(), ptr(self.b[li]), ptr(self.gb[li]), ptr(self.mb[li]),
                      ptr(self.vb[li]), lr, b1, b2, eps, self.step_count, None, stream_ptr())

    def mse(self, pred: torch.Tensor, target: torch.Tensor, mask: torch.Tensor | None = None):
        """(loss, dLoss/dpred) with the mean over the masked entries."""
        n = pred.numel()
        cnt = float(mask.sum().item()) if mask is not None else float(n)
        g = torch.empty_like(pred)
        part = torch.empty((n + 255) // 256, dtype=torch.float32, device=pred.device)
        _lib.call("ssx_mse_f32", n, ptr(pred), ptr(target), ptr(mask), 1.0 / max(cnt, 1.0), ptr(g), ptr(part),
                  stream_ptr())
        return float(part.sum().item()), g

    def save(self, path) -> None:
        st = self.state_dict()
        np.savez(path, dims=np.asarray(st["dims"], dtype=np.int64),
                 **{f"W{i}": w for i, w in enumerate(st["W"])}, **{f"b{i}": b for i, b in enumerate(st["b"])})

    @classmethod
    def load(cls, path, device=None) -> "NodeMLP":
        d = np.load(path, allow_pickle=False)
        dims = [int(x) for x in d["dims"]]
        net = cls(n_in=dims[0], hidden=dims[1], layers=len(dims) - 1, n_out=dims[-1], device=device)
        net.load_state_dict({"W": [d[f"W{i}"] for i in range(len(dims) - 1)],
                             "b": [d[f"b{i}"] for i in range(len(dims) - 1)]})
        return net

    def state_dict(self):
        return {"dims": self.dims, "W": [w.cpu().numpy() for w in self.W], "b": [b.cpu().numpy() for b in self.b]}

    def load_state_dict(self, st):
        for li, (w, b) in enumerate(zip(st["W"], st["b"])):
            self.W[li].copy_(torch.from_numpy(np.asarray(w, dtype=np.float32)))
            self.b[li].copy_(torch.from_numpy(np.asarray(b, dtype=np.float32)))
            self.Wb[li] = _bf16(self.W[li])


# ---------------------------------------------------------------------------------------------
# features, samples, training, the continuation hook
# ---------------------------------------------------------------------------------------------

@dataclass
class NodeInfo:
    """Static per-node inputs of one mesh (device tensors)."""

    coord: torch.Tensor     # (n_nodes, 3) normalised
    mat: torch.Tensor       # (n_nodes, 2)
    free: torch.Tensor      # (n_nodes, 3) float mask (zero-padded in 2D)
    dim: int
    elem: torch.Tensor | None = None       # (n_elem, n_p) int64 node table (smoothing)
    inv_count: torch.Tensor | None = None  # (n_nodes, 1) 1 / number of elements at each node


def node_info(mesh, materials, device=None) -> NodeInfo:
    from .materials import as_specs
    dev = device or torch.device("cuda", 0)
    c = np.zeros((mesh.n_nodes, 3))
    c[:, : mesh.dim] = mesh.coord.T
    c = (c - c.mean(0)) / max(float(np.ptp(mesh.coord)), 1e-12)
    specs = as_specs(materials)
    mid = np.asarray(mesh.material, dtype=np.int64)
    if len(specs) == 1:
        mid = np.zeros_like(mid)
    cm = np.array([[s.c0 / 20.0, np.tan(np.deg2rad(s.phi))] for s in specs])
    nm = np.zeros((mesh.n_nodes, 2))
    # node material: that of the last element listing it (vectorised scatter)
    nm[mesh.elem.T.reshape(-1)] = np.repeat(cm[mid], mesh.elem.shape[0], axis=0)
    free = np.zeros((mesh.n_nodes, 3))
    free[:, : mesh.dim] = mesh.q_mask.T
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)  # noqa: E731
    elem = np.ascontiguousarray(np.asarray(mesh.elem, dtype=np.int64).T)
    cnt = np.bincount(elem.reshape(-1), minlength=mesh.n_nodes).astype(np.float32)
    return NodeInfo(t(c), t(nm), t(free), mesh.dim, torch.from_numpy(elem).to(dev),
                    t((1.0 / np.maximum(cnt, 1.0))[:, None]))


def smooth_nodal(info: NodeInfo, v: torch.Tensor, passes: int) -> torch.Tensor:
    """Low-pass filter of a nodal field (n_nodes, 3): `passes` rounds of element averaging
    (mean over an element's nodes) followed by nodal averaging over the elements at each node."""
    if passes <= 0 or info.elem is None:
        return v
    flat = info.elem.reshape(-1)
    n_p = info.elem.shape[1]
    for _ in range(passes):
        ve = v[info.elem].mean(dim=1)                                  # (n_elem, 3)
        acc = torch.zeros_like(v)
        acc.index_add_(0, flat, ve.repeat_interleave(n_p, dim=0))
        v = acc * info.inv_count
    return v


def smooth_nodal_T(info: NodeInfo, g: torch.Tensor, passes: int) -> torch.Tensor:
    """Transpose of `smooth_nodal` (for gradients through the smoother): one pass of smooth_nodal
    is S = diag(1/count) E^T E / n_p with E the element-node incidence, so
    S^T g = E^T E (g / count) / n_p."""
    if passes <= 0 or info.elem is None:
        return g
    flat = info.elem.reshape(-1)
    n_p = info.elem.shape[1]
    for _ in range(passes):
        ge = (g * info.inv_count)[info.elem].sum(dim=1) / n_p      # (n_elem, 3)
        acc = torch.zeros_like(g)
        acc.index_add_(0, flat, ge.repeat_interleave(n_p, dim=0))
        g = acc
    return g


def _pad3(v: torch.Tensor, dim: int) -> torch.Tensor:
    """flat dof vector (n_nodes*dim) -> (n_nodes, 3) float32."""
    x = v.view(-1, dim).to(torch.float32)
    if dim == 3:
        return x
    return torch.nn.functional.pad(x, (0, 3 - dim))


LOCAL_FLOOR = 0.02   # floor of the local scale, as a fraction of the largest increment


def features(info: NodeInfo, dU1: torch.Tensor, dU2: torch.Tensor, a: float, lam: float, dlam: float):
    """(X (n_nodes, 16), scale s, secant increment (n_nodes, 3), local scale (n_nodes, 3)).

    The network's output is a correction RELATIVE to the local secant increment: the increment is
    sec + loc * out with loc = |sec| + LOCAL_FLOOR * s, so the error of the prediction shrinks with
    the local motion (far from the failure zone the guess stays close to the secant instead of
    carrying errors of the size of the largest increment, which the stiffness would amplify in the
    residual)."""
    d1 = _pad3(dU1, info.dim)
    d2 = _pad3(dU2, info.dim)
    s = float(d1.abs().max().item()) or 1.0
    sec = a * d1
    loc = sec.abs() + LOCAL_FLOOR * abs(a) * s + 1e-30
    n = d1.shape[0]
    X = torch.empty((n, N_FEAT), dtype=torch.float32, device=d1.device)
    X[:, 0:3] = info.coord
    X[:, 3:5] = info.mat
    X[:, 5] = lam
    X[:, 6] = dlam * 10.0
    X[:, 7:10] = d1 / s
    X[:, 10:13] = d2 / s
    X[:, 13:16] = sec / s
    return X, s, sec, loc


def samples_from_history(info: NodeInfo, step_U, omega_hist, lambda_hist):
    """Training pairs (X, target correction, mask) from one accepted-state history."""
    out = []
    for k in range(2, len(step_U) - 1):
        U0, U1, U2, U3 = step_U[k - 2], step_U[k - 1], step_U[k], step_U[k + 1]
        dU2, dU1, dU_next = U1 - U0, U2 - U1, U3 - U2
        den = omega_hist[k] - omega_hist[k - 1]
        a = 0.0 if abs(den) < 1e-30 else (omega_hist[k + 1] - omega_hist[k]) / den
        X, s, sec, loc = features(info, dU1, dU2, a, float(lambda_hist[k]), float(lambda_hist[k] - lambda_hist[k - 1]))
        target = torch.clamp((_pad3(dU_next, info.dim) - sec) / loc, -4.0, 4.0)
        out.append((X, target.contiguous(), info.free))
    return out


def train(net: NodeMLP, samples, *, epochs: int = 200, lr: float = 2e-3, log=None, info: NodeInfo | None = None,
          smoothing: int = 0) -> list[float]:
    """Adam on the masked MSE.  With `info` and `smoothing` > 0 the loss is taken on the smoothed
    prediction (the field the predictor actually adds), the gradient flowing back through the
    transpose of the smoother."""
    losses = []
    for ep in range(epochs):
        tot = 0.0
        for X, T, mask in samples:
            pred = net.forward(X)
            if info is not None and smoothing > 0:
                sp = smooth_nodal(info, pred, smoothing).contiguous()
                loss, g = net.mse(sp, T, mask)
                g = smooth_nodal_T(info, g, smoothing).contiguous()
            else:
                loss, g = net.mse(pred, T, mask)
            net.backward(g)
            net.adam(lr)
            tot += loss
        losses.append(tot / max(len(samples), 1))
        if log is not None and (ep % 20 == 0 or ep == epochs - 1):
            log(ep, losses[-1])
    return losses


class LearnedPredictor:
    """Continuation hook: returns U + predicted increment, or None while fewer than 3 states exist.
    The increment is the secant increment plus the network's relative correction (smoothed over
    the mesh by `smoothing` passes of element/nodal averaging) times the local secant scale."""

    def __init__(self, net: NodeMLP, info: NodeInfo, smoothing: int = 2):
        self.net = net
        self.info = info
        self.smoothing = int(smoothing)
        self.calls = 0

    def __call__(self, ctx):
        hist = ctx.get("states")
        if not hist or len(hist) < 3:
            return None
        (U0, w0, l0), (U1, w1, l1), (U2, w2, l2) = hist[-3], hist[-2], hist[-1]
        den = w2 - w1
        a = 0.0 if abs(den) < 1e-30 else (ctx["omega_target"] - w2) / den
        X, s, sec, loc = features(self.info, U2 - U1, U1 - U0, a, l2, l2 - l1)
        # per-node predictions carry point-to-point noise; the relative correction is low-pass
        # filtered over the mesh before it is scaled by the local secant increment
        corr = smooth_nodal(self.info, self.net.forward(X, keep=False), self.smoothing)
        inc = (sec + loc * corr) * self.info.free
        self.calls += 1
        dim = self.info.dim
        return U2 + inc[:, :dim].reshape(-1).to(torch.float64)
