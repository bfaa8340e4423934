"""Helpers shared by the GPU tests."""

import torch


def misaligned(K):
    """The same CSR matrix with indices and values in views 4 / 8 bytes off the 16-byte grid (the
    SpMV then reads one entry per lane)."""
    from slope_stability_petsc4py_amd.device import DeviceCSR

    def shift(t):
        buf = torch.empty(t.numel() + 1, dtype=t.dtype, device=t.device)
        buf[1:].copy_(t)
        return buf[1:]
    return DeviceCSR(K.n_rows, K.indptr, shift(K.indices), shift(K.values),
                     None if K.values32 is None else shift(K.values32), K.lanes)
