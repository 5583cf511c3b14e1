"""Device-side layouts for the KRR path.

Representations live in HBM as row-major ``(n, ld)`` FP64 tensors with ``ld`` rounded up to
a multiple of 16 and the pad columns zero: zero features change neither an L1 nor an L2
distance, and a 16-wide pad lets the MFMA path read whole 16-feature chunks.
"""

from __future__ import annotations

import numpy as np

FEATURE_PAD = 16


def padded_ld(d: int) -> int:
    return max(FEATURE_PAD, ((d + FEATURE_PAD - 1) // FEATURE_PAD) * FEATURE_PAD)


def device():
    import torch

    if not torch.cuda.is_available():
        from ._lib import QmlhError

        raise QmlhError("no HIP device visible: qml_amd runs only on the GPU")
    return torch.device("cuda", torch.cuda.current_device())


def rows_to_device(X, ld: int | None = None):
    """Copy an (n, d) array into a zero-padded (n, ld) FP64 device tensor."""
    import torch

    if isinstance(X, torch.Tensor):
        X = X.detach()
        n, d = X.shape
        ld = padded_ld(d) if ld is None else ld
        out = torch.zeros((n, ld), dtype=torch.float64, device=device())
        out[:, :d] = X.to(device=out.device, dtype=torch.float64)
        return out
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    ld = padded_ld(d) if ld is None else ld
    out = torch.zeros((n, ld), dtype=torch.float64, device=device())
    if n and d:
        out[:, :d].copy_(torch.from_numpy(X))
    return out


def ragged_to_device(Z, R, offsets):
    """Upload a ragged molecule batch (Z (total,), R (total, 3), offsets (nmol+1,))."""
    import torch

    dev = device()
    Zd = torch.from_numpy(np.ascontiguousarray(Z, dtype=np.float64)).to(dev)
    Rd = torch.from_numpy(np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)).to(dev)
    Od = torch.from_numpy(np.ascontiguousarray(offsets, dtype=np.int64)).to(dev)
    return Zd, Rd, Od
