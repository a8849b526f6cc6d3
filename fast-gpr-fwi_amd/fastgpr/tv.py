"""Total-variation denoiser used as the prior inside ``GPRFWI.backward``
(reference: src/tv.py:135-217, called at compute.py:916-933).

ADMM for  min_X  |D X|_1 + 1/(2 lamda) |X - Y|^2  on the max-normalised
image, zero-padded by one cell, with periodic forward differences D and
soft-thresholding of the split variables.  The fixed-point sweep, defaults
(lamda = 0.02, rho = 1, 500 iterations, stop when ||X - X_prev|| < 1e-5)
and the float32 working precision are those of the reference.

Device tensors run the HIP implementation (``fgpr_tv``: every iteration and
the stopping test on the device, no host synchronisation per iteration);
``total_variation_torch`` is the same iteration in torch operations.
"""
from __future__ import annotations

import ctypes
import threading

import torch


def _dt(z: torch.Tensor, dim: int) -> torch.Tensor:
    """Adjoint of the periodic forward difference along ``dim``: z - z[+1]."""
    return z - torch.roll(z, shifts=-1, dims=dim)


def _d(x: torch.Tensor, dim: int) -> torch.Tensor:
    """Periodic backward difference along ``dim``: x - x[-1]."""
    return x - torch.roll(x, shifts=1, dims=dim)


def _shrink(t: torch.Tensor, thr: float) -> torch.Tensor:
    return torch.sign(t) * torch.clamp(torch.abs(t) - thr, min=0)


def total_variation(data: torch.Tensor, lamda: float = 0.02, rho: float = 1.0,
                    num_iter: int = 500, tol: float = 1e-5) -> torch.Tensor:
    """Denoise a (M, N, 1) or (M, N) model; returns (M, N, 1).

    Like the reference, the trailing singleton axis of ``data`` is squeezed
    in place.  Device tensors go through the HIP kernels.
    """
    if not data.is_cuda:
        return total_variation_torch(data, lamda, rho, num_iter, tol)
    from . import _native
    from . import trace as _trace
    from .solver import _stream
    data.squeeze_(-1)
    M, N = data.shape
    with _BUFFER_LOCK:          # one caller at a time enqueues on the shared buffers
        x, work, out = _buffers(data.device, M, N)
        x.copy_(data.detach())                # float32, rounded as .to(torch.float32) rounds
        nbytes = work.numel()
        with _trace.span(f"fgpr tv {M}x{N}"):
            _native.check(_native.lib.fgpr_tv(ctypes.c_void_p(x.data_ptr()), M, N, float(lamda),
                                              float(rho), int(num_iter), float(tol),
                                              ctypes.c_void_p(work.data_ptr()), nbytes,
                                              ctypes.c_void_p(out.data_ptr()),
                                              _stream(x.device)), "tv")
        return out.clone().unsqueeze(-1)


# Input, workspace and output buffers kept per (device, stream, grid): the
# same pointers on every call, so the library replays the call's captured
# graph instead of capturing a new one.  Calls on one stream are ordered, so
# a call may reuse the buffers of the one before; the result is a copy.  The
# lock keeps two threads on one stream from interleaving their copy-in,
# launches and copy-out on the same buffers.
_BUFFERS: dict = {}
_BUFFER_SETS = 8
_BUFFER_LOCK = threading.Lock()


def _buffers(device, M: int, N: int):
    from . import _native
    key = (str(device), torch.cuda.current_stream(device).cuda_stream, M, N)
    b = _BUFFERS.get(key)
    if b is None:
        nbytes = int(_native.lib.fgpr_tv_workspace_bytes(M, N))
        b = (torch.empty((M, N), dtype=torch.float32, device=device),
             torch.empty(nbytes, dtype=torch.uint8, device=device),
             torch.empty((M, N), dtype=torch.float32, device=device))
        if len(_BUFFERS) >= _BUFFER_SETS:
            _BUFFERS.pop(next(iter(_BUFFERS)))
        _BUFFERS[key] = b
    return b


def total_variation_torch(data: torch.Tensor, lamda: float = 0.02, rho: float = 1.0,
                          num_iter: int = 500, tol: float = 1e-5) -> torch.Tensor:
    """The same iteration in torch operations (any device)."""
    data.squeeze_(-1)
    dmax = data.max()
    y = data / (dmax + 1e-6)
    M, N = y.shape
    X = torch.zeros((M + 2, N + 2), device=data.device)
    X[1:-1, 1:-1] = y
    Y = X.clone()
    Zx = torch.zeros_like(X)
    Zy = torch.zeros_like(X)
    Ux = torch.zeros_like(X)
    Uy = torch.zeros_like(X)
    lr = lamda * rho
    for _ in range(num_iter):
        X_prev = X.clone()
        rhs = Y + lr * (_dt(Zx, 1) + _dt(Zy, 0)) - lamda * (_dt(Ux, 1) + _dt(Uy, 0))
        nb = X[:-2, 1:-1] + X[2:, 1:-1] + X[1:-1, :-2] + X[1:-1, 2:]
        X[1:-1, 1:-1] = (nb * lr + rhs[1:-1, 1:-1]) / (1 + 4 * lr)
        dxX = _d(X, 1)
        dyX = _d(X, 0)
        Zx = _shrink((Ux + rho * dxX) / rho, 1 / rho)
        Zy = _shrink((Uy + rho * dyX) / rho, 1 / rho)
        Ux += rho * (dxX - Zx)
        Uy += rho * (dyX - Zy)
        if torch.norm(X - X_prev) < tol:
            break
    return (X[1:-1, 1:-1] * dmax).unsqueeze(-1)
