"""Multiscale trace filtering (reference: src/multiscale.py).

``design_fir_filter`` builds the Hamming-windowed sinc low-pass taps in
float32 torch exactly as the reference does (multiscale.py:3-17, including
its centre tap 2 fc / fs for even tap counts); ``apply_filter`` reflect-pads
each trace by taps - 1 samples on the left and correlates (multiscale.py:
19-50).  Device tensors run the HIP kernel ``fgpr_fir`` directly on the
(shots, steps, receivers) trace layout (no permute/reshape copies); CPU
tensors use torch's conv1d like the reference.
"""
from __future__ import annotations

import ctypes

import torch


def design_fir_filter(cutoff: float, fs: float, numtaps: int) -> torch.Tensor:
    """Hamming-window FIR low-pass taps, normalised to unit sum."""
    n = torch.arange(numtaps, dtype=torch.float32)
    window = 0.54 - 0.46 * torch.cos(2 * torch.pi * n / (numtaps - 1))
    m = n - (numtaps - 1) / 2
    sinc = torch.sin(2 * torch.pi * (cutoff / fs) * m) / (torch.pi * m)
    sinc[(numtaps - 1) // 2] = 2 * cutoff / fs
    h = window * sinc
    return h / h.sum()


def _conv_torch(x3: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    L = h.numel()
    pad = torch.nn.functional.pad(x3, (L - 1, 0), mode="reflect")
    return torch.nn.functional.conv1d(pad, h.view(1, 1, -1), padding=0)


def apply_filter(data: torch.Tensor, fs: float, cutoff: float) -> torch.Tensor:
    """Low-pass ``data`` (1-D trace or (shots, steps, receivers) traces)."""
    numtaps = int(1 * (fs / cutoff))
    h = design_fir_filter(cutoff, fs, numtaps)
    if data.ndim not in (1, 3):
        raise ValueError(f"unsupported data dimension {data.ndim}: expected a 1-D or 3-D tensor")
    if not data.is_cuda:
        h = h.to(data.device)
        if data.ndim == 1:
            return _conv_torch(data.view(1, 1, -1), h).view(-1)
        S, nt, nrx = data.shape
        x3 = data.permute(0, 2, 1).reshape(-1, 1, nt)
        return _conv_torch(x3, h).view(S, nrx, nt).permute(0, 2, 1)
    from . import _native
    from .solver import _stream
    x = data.detach().to(torch.float32).contiguous()
    S, nt, nrx = (1, x.numel(), 1) if x.ndim == 1 else tuple(x.shape)
    if numtaps - 1 >= nt:
        raise ValueError(f"filter of {numtaps} taps needs traces longer than {numtaps - 1} samples")
    taps = h.to(x.device)
    out = torch.empty_like(x)
    _native.check(_native.lib.fgpr_fir(ctypes.c_void_p(x.data_ptr()), S, nt, nrx,
                                       ctypes.c_void_p(taps.data_ptr()), numtaps,
                                       ctypes.c_void_p(out.data_ptr()), _stream(x.device)), "fir")
    return out
