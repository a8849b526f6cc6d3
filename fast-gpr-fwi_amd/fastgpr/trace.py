"""roctx ranges around the solver's host phases, for rocprofv3's marker
trace (``rocprofv3 --marker-trace --kernel-trace ...``); off unless the
environment sets FASTGPR_ROCTX=1 when this module is first imported.

A range covers the host issuing a phase: sweeps are asynchronous, so a range
closes once its launches are queued (or its graphs launched), not when the
device finishes them; the kernel trace gives the device side.  Ranges:
``fgpr forward`` / ``fgpr adjoint`` / ``fgpr exact adjoint`` (one sweep of a
batch), ``fgpr checkpoint forward`` and ``fgpr recompute`` (checkpointed
runs), ``fgpr tv``, ``fgpr uneven capture`` / ``fgpr uneven replay`` and
``fgpr admit`` (serving).  The reference has no tracing (``time.time()``
around its TV call, compute.py:917, 933).
"""
from __future__ import annotations

import contextlib
import os

import torch

ENABLED = os.environ.get("FASTGPR_ROCTX", "0") == "1"


@contextlib.contextmanager
def span(name: str):
    """One roctx range (torch.cuda.nvtx is roctx on ROCm)."""
    if not ENABLED:
        yield
        return
    torch.cuda.nvtx.range_push(name)
    try:
        yield
    finally:
        torch.cuda.nvtx.range_pop()


def grid_tag(G) -> str:
    return f"{G.NI}x{G.NJ} S{G.S} nt{G.iterations}"
