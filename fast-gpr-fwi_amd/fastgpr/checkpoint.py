"""Checkpoint / recompute gradients for wavefield histories that do not fit.

Storing Ez after every step costs nt * S * NI * NJ * 4 bytes (64 GB per shot
for a 2000 x 1000 grid over 8000 steps).  With an interval C the forward
sweep keeps only the full solver state (Ez, Hx, Hy and the CPML memory
variables of every shot) at steps C, 2C, ...  The adjoint then walks the
segments backwards: restore the state at the segment start, recompute that
segment's snapshots into a ring of C + 1 slots (slot = step mod (C + 1), so
one captured graph serves every segment), and run the adjoint over it.
Memory: about nt / C states plus C + 1 snapshot slots; cost: one extra
forward sweep.  The recomputed snapshots are bitwise those of the first
sweep, so the gradient equals the full-storage one exactly.

The reference stores the whole history (compute.py:551-553, 710-712); its
``nseg`` segmenting is unfinished (compute.py:543-553) and raises when used.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _native
from . import solver as _s
from ._native import check, lib


def interval(G, checkpoint) -> int | None:
    """Steps between saved states, or None for full snapshot storage.

    "auto" checkpoints only when the full history would not fit in the free
    device memory (counting what torch's caching allocator holds unused) less
    a reserve for the rest of the gradient (5 % of the device, at least
    8 GiB), with C about sqrt(3 nt) rounded up to a whole number of
    captured-graph chunks.  A history under a quarter of the device's memory
    is stored without asking the device for its free memory."""
    if checkpoint is None or checkpoint is False:
        return None
    nt = G.iterations
    if checkpoint == "auto":
        slot = G.S * G.NI * _native.pitch(G.NJ)
        if G.time_block and G.options.get("gradient", "reference") != "exact":
            _, _, ti, tj = _native.tb_tiles(G.NI, G.NJ, *G.time_block)
            slot = G.S * ti * tj * G.time_block[0] * 64          # time-blocked layout
        full = nt * slot * 4
        total = torch.cuda.get_device_properties(G.device).total_memory
        if full <= 0.25 * total:
            return None
        free, _ = torch.cuda.mem_get_info(G.device)
        free += torch.cuda.memory_reserved(G.device) - torch.cuda.memory_allocated(G.device)
        if full <= free - max(8 << 30, 0.05 * total):
            return None
        c = int(math.ceil(math.sqrt(3.0 * nt)))
        q = max(G.graph_chunk, 1)
        return max(q, ((c + q - 1) // q) * q)
    c = int(checkpoint)
    if c < 1:
        raise ValueError("checkpoint interval must be >= 1")
    return c


class States:
    """Solver states at steps k*C (k = 1 .. nseg-1), on the device."""

    def __init__(self, b: "_s.Batch", C: int):
        self.C = C
        self.nseg = (b.nt + C - 1) // C
        n = max(self.nseg - 1, 0)
        self.fields = torch.empty((n,) + tuple(b.fields.shape), dtype=torch.float32,
                                  device=b.fields.device)
        self.psi = torch.empty((n,) + tuple(b.psi.shape), dtype=torch.float32, device=b.psi.device)

    def save(self, k: int, b: "_s.Batch"):
        self.fields[k - 1].copy_(b.fields)
        self.psi[k - 1].copy_(b.psi)

    def restore(self, k: int, b: "_s.Batch"):
        if k == 0:
            b.reset_state()
        else:
            b.fields.copy_(self.fields[k - 1])
            b.psi.copy_(self.psi[k - 1])

    def nbytes(self) -> int:
        return self.fields.numel() * 4 + self.psi.numel() * 4


@dataclass
class Forward:
    batch: "_s.Batch"
    states: States
    coef: torch.Tensor
    scale: torch.Tensor


def forward(G, er, se, mr, C: int):
    """Forward sweep saving the state every C steps.  Returns (traces, Forward)."""
    dev = G.device
    se_f = torch.clamp(se.detach(), min=0, max=1e6)          # compute.py:729
    coef = _s.cell_coeffs(er, se_f, mr, G.dt32, G.dx32, False, dev)
    b = _s.Batch(G, store=False, tb=G.time_block)
    scale = _s.emitting_sources(G, b)
    st = States(b, C)
    p = b.problem(coef, srcscale=scale)
    for k in range(st.nseg):
        if k >= 1:
            st.save(k, b)
        check(lib.fgpr_forward(ctypes.byref(p), k * C, min((k + 1) * C, b.nt),
                               _s._MODES[G.exec_mode], G.graph_chunk, _s._stream(dev)), "forward")
    return b.traces, Forward(b, st, coef, scale)


def adjoint(G, fw: Forward, er, se, mr, grad_output, want_er=True, want_se=True, exact=False):
    """Adjoint sweep segment by segment with recomputed snapshots.

    Same results as ``solver.adjoint`` (reference mode) or
    ``solver.adjoint_exact`` (``exact=True``) on a full history."""
    dev = G.device
    b, st = fw.batch, fw.states
    C, nt, S, NI, NJ = st.C, b.nt, b.S, b.NI, b.NJ
    stream = _s._stream(dev)
    mode = _s._MODES[G.exec_mode]
    ring = torch.empty((C + 1,) + b.snap_slot_shape(), dtype=torch.float32, device=dev)
    traces_scratch = torch.empty_like(b.traces)
    pf = b.problem(fw.coef, srcscale=fw.scale)
    pf.snaps, pf.nsnap, pf.snap_mod = ring.data_ptr(), C + 1, C + 1
    pf.traces = traces_scratch.data_ptr()

    ba = _s.Batch(G, store=False, tb=None if exact else G.time_block)
    if exact:
        coef_a = fw.coef
    else:
        coef_a = _s.cell_coeffs(er, se, mr, G.dt32, G.dx32, True, dev)      # Ucgeta
    resid = grad_output.detach().to(device=dev, dtype=torch.float32).contiguous()
    if resid.shape != (S, nt, b.nrx):
        raise ValueError(f"grad_output shape {tuple(resid.shape)} != {(S, nt, b.nrx)}")
    ge = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    gs = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    if ba.tb:
        ba.gshot_buffer()
    pa = ba.problem(coef_a, resid=resid, grad_eps=ge, grad_sig=gs)
    pa.snaps, pa.nsnap, pa.snap_mod = ring.data_ptr(), C + 1, C + 1
    scratch = None
    if exact:
        scratch = torch.zeros((7, S, NI + 1, b.pitch), dtype=torch.float32, device=dev)
        pa.scratch = scratch.data_ptr()
        pa.tb_k = 0
    gmode = _native.GRAD_EXACT if exact else _native.GRAD_REFERENCE

    for k in reversed(range(st.nseg)):
        s0, e = k * C, min((k + 1) * C, nt)
        st.restore(k, b)
        if k >= 1:
            # Ez after step s0 - 1 is the saved state's Ez
            b.seed_slot(pf, (s0 - 1) % (C + 1), ring, st.fields[k - 1][0])
        check(lib.fgpr_forward(ctypes.byref(pf), s0, e, mode, G.graph_chunk, stream), "recompute")
        n_lo = s0 if exact else max(s0, 1)
        if e - 1 >= n_lo:
            if ba.tb:
                ba.gshot.zero_()
            check(lib.fgpr_adjoint(ctypes.byref(pa), e - 1, n_lo, gmode, mode, G.graph_chunk,
                                   stream), "adjoint")
    if exact:
        return _s.finish_exact(G, ge, gs, want_er, want_se)
    return _s.finish_reference(G, ge, gs, want_er, want_se)
