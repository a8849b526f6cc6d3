"""Gradients whose wavefield history does not fit: keep what fits, recompute
the rest.

Storing Ez after every step costs nt * S * NI * NJ * 4 bytes (64 GB per shot
for a 2000 x 1000 grid over 8000 steps).  A ``History`` keeps the snapshots
in a ring of R slots (step n in slot n mod R):

* the forward sweep writes its last R snapshots into the ring and, before
  that, saves the full solver state (Ez, Hx, Hy and the CPML memory of every
  shot) at the start of each recompute segment (every C = R - 1 steps);
* the adjoint first walks the stored tail, steps nt-1 down to h = nt - R + 1,
  straight from the ring;
* then it walks the head segments [s0, e) backwards: restore the state at
  s0, recompute the segment's snapshots into the ring (the slot before s0 is
  seeded from the saved Ez), run the adjoint over it.

Only the h head steps are computed twice, so the recompute shrinks as the
ring grows: with R >= nt the history is stored whole (the reference's
behaviour) and nothing is recomputed.  The recomputed snapshots are bitwise
those of the first sweep, so every split gives the full-storage gradient
exactly.

The reference stores the whole history (compute.py:551-553, 710-712); its
``nseg`` segmenting is unfinished (compute.py:543-553) and raises when used.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _native
from . import solver as _s
from . import trace as _trace
from ._native import check, lib


@dataclass(frozen=True)
class History:
    """Snapshot ring of R slots for an nt-step gradient whose adjoint runs
    down to step ``lowest`` (1 for the reference gradient, 0 for the exact
    one)."""
    nt: int
    R: int
    lowest: int = 1

    def __post_init__(self):
        if self.R < 2 or self.nt < 1:
            raise ValueError(f"snapshot ring of {self.R} slots for {self.nt} steps")

    @property
    def C(self) -> int:
        """Steps per recompute segment."""
        return self.R - 1

    @property
    def h(self) -> int:
        """First step of the stored tail (adjoint steps h .. nt-1 read the
        ring as the forward left it)."""
        return max(self.nt - self.R + 1, 0)

    def tail(self):
        """(n_hi, n_lo) of the adjoint over the stored tail, or None."""
        lo = max(self.h, self.lowest)
        return (self.nt - 1, lo) if self.nt - 1 >= lo else None

    def segments(self) -> list:
        """Head segments [s0, e) in forward order whose adjoint range
        [max(s0, lowest), e - 1] is not empty."""
        out = []
        for s0 in range(0, self.h, self.C):
            e = min(s0 + self.C, self.h)
            if e - 1 >= max(s0, self.lowest):
                out.append((s0, e))
        return out

    def saves(self) -> list:
        """Steps at which the forward saves the solver state (the starts of
        the head segments after the first)."""
        return [s0 for s0, _ in self.segments() if s0 > 0]

    @property
    def full(self) -> bool:
        return self._head()[0] == 0

    def first_stored(self) -> int:
        """First step whose snapshot the forward must write to the ring."""
        return max(self.nt - self.R, 0)

    def _head(self) -> tuple:
        """(segments, steps they cover) of the head in closed form: every
        segment [s0, e) with s0 > 0 has a non-empty adjoint range; the first,
        [0, min(C, h)), has one when min(C, h) - 1 >= lowest."""
        if self.h <= 0:
            return 0, 0
        n = -(-self.h // self.C)
        first = min(self.C, self.h)
        if first - 1 >= self.lowest:
            return n, self.h
        return n - 1, self.h - first

    def n_saves(self) -> int:
        """len(saves()) without building the list."""
        return max(-(-self.h // self.C) - 1, 0) if self.h > 0 else 0

    def recompute_steps(self) -> int:
        return self._head()[1]

    def nbytes(self, slot_bytes: int, state_bytes: int) -> int:
        return self.R * slot_bytes + self.n_saves() * state_bytes


def plan(nt: int, slot_bytes: int, state_bytes: int, budget: int, lowest: int = 1) -> History:
    """The largest ring whose slots and saved states fit in ``budget`` bytes
    (the whole history when it fits).  Raises MemoryError when even the
    smallest footprint does not fit."""
    if nt * slot_bytes <= budget:
        return History(nt, nt, lowest)
    R = min(nt, int(budget // max(slot_bytes, 1)))
    best = None
    while R >= 2:
        hist = History(nt, R, lowest)
        over = hist.nbytes(slot_bytes, state_bytes) - budget
        if over <= 0:
            return hist
        best = hist
        R -= max(1, -(-over // max(slot_bytes, 1)))
    need = best.nbytes(slot_bytes, state_bytes) if best else 2 * slot_bytes
    raise MemoryError(f"snapshot history needs at least ~{need} bytes of device memory, "
                      f"{budget} available")


def slot_bytes(G, tb=None) -> int:
    """Bytes of one snapshot slot of G's batch (time-blocked layout when the
    time-blocked reference adjoint reads it)."""
    tb = G.time_block if tb is None else tb
    if tb and tb != "auto" and G.options.get("gradient", "reference") != "exact":
        return 4 * G.S * _native.tb_snap_floats(G.NI, G.NJ, tb)
    return 4 * G.S * G.NI * _native.pitch(G.NJ)


def state_bytes(G) -> int:
    """Bytes of one saved solver state (fields and CPML memory)."""
    pitch = _native.pitch(G.NJ)
    return 4 * (3 * G.S * (G.NI + 1) * pitch + max(_native.psi_floats(G.NI, G.NJ, G.pmlthick, G.S), 1))


def device_budget(device) -> int:
    """Device bytes a gradient's snapshot history and working buffers may
    take: the free memory, counting what torch's caching allocator holds
    unused, less a reserve for everything else (5 % of the device, at least
    8 GiB)."""
    device = torch.device(device)
    total = torch.cuda.get_device_properties(device).total_memory
    free, _ = torch.cuda.mem_get_info(device)
    free += torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    return int(free - max(8 << 30, 0.05 * total))


def history(G, checkpoint) -> History | None:
    """How the snapshots of G's gradient are kept; None = every snapshot in
    one buffer (the reference's full history).

    ``checkpoint`` = None / False: full history.  An int C: a ring of C + 1
    slots (recompute segments of C steps).  "auto": the full history when it
    fits in the free device memory (counting what torch's caching allocator
    holds unused) less a reserve for the rest of the gradient (5 % of the
    device, at least 8 GiB); otherwise the largest ring that does.  A history
    under a quarter of the device's memory is stored without asking the
    device for its free memory.  (Should the cached memory be too fragmented
    for the buffer, ``solver.device_empty`` releases it and tries once
    more.)"""
    if checkpoint is None or checkpoint is False:
        return None
    nt = G.iterations
    lowest = 0 if G.options.get("gradient", "reference") == "exact" else 1
    if checkpoint == "auto":
        sb = slot_bytes(G)
        full = nt * sb
        total = torch.cuda.get_device_properties(G.device).total_memory
        if full <= 0.25 * total:
            return None
        budget = device_budget(G.device)
        if full <= budget:
            return None
        return plan(nt, sb, state_bytes(G), budget, lowest)
    c = int(checkpoint)
    if c < 1:
        raise ValueError("checkpoint interval must be >= 1")
    return History(nt, min(c + 1, nt + 1), lowest)


class States:
    """Solver states saved at the given steps, on the device."""

    def __init__(self, b: "_s.Batch", steps, alloc=None):
        self.steps = list(steps)
        n = len(self.steps)
        if alloc is None:
            alloc = lambda shape, dt: torch.empty(shape, dtype=dt, device=b.fields.device)
        self.fields = alloc((n,) + tuple(b.fields.shape), torch.float32)
        self.psi = alloc((n,) + tuple(b.psi.shape), torch.float32)

    def save(self, step: int, b: "_s.Batch"):
        k = self.steps.index(step)
        self.fields[k].copy_(b.fields)
        self.psi[k].copy_(b.psi)

    def restore(self, step: int, b: "_s.Batch"):
        if step == 0:
            b.reset_state()
            return
        k = self.steps.index(step)
        b.fields.copy_(self.fields[k])
        b.psi.copy_(self.psi[k])

    def ez(self, step: int) -> torch.Tensor:
        """Saved Ez (S, NI+1, pitch) of the state at ``step``: Ez after step
        ``step`` - 1."""
        return self.fields[self.steps.index(step)][0]

    def nbytes(self) -> int:
        return self.fields.numel() * 4 + self.psi.numel() * 4


def forward_cuts(hist: History) -> list:
    """(n0, n1, save_at_n0, store) pieces of the first forward sweep: it stops
    where a state is saved and where the snapshots start to be kept."""
    marks = set(hist.saves()) | {hist.first_stored(), hist.nt}
    marks.discard(0)
    out, n = [], 0
    for m in sorted(marks):
        if m <= n:
            continue
        out.append((n, m, n in hist.saves(), n >= hist.first_stored()))
        n = m
    return out


@dataclass
class Forward:
    batch: "_s.Batch"
    states: States
    hist: History
    ring: torch.Tensor
    coef: torch.Tensor
    scale: torch.Tensor


def _ring_problem(b, coef, scale, ring, R):
    p = b.problem(coef, srcscale=scale)
    p.snaps, p.nsnap, p.snap_mod = ring.data_ptr(), R, R
    return p


def forward(G, er, se, mr, hist: History):
    """Forward sweep keeping the last R snapshots in the ring and the states
    the head segments restart from.  Returns (traces, Forward)."""
    dev = G.device
    se_f = torch.clamp(se.detach(), min=0, max=1e6)          # compute.py:729
    coef = _s.cell_coeffs(er, se_f, mr, G.dt32, G.dx32, False, dev)
    b = _s.Batch(G, store=False, tb=G.time_block)
    scale = _s.emitting_sources(G, b)
    st = States(b, hist.saves())
    ring = _s.device_empty((hist.R,) + b.snap_slot_shape(), torch.float32, dev)
    plain = b.problem(coef, srcscale=scale)
    kept = _ring_problem(b, coef, scale, ring, hist.R)
    with _trace.span(f"fgpr checkpoint forward {_trace.grid_tag(G)} ring {hist.R}"):
        for n0, n1, save, store in forward_cuts(hist):
            if save:
                st.save(n0, b)
            check(lib.fgpr_forward(ctypes.byref(kept if store else plain), n0, n1,
                                   _s._MODES[G.exec_mode], G.graph_chunk, _s._stream(dev)),
                  "forward")
    return b.traces, Forward(b, st, hist, ring, coef, scale)


def adjoint(G, fw: Forward, er, se, mr, grad_output, want_er=True, want_se=True, exact=False):
    """Adjoint over the stored tail, then segment by segment over the head
    with recomputed snapshots.  Same results as ``solver.adjoint``
    (reference mode) or ``solver.adjoint_exact`` (``exact=True``) on a full
    history."""
    dev = G.device
    b, st, hist, ring = fw.batch, fw.states, fw.hist, fw.ring
    nt, S, NI, NJ, R = b.nt, b.S, b.NI, b.NJ, hist.R
    stream = _s._stream(dev)
    mode = _s._MODES[G.exec_mode]
    traces_scratch = torch.empty_like(b.traces)
    pf = _ring_problem(b, fw.coef, fw.scale, ring, R)
    pf.traces = traces_scratch.data_ptr()

    ba = _s.Batch(G, store=False, tb=None if exact else G.time_block)
    coef_a = fw.coef if exact else _s.cell_coeffs(er, se, mr, G.dt32, G.dx32, True, dev)  # Ucgeta
    resid = grad_output.detach().to(device=dev, dtype=torch.float32).contiguous()
    if resid.shape != (S, nt, b.nrx):
        raise ValueError(f"grad_output shape {tuple(resid.shape)} != {(S, nt, b.nrx)}")
    ge = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    gs = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    if ba.tb:
        ba.gshot_buffer()
    pa = ba.problem(coef_a, resid=resid, grad_eps=ge, grad_sig=gs)
    pa.snaps, pa.nsnap, pa.snap_mod = ring.data_ptr(), R, R
    scratch = None
    if exact:
        scratch = torch.empty((4, S, NI + 1, b.pitch), dtype=torch.float32, device=dev)
        pa.scratch = scratch.data_ptr()
        pa.tb_k = 0
    gmode = _native.GRAD_EXACT if exact else _native.GRAD_REFERENCE

    def adj(n_hi, n_lo):
        if ba.tb:
            ba.gshot.zero_()
        with _trace.span(f"fgpr {'exact ' if exact else ''}adjoint {_trace.grid_tag(G)} "
                         f"steps {n_hi}-{n_lo}"):
            check(lib.fgpr_adjoint(ctypes.byref(pa), n_hi, n_lo, gmode, mode, G.graph_chunk,
                                   stream), "adjoint")

    t = hist.tail()
    if t is not None:
        adj(*t)
    for s0, e in reversed(hist.segments()):
        st.restore(s0, b)
        if s0 >= 1:
            b.seed_slot(pf, (s0 - 1) % R, ring, st.ez(s0))      # Ez after step s0 - 1
        with _trace.span(f"fgpr recompute {_trace.grid_tag(G)} steps {s0}-{e}"):
            check(lib.fgpr_forward(ctypes.byref(pf), s0, e, mode, G.graph_chunk, stream),
                  "recompute")
        adj(e - 1, max(s0, hist.lowest))
    if exact:
        return _s.finish_exact(G, ge, gs, want_er, want_se)
    return _s.finish_reference(G, ge, gs, want_er, want_se)
