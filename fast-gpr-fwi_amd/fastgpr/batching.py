"""Uneven batches: many solves -- different grids, shot counts, step counts
and phases -- advanced together, one multi-problem time-blocked launch per
tick.

A ``Job`` is one gradient (forward sweep, residual, adjoint sweep, FP64
reduction and L2 normalisation) or one forward modelling of one model with
its shots.  The ``Scheduler`` asks every active job for its next block of up
to K steps and packs all blocks of the same kind (forward-like / adjoint) and
time-blocking variant into one launch of ``k_tb_multi``.  Jobs keep their own
buffers, so each result is bitwise the one the job would produce alone.
Histories that do not fit are kept in part (``checkpoint.History``): the
last snapshots stay in a ring, the adjoint reads them directly, and the
earlier ones are recomputed segment by segment from saved solver states.

A job does no device work when it is built: everything it does -- clearing
its state, the coefficients, every block and every phase change -- is an
operation queued in stream order, so a job can live in memory that another
job is still using until its first operation runs (``UnevenBatch`` waves).
The ``Scheduler`` runs ticks as they are planned (serving, where the job set
changes between ticks) or plans a fixed job set to its end, uploads every
launch table at once and captures the whole schedule in one HIP graph
(``Scheduler.capture``): replaying it runs the batch with no host work per
tick.

The reference solves one model per call and one step per launch
(compute.py:574-712, 1048-1172); this module has no counterpart there.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from . import solver as _s
from . import trace as _trace
from .checkpoint import History, States, plan
from ._native import check, lib

MULTI_VARIANTS = [(32, 512, 2), (32, 512, 4), (32, 512, 6), (64, 1024, 2), (64, 1024, 4),
                  (64, 1024, 6), (64, 1024, 8), (64, 512, 2), (64, 512, 4), (64, 512, 6),
                  (64, 512, 8)]


ALIGN = 256
_SLACK = 64 * ALIGN          # alignment padding of a job's pieces


def _aligned(n: int) -> int:
    return (int(n) + ALIGN - 1) // ALIGN * ALIGN


def job_bytes(NI, NJ, S, nt, nrx, P, R, gradient=True, tb=(64, 1024, 6)) -> int:
    """Device bytes of the large buffers one job holds: forward-side fields,
    CPML memory and their ping-pong partners, coefficients and traces; for a
    gradient also the adjoint side, a snapshot ring of R slots in the
    time-blocked layout of variant ``tb`` with the states its recompute
    segments restart from, residual and recomputed traces, per-shot
    and summed FP64 gradients and the normalised outputs.  ``R`` may also be
    a ``History``.  Includes the alignment padding of the pieces."""
    pitch = _native.pitch(NJ)
    slot = S * _native.tb_snap_floats(NI, NJ, tb) * 4
    field = 3 * S * (NI + 1) * pitch * 4
    psi = max(_native.psi_floats(NI, NJ, P, S), 1) * 4
    coef = 5 * (NI + 1) * pitch * 4
    tr = S * nt * nrx * 4
    b = 2 * (field + psi) + coef + tr
    if gradient:
        hist = R if isinstance(R, History) else History(nt, max(2, min(int(R), nt + 1)))
        b += 2 * (field + psi) + coef + 2 * tr
        b += hist.nbytes(slot, field + psi)
        b += 2 * S * NI * NJ * 8 + 2 * NI * NJ * 8 + 2 * NI * NJ * 4
    return int(b) + _SLACK


def _slot_state(NI, NJ, S, P, tb):
    pitch = _native.pitch(NJ)
    slot = S * _native.tb_snap_floats(NI, NJ, tb) * 4
    state = 3 * S * (NI + 1) * pitch * 4 + max(_native.psi_floats(NI, NJ, P, S), 1) * 4
    return slot, state


def job_history(NI, NJ, S, nt, nrx, P, budget, tb=(64, 1024, 6)) -> History:
    """The largest snapshot ring with which a gradient job fits in
    ``budget`` bytes (the whole history when it fits)."""
    slot, state = _slot_state(NI, NJ, S, P, tb)
    h2 = History(nt, 2)
    rest = job_bytes(NI, NJ, S, nt, nrx, P, h2, True, tb) - h2.nbytes(slot, state)
    return plan(nt, slot, state, budget - rest, lowest=1)


def plan_histories(specs, budget: int) -> list:
    """Snapshot histories for gradient jobs that run together within
    ``budget`` device bytes.  ``specs``: (NI, NJ, S, nt, nrx, P, tb) per job.
    Every job's fixed footprint is reserved first; then whole histories go
    to the jobs in increasing size while they fit, and the remaining jobs
    share what is left in proportion to their histories, each keeping the
    largest ring its share holds (the recompute per byte kept is about the
    same for every grid, so this keeps the total recompute near its least).
    Raises MemoryError when the jobs do not fit."""
    fixed, hb, sb, stb = [], [], [], []
    for NI, NJ, S, nt, nrx, P, tb in specs:
        slot, state = _slot_state(NI, NJ, S, P, tb)
        h2 = History(nt, 2)
        fixed.append(job_bytes(NI, NJ, S, nt, nrx, P, h2, True, tb) - h2.nbytes(slot, state))
        hb.append(nt * slot)
        sb.append(slot)
        stb.append(state)
    left = budget - sum(fixed)
    if left <= 0:
        raise MemoryError(f"jobs need {sum(fixed)} bytes before their histories, budget {budget}")
    n = len(specs)
    full = [False] * n
    for k in sorted(range(n), key=lambda k: hb[k]):
        if hb[k] > left:
            break
        full[k] = True
        left -= hb[k]
    out = [None] * n
    total = sum(hb[k] for k in range(n) if not full[k])
    for k in range(n):
        nt = specs[k][3]
        out[k] = History(nt, nt) if full[k] else plan(nt, sb[k], stb[k], int(left * hb[k] / total))
    return out


class Region:
    """Bump allocator over a byte range of one device buffer: typed,
    256-byte aligned views, uninitialised.  Raises MemoryError past its end."""

    def __init__(self, buf: torch.Tensor, offset: int, nbytes: int):
        self.buf, self.base, self.end, self.top = buf, int(offset), int(offset) + int(nbytes), int(offset)

    def __call__(self, shape, dtype) -> torch.Tensor:
        shape = tuple(int(x) for x in shape)
        n = int(np.prod(shape)) * torch.empty((), dtype=dtype).element_size()
        if self.top + n > self.end:
            raise MemoryError(f"region of {self.end - self.base} bytes exhausted")
        t = self.buf[self.top:self.top + max(n, 1)][:n].view(dtype).view(shape)
        self.top += _aligned(max(n, 1))
        return t

    @property
    def used(self) -> int:
        return self.top - self.base


class Arena:
    """One device allocation carved into blocks: best fit over a free list
    kept sorted by offset, neighbouring free blocks merged on release.  The
    caller owns the ordering: a block is released only after the device work
    that uses it has finished (or is ordered before the next user's work on
    the same stream)."""

    def __init__(self, device, nbytes: int):
        nbytes = _aligned(nbytes)
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.size = nbytes
        self.free_list = [(0, nbytes)]          # (offset, size), sorted, disjoint, merged
        self.used = {}

    def alloc(self, nbytes: int):
        """Offset of a block of at least ``nbytes`` bytes, or None."""
        need = _aligned(nbytes)
        best = None
        for k, (off, size) in enumerate(self.free_list):
            if size >= need and (best is None or size < self.free_list[best][1]):
                best = k
        if best is None:
            return None
        off, size = self.free_list[best]
        if size == need:
            del self.free_list[best]
        else:
            self.free_list[best] = (off + need, size - need)
        self.used[off] = need
        return off

    def release(self, off: int):
        size = self.used.pop(off)
        fl = self.free_list
        k = 0
        while k < len(fl) and fl[k][0] < off:
            k += 1
        fl.insert(k, (off, size))
        if k + 1 < len(fl) and fl[k][0] + fl[k][1] == fl[k + 1][0]:
            fl[k] = (fl[k][0], fl[k][1] + fl[k + 1][1])
            del fl[k + 1]
        if k > 0 and fl[k - 1][0] + fl[k - 1][1] == fl[k][0]:
            fl[k - 1] = (fl[k - 1][0], fl[k - 1][1] + fl[k][1])
            del fl[k]

    def region(self, off: int) -> Region:
        return Region(self.buf, off, self.used[off])

    def largest_free(self) -> int:
        return max((size for _, size in self.free_list), default=0)

    def free_bytes(self) -> int:
        return sum(size for _, size in self.free_list)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class Job:
    """One solve advanced block by block (see module docstring).

    ``history``: None (keep every snapshot), an int C (a ring of C + 1
    slots) or a ``checkpoint.History``.  ``alloc(shape, dtype)`` supplies the
    large buffers (default: torch tensors); nothing is written to them until
    the job's first operation runs on its stream.  ``keep`` (default
    ``alloc``) supplies the results: the traces and the normalised
    gradients.

    ``shots`` = (lo, hi) and ``group`` (a ``ShotGroup``) make the job one
    shot group of a model split over several jobs: it computes shots lo..hi-1
    into its slices of the group's traces and per-shot partial gradients and
    leaves the shot reduction and the normalisation to the group's
    ``finish_ops`` (one pass in shot order once every group is done, as for
    one batch).  Such a job must keep its whole snapshot history (one adjoint
    phase)."""

    def __init__(self, G, er, se, mr, obs=None, tb=(64, 1024, 6), history=None,
                 want=("er", "se"), tag=None, alloc=None, keep=None, shots=None, group=None):
        if tuple(tb) not in MULTI_VARIANTS:
            raise ValueError(f"time-blocking variant {tb} has no multi-problem kernel")
        dev = G.device
        self.G = G
        self.tag = tag
        self.tb = tuple(tb)
        self.K = self.tb[2]
        self.gradient = obs is not None
        self.want = want
        nt = G.iterations
        self.nt = nt
        if alloc is None:
            alloc = lambda shape, dt: torch.empty(shape, dtype=dt, device=dev)
        keep = alloc if keep is None else keep
        if isinstance(history, History):
            hist = history
        else:
            hist = History(nt, nt + 1 if history is None else max(2, min(int(history) + 1, nt + 1)))
        if hist.nt != nt or hist.lowest != 1:
            raise ValueError("history plan does not match the job")
        self.hist = hist
        NI, NJ = G.NI, G.NJ
        f32 = lambda x: x.detach().reshape(NI, NJ).to(device=dev, dtype=torch.float32).contiguous()
        # model copies owned by the job (read by its coefficient operations)
        self.er, self.se, self.mr = f32(er), f32(se), f32(mr)
        self.se_f = torch.empty_like(self.se)
        nrx = int(G.receiver.shape[1])
        lo, hi = shots if shots is not None else (0, G.S)
        self.group = group
        traces = group.traces[lo:hi] if group is not None else keep((hi - lo, nt, nrx), torch.float32)
        self.fb = _s.Batch(G, store=False, tb=self.tb, alloc=alloc, traces=traces, shots=(lo, hi))
        self.scale = _s.emitting_sources(G, self.fb)
        S = self.fb.S
        pitch = self.fb.pitch
        self.coef_f = alloc((5, NI + 1, pitch), torch.float32)
        self.pf_plain = self.fb.problem(self.coef_f, srcscale=self.scale)   # no snapshots
        self.result = None
        self.done_event = None
        self.cancelled = False
        self.finite_check = False        # queue a finiteness test of the results at the end
        self._finite = None
        self.stream_idx = 0
        self.ip = 0
        self.n = 0
        self.parity = 0
        self.marks = set()
        if not self.gradient:
            self.phases = [("fwd", 0, nt)]
            return
        R = hist.R
        if group is not None and not hist.full:
            raise ValueError("a shot group of a split model must keep its whole history")
        self.obs = obs.detach().to(device=dev, dtype=torch.float32)[lo:hi].reshape(S, nt, nrx).contiguous()
        self.ring = alloc((R,) + self.fb.snap_slot_shape(), torch.float32)
        self.pf = self.fb.problem(self.coef_f, srcscale=self.scale)
        self.pf.snaps, self.pf.nsnap, self.pf.snap_mod = self.ring.data_ptr(), R, R
        self.rec_traces = alloc((S, nt, nrx), torch.float32)
        self.pf_rec = self.fb.problem(self.coef_f, srcscale=self.scale)
        self.pf_rec.snaps, self.pf_rec.nsnap, self.pf_rec.snap_mod = self.ring.data_ptr(), R, R
        self.pf_rec.traces = self.rec_traces.data_ptr()
        self.states = States(self.fb, hist.saves(), alloc=alloc)
        self.first_stored = hist.first_stored()
        self.marks = set(hist.saves()) | {self.first_stored}
        self.resid = alloc((S, nt, nrx), torch.float32)
        self.ab = _s.Batch(G, store=False, tb=self.tb, alloc=alloc, traces=self.resid, shots=(lo, hi))
        self.coef_a = alloc((5, NI + 1, pitch), torch.float32)
        if group is None:
            self.ge = alloc((NI, NJ), torch.float64)
            self.gs = alloc((NI, NJ), torch.float64)
            self.ab.gshot_buffer(zero=False)
            self.out = {name: keep((NI, NJ, 1), torch.float32) for name in ("er", "se") if name in want}
            self.pa = self.ab.problem(self.coef_a, resid=self.resid, grad_eps=self.ge, grad_sig=self.gs)
        else:
            # per-shot partials into the group's buffer; no gradient sums here
            self.ge = self.gs = None
            self.ab.gshot = group.gshot[lo:hi]
            self.out = group.out
            self.pa = self.ab.problem(self.coef_a, resid=self.resid)
        self.pa.snaps, self.pa.nsnap, self.pa.snap_mod = self.ring.data_ptr(), R, R
        self.phases = [("fwd", 0, nt)]
        t = hist.tail()
        if t is not None:
            self.phases.append(("adj", t[0], t[1]))
        for s0, e in reversed(hist.segments()):
            self.phases.append(("rec", s0, e))
            self.phases.append(("adj", e - 1, max(s0, 1)))

    # --- operations (run on the job's stream, in the order returned) ----------
    def start_ops(self) -> list:
        """Clear the state and compute the coefficients: the first operations
        of a run (and of every replay of a captured run)."""
        G = self.G

        def start(stream):
            dev = G.device
            self.fb.reset_state()
            self.fb.traces.zero_()                     # dead receivers keep zero traces
            torch.clamp(self.se, min=0, max=1e6, out=self.se_f)      # compute.py:729
            _s.cell_coeffs(self.er, self.se_f, self.mr, G.dt32, G.dx32, False, dev,
                           out=self.coef_f)
            if self.gradient:
                self.ab.reset_state()
                if self.group is None:
                    self.ge.zero_()
                    self.gs.zero_()
                _s.cell_coeffs(self.er, self.se, self.mr, G.dt32, G.dx32, True, dev,
                               out=self.coef_a)          # Ucgeta
        return [start]

    def reset(self):
        """Rewind the phase machine to the first block (host side only)."""
        self.ip, self.n, self.parity = 0, self.phases[0][1], 0
        self.result, self.done_event = None, None

    @property
    def done(self) -> bool:
        return self.ip >= len(self.phases)

    def _enter_phase(self) -> list:
        if self.done:
            return []
        kind, a, b = self.phases[self.ip]
        self.parity = 0
        self.n = a
        if kind == "rec":
            def rec(stream, a=a):
                self.states.restore(a, self.fb)
                if a >= 1:
                    self.fb.seed_slot(self.pf_rec, (a - 1) % self.hist.R, self.ring,
                                      self.states.ez(a))
            return [rec]
        if kind == "adj":
            return [lambda stream: self.ab.gshot.zero_()]
        return []

    def _fwd_problem(self, kind):
        if not self.gradient:
            return self.pf_plain
        if kind == "rec":
            return self.pf_rec
        return self.pf if self.n >= self.first_stored else self.pf_plain

    def next_block(self):
        """(adjoint?, TBItem) for this job's next block, or None when done."""
        if self.done:
            return None
        kind, a, b = self.phases[self.ip]
        item = _native.TBItem()
        if kind in ("fwd", "rec"):
            stop = b
            if kind == "fwd":
                ahead = [m for m in self.marks if m > self.n]
                if ahead:
                    stop = min(b, min(ahead))       # a state to save, or snapshots begin
            ns = min(self.K, stop - self.n)
            self._prob = self._fwd_problem(kind)
            adj = False
        else:
            ns = min(self.K, self.n - b + 1)
            self._prob = self.pa
            adj = True
        item.prob = ctypes.pointer(self._prob)
        item.n0, item.nsteps, item.parity = self.n, ns, self.parity
        self._ns = ns
        return adj, item

    def after_block(self) -> list:
        """Advance the phase machine past the block just launched; returns the
        operations that follow it on the job's stream."""
        kind, a, b = self.phases[self.ip]
        self.parity ^= 1
        ops = []
        if kind in ("fwd", "rec"):
            self.n += self._ns
            boundary = kind == "fwd" and self.n in self.marks and self.n < b
            if self.n >= b or boundary:
                prob, par = self._prob, self.parity
                ops.append(lambda st, prob=prob, par=par:
                           check(lib.fgpr_tb_settle(ctypes.byref(prob), par, st), "settle"))
                self.parity = 0
            if boundary:
                if self.n in self.states.steps:
                    n = self.n
                    ops.append(lambda st, n=n: self.states.save(n, self.fb))
                return ops
            if self.n < b:
                return ops
            if kind == "fwd" and self.gradient:
                # dL/dtraces of 0.5 ||d - obs||^2
                ops.append(lambda st: torch.sub(self.fb.traces, self.obs, out=self.resid))
        else:
            self.n -= self._ns
            if self.n >= b:
                return ops
            par = self.parity
            ops.append(lambda st, par=par:
                       check(lib.fgpr_tb_settle(ctypes.byref(self.pa), par, st), "settle"))
            if self.group is None:         # a shot group's reduction is its group's
                ops.append(lambda st: check(lib.fgpr_reduce_shots(ctypes.byref(self.pa), st),
                                            "reduce"))
        self.ip += 1
        if self.done:
            ops.extend(self._finish_ops())
        else:
            ops.extend(self._enter_phase())
        return ops

    def _finish_ops(self) -> list:
        def finish(st):
            if self.gradient and self.group is None:
                NI, NJ = self.fb.NI, self.fb.NJ
                for name, g in (("er", self.ge), ("se", self.gs)):
                    if name in self.out:
                        check(lib.fgpr_normalise(_vp(g), NI * NJ, 1.0, _vp(self.out[name]), st),
                              "normalise")
            if self.finite_check:
                # one device flag for "every trace and gradient is finite",
                # copied to pinned host memory behind the job's final event
                ok = torch.isfinite(self.fb.traces).all()
                for o in (self.out.values() if self.gradient else ()):
                    ok = ok & torch.isfinite(o).all()
                self._finite = torch.empty(1, dtype=torch.bool, pin_memory=True)
                self._finite.copy_(ok.reshape(1), non_blocking=True)
        return [finish]

    def cancel(self):
        """Stop scheduling this job (host side; already queued work still runs)."""
        self.cancelled = True
        self.ip = len(self.phases)

    def finite(self) -> bool:
        """The queued finiteness test's answer (True when none was queued);
        valid once the job's final event has passed."""
        return True if self._finite is None else bool(self._finite.item())

    def results(self) -> dict:
        out = dict(self.out) if self.gradient else {}
        out["traces"] = self.fb.traces
        return out

    def cell_updates(self) -> int:
        return (2 if self.gradient else 1) * self.fb.S * self.fb.NI * self.fb.NJ * self.nt


class ShotGroup:
    """Shared state of one model split into shot-group jobs: whole-model
    traces (S, nt, nrx), per-shot FP64 partial gradients (S, 2, NI, NJ), the
    FP64 sums and the normalised outputs, from ``keep``.  ``finish_ops``
    (after every group job is done) reduces the partials over all shots in
    shot order and normalises -- what one job of all the shots does at its
    end, so the results are bitwise the same."""

    def __init__(self, G, keep, want=("er", "se")):
        S, nt, nrx = G.S, G.iterations, int(G.receiver.shape[1])
        self.G, self.S = G, S
        self.traces = keep((S, nt, nrx), torch.float32)
        self.gshot = keep((S, 2, G.NI, G.NJ), torch.float64)
        self.ge = keep((G.NI, G.NJ), torch.float64)
        self.gs = keep((G.NI, G.NJ), torch.float64)
        self.out = {name: keep((G.NI, G.NJ, 1), torch.float32) for name in ("er", "se")
                    if name in want}
        self.template = None     # a copy of one group job's problem struct

    def adopt(self, job) -> None:
        """Take the problem struct of one of the group's jobs (its grid, CPML
        and coefficient pointers, which every group job shares) as the
        template of the reduction.  A copy, so the group holds no reference
        to its jobs."""
        if self.template is None:
            self.template = bytes(job.pa)

    @staticmethod
    def bytes(G, want=("er", "se")) -> int:
        """Bytes ``keep`` supplies for a group (with alignment padding)."""
        S, nt, nrx = G.S, G.iterations, int(G.receiver.shape[1])
        sizes = [S * nt * nrx * 4, S * 2 * G.NI * G.NJ * 8, G.NI * G.NJ * 8, G.NI * G.NJ * 8]
        sizes += [G.NI * G.NJ * 4] * len(want)
        return sum(_aligned(b) for b in sizes)

    def finish_ops(self) -> list:
        def finish(st):
            self.ge.zero_()
            self.gs.zero_()
            p = _native.Problem.from_buffer_copy(self.template)
            p.S = self.S
            p.gshot = self.gshot.data_ptr()
            p.grad_eps, p.grad_sig = self.ge.data_ptr(), self.gs.data_ptr()
            check(lib.fgpr_reduce_shots(ctypes.byref(p), st), "reduce")
            NI, NJ = self.G.NI, self.G.NJ
            for name, g in (("er", self.ge), ("se", self.gs)):
                if name in self.out:
                    check(lib.fgpr_normalise(_vp(g), NI * NJ, 1.0, _vp(self.out[name]), st),
                          "normalise")
        return [finish]

    def results(self) -> dict:
        out = dict(self.out)
        out["traces"] = self.traces
        return out


def _run_ops(ops, stream: torch.cuda.Stream):
    handle = ctypes.c_void_p(stream.cuda_stream)
    with torch.cuda.stream(stream):
        for op in ops:
            op(handle)


class Scheduler:
    """Advances a set of jobs together: per tick, one multi-problem launch per
    (stream, variant, direction) group holding one block of every active job
    of that group.

    Jobs are spread over side streams (by default as many as have a hardware
    queue of their own, at most three), each new job on the stream with the
    least active work, so that one stream's launch tail is filled by another
    stream's launch.  Every operation of a job runs on its stream, ordered
    after the work queued on the calling stream when the job was added; a
    job's ``done_event`` is recorded on its stream, and ``join`` (the end of
    ``run``) makes the calling stream wait for all of them.  A job's
    arithmetic does not depend on its stream or on the other jobs."""

    def __init__(self, device, streams: int | None = None):
        self.device = torch.device(device)
        self.jobs: list[Job] = []
        self.launches = 0
        cap = _s.max_shot_groups()
        n = cap if streams is None else max(1, min(int(streams), cap))
        self.streams = _s._side_streams(self.device, n)

    def _place(self, job: Job):
        load = [0] * len(self.streams)
        for j in self.active():
            load[j.stream_idx] += j.cell_updates()
        job.stream_idx = min(range(len(self.streams)), key=lambda i: load[i])

    def add(self, job: Job, start: bool = True):
        """Queue a job: its stream waits for the calling stream's work so far,
        then runs the job's start operations (``start``; ``capture`` records
        them in the graph instead)."""
        self._place(job)
        self.jobs.append(job)
        if start:
            st = self.streams[job.stream_idx]
            st.wait_stream(torch.cuda.current_stream(self.device))
            _run_ops(job.start_ops() + job._enter_phase(), st)

    def active(self):
        return [j for j in self.jobs if not j.done]

    def plan_tick(self) -> list:
        """The next tick on the host: [(stream index, variant, adjoint?,
        [TBItem], [operations after the launch])], advancing every active job's
        phase machine (nothing is launched)."""
        groups: dict = {}
        for j in self.active():
            blk = j.next_block()
            if blk is None:
                continue
            adj, item = blk
            groups.setdefault((j.stream_idx, j.tb, adj), []).append((j, item))
        out = []
        for (si, tb, adj), members in groups.items():
            ops = []
            for j, _ in members:
                ops.extend(j.after_block())
                if j.done:
                    ops.append(("done", j))
            out.append((si, tb, adj, [m[1] for m in members], ops))
        return out

    @staticmethod
    def _pack(items, adj, host: torch.Tensor, off: int) -> int:
        n = len(items)
        arr = (_native.TBItem * n)(*items)
        nbytes = lib.fgpr_tb_table_bytes(n)
        total = ctypes.c_int()
        check(lib.fgpr_tb_pack(arr, n, 1 if adj else 0, ctypes.c_void_p(host.data_ptr() + off),
                               nbytes, ctypes.byref(total)), "pack")
        return total.value

    def _launch(self, st, dev_ptr: int, n: int, total: int, tb, adj, ops):
        stream = ctypes.c_void_p(st.cuda_stream)
        check(lib.fgpr_tb_launch_table(ctypes.c_void_p(dev_ptr), n, total, tb[0], tb[1], tb[2],
                                       1 if adj else 0, stream), "launch")
        self.launches += 1
        with torch.cuda.stream(st):
            for op in ops:
                if isinstance(op, tuple):
                    if op[1].done_event is None:
                        op[1].done_event = torch.cuda.Event()
                    op[1].done_event.record(st)
                else:
                    op(stream)

    def tick(self) -> int:
        """Plan one tick and launch it; returns the number of blocks advanced."""
        plan_ = self.plan_tick()
        for si, tb, adj, items, ops in plan_:
            st = self.streams[si]
            n = len(items)
            host = torch.empty(lib.fgpr_tb_table_bytes(n), dtype=torch.uint8, pin_memory=True)
            total = self._pack(items, adj, host, 0)
            with torch.cuda.stream(st):
                dev_table = host.to(self.device, non_blocking=True)
            self._launch(st, dev_table.data_ptr(), n, total, tb, adj, ops)
            dev_table.record_stream(st)
        return sum(len(g[3]) for g in plan_)

    def tick_events(self) -> list:
        """One event per stream, recorded after the work queued so far."""
        evs = []
        for st in self.streams:
            ev = torch.cuda.Event()
            ev.record(st)
            evs.append(ev)
        return evs

    def mark_done(self, job: Job):
        """Record ``job``'s final event after the work queued on its stream
        (a cancelled job)."""
        job.done_event = torch.cuda.Event()
        job.done_event.record(self.streams[job.stream_idx])

    def join(self):
        """The calling stream waits for the work queued on every stream."""
        main = torch.cuda.current_stream(self.device)
        for st in self.streams:
            main.wait_stream(st)

    def run(self):
        while self.active():
            self.tick()
        self.join()

    def capture(self, waves=None, final_ops=None) -> "CapturedBatch":
        """Plan jobs to their end and capture the whole schedule -- start
        operations, every launch and every phase change, on the jobs'
        streams -- in one HIP graph.  ``waves``: lists of added (not yet
        started) jobs run one after another (all streams join between waves,
        so a wave may reuse the memory of the one before); default: all jobs
        in one wave.  ``final_ops`` run on the capturing stream after every
        wave has joined.  The launch tables of all ticks are packed on the
        host and uploaded once."""
        with _trace.span("fgpr uneven capture"):
            return self._capture(waves, final_ops)

    def _capture(self, waves, final_ops):
        waves = [list(self.jobs)] if waves is None else [list(w) for w in waves]
        everyone = self.jobs
        plans = []
        try:
            for wave in waves:
                self.jobs = wave
                for j in wave:
                    j.reset()
                ticks = []
                while self.active():
                    ticks.append(self.plan_tick())
                plans.append(ticks)
        finally:
            self.jobs = everyone
        rows, off = [], 0
        for ticks in plans:
            for t in ticks:
                for g in t:
                    rows.append(off)
                    off += _aligned(lib.fgpr_tb_table_bytes(len(g[3])))
        host = torch.empty(max(off, ALIGN), dtype=torch.uint8, pin_memory=True)
        totals, k = [], 0
        for ticks in plans:
            for t in ticks:
                for si, tb, adj, items, ops in t:
                    totals.append(self._pack(items, adj, host, rows[k]))
                    k += 1
        table = host.to(self.device)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        launches0 = self.launches
        with torch.cuda.graph(graph):
            cap = torch.cuda.current_stream(self.device)
            k = 0
            for wave, ticks in zip(waves, plans):
                for st in self.streams:
                    st.wait_stream(cap)
                for j in wave:
                    j.reset()
                    _run_ops(j.start_ops() + j._enter_phase(), self.streams[j.stream_idx])
                for t in ticks:
                    for si, tb, adj, items, ops in t:
                        ops = [op for op in ops if not isinstance(op, tuple)]
                        self._launch(self.streams[si], table.data_ptr() + rows[k], len(items),
                                     totals[k], tb, adj, ops)
                        k += 1
                for st in self.streams:
                    cap.wait_stream(st)
            if final_ops:
                _run_ops(list(final_ops), cap)
        for wave in waves:
            for j in wave:
                j.ip = len(j.phases)          # the host phase machines ran to the end
        return CapturedBatch(graph, table, host, [j for w in waves for j in w],
                             sum(len(t) for t in plans), self.launches - launches0)


class CapturedBatch:
    """A job set's whole schedule in one HIP graph (``Scheduler.capture``).
    ``replay`` runs it on the calling stream; each replay starts from the
    jobs' start operations, so it recomputes every result from the current
    models and observed traces (``Job.set_model`` / ``Job.set_observed``)."""

    def __init__(self, graph, table, host, jobs, ticks, launches):
        self.graph, self.table, self.host, self.jobs = graph, table, host, jobs
        self.ticks, self.launches = ticks, launches

    def replay(self):
        with _trace.span("fgpr uneven replay"):
            self.graph.replay()

    def results(self) -> list:
        return [j.results() for j in self.jobs]
