"""Uneven batches: many solves -- different grids, shot counts, step counts
and phases -- advanced together, one multi-problem time-blocked launch per
tick.

A ``Job`` is one gradient (forward sweep, residual, adjoint sweep, FP64
reduction and L2 normalisation) or one forward modelling of one model with
its shots.  The ``Scheduler`` asks every active job for its next block of up
to K steps and packs all blocks of the same kind (forward-like / adjoint) and
time-blocking variant into one launch of ``k_tb_multi``.  Jobs keep their own
buffers, so each result is bitwise the one the job would produce alone.
Long histories are checkpointed: the forward keeps the state every C steps,
and the adjoint replays each segment's snapshots into a ring of C + 1 slots
before walking it backwards (as ``fastgpr.checkpoint``).

The reference solves one model per call and one step per launch
(compute.py:574-712, 1048-1172); this module has no counterpart there.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _native
from . import solver as _s
from .checkpoint import States
from ._native import check, lib

MULTI_VARIANTS = [(32, 512, 2), (32, 512, 4), (32, 512, 6), (64, 1024, 2), (64, 1024, 4),
                  (64, 1024, 6), (64, 512, 2), (64, 512, 4), (64, 512, 6), (64, 512, 8)]


def job_bytes(NI, NJ, S, nt, nrx, P, C, gradient=True, tb=(64, 1024, 6)) -> int:
    """Device bytes one job holds (fields and ping-pong partners for the
    forward and adjoint sides, CPML memory, coefficients, snapshot ring in the
    time-blocked layout of variant ``tb``, saved states, traces / residual,
    FP64 gradients)."""
    pitch = _native.pitch(NJ)
    _, _, ti, tj = _native.tb_tiles(NI, NJ, *tb)
    slot = S * max(ti * tj * tb[0] * 64, NI * pitch)
    field = 3 * S * (NI + 1) * pitch * 4
    psi = max(_native.psi_floats(NI, NJ, P, S), 1) * 4
    sides = 2 if gradient else 1
    b = sides * 2 * (field + psi) + sides * 5 * (NI + 1) * pitch * 4
    b += 2 * S * nt * nrx * 4
    if gradient:
        nseg = (nt + C - 1) // C
        b += (C + 1) * slot * 4
        b += max(nseg - 1, 0) * (field + psi)
        b += 2 * S * NI * NJ * 8 + 2 * NI * NJ * 8 + 2 * NI * NJ * 4
    return int(b)


class Job:
    """One solve advanced block by block (see module docstring)."""

    def __init__(self, G, er, se, mr, obs=None, tb=(64, 1024, 6), checkpoint=None,
                 want=("er", "se"), tag=None):
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
        C = nt if checkpoint is None else max(1, min(int(checkpoint), nt))
        self.C = C
        self.nseg = (nt + C - 1) // C
        se_f = torch.clamp(se.detach(), min=0, max=1e6)          # compute.py:729
        self.coef_f = _s.cell_coeffs(er, se_f, mr, G.dt32, G.dx32, False, dev)
        self.fb = _s.Batch(G, store=False, tb=self.tb)
        self.scale = _s.emitting_sources(G, self.fb)
        S, NI, NJ, pitch = self.fb.S, self.fb.NI, self.fb.NJ, self.fb.pitch
        self.pf_plain = self.fb.problem(self.coef_f, srcscale=self.scale)   # no snapshots
        self.result = None
        self.done_event = None
        self.stream_idx = 0
        self.ip = 0
        self.n = 0
        self.parity = 0
        if not self.gradient:
            self.phases = [("fwd", 0, nt)]
            self._enter_phase()
            return
        self.obs = obs.detach().to(device=dev, dtype=torch.float32).contiguous()
        self.ring = torch.empty((C + 1,) + self.fb.snap_slot_shape(), dtype=torch.float32, device=dev)
        self.pf = self.fb.problem(self.coef_f, srcscale=self.scale)
        self.pf.snaps, self.pf.nsnap, self.pf.snap_mod = self.ring.data_ptr(), C + 1, C + 1
        self.rec_traces = torch.empty_like(self.fb.traces)
        self.pf_rec = self.fb.problem(self.coef_f, srcscale=self.scale)
        self.pf_rec.snaps, self.pf_rec.nsnap, self.pf_rec.snap_mod = self.ring.data_ptr(), C + 1, C + 1
        self.pf_rec.traces = self.rec_traces.data_ptr()
        self.states = States(self.fb, C) if self.nseg > 1 else None
        self.ab = _s.Batch(G, store=False, tb=self.tb)
        self.coef_a = _s.cell_coeffs(er, se, mr, G.dt32, G.dx32, True, dev)      # Ucgeta
        self.resid = torch.empty((S, nt, self.fb.nrx), dtype=torch.float32, device=dev)
        self.ge = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
        self.gs = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
        self.ab.gshot_buffer()
        self.pa = self.ab.problem(self.coef_a, resid=self.resid, grad_eps=self.ge, grad_sig=self.gs)
        self.pa.snaps, self.pa.nsnap, self.pa.snap_mod = self.ring.data_ptr(), C + 1, C + 1
        if self.nseg == 1:
            self.phases = [("fwd", 0, nt), ("adj", nt - 1, 1)]
        else:
            self.phases = [("fwd", 0, nt)]
            for k in reversed(range(self.nseg)):
                s0, e = k * C, min((k + 1) * C, nt)
                self.phases.append(("rec", s0, e))
                if e - 1 >= max(s0, 1):
                    self.phases.append(("adj", e - 1, max(s0, 1)))
        self._enter_phase()

    # --- phase machine -------------------------------------------------------
    @property
    def done(self) -> bool:
        return self.ip >= len(self.phases)

    def _enter_phase(self):
        if self.done:
            return
        kind, a, b = self.phases[self.ip]
        self.parity = 0
        if kind in ("fwd", "rec"):
            self.n = a
            if kind == "rec":
                k = a // self.C
                self.states.restore(k, self.fb)
                if k >= 1:
                    self.fb.seed_slot(self.pf_rec, (a - 1) % (self.C + 1), self.ring,
                                      self.states.fields[k - 1][0])
        else:
            self.n = a
            self.ab.gshot.zero_()

    def _fwd_problem(self, kind):
        if not self.gradient:
            return self.pf_plain
        if kind == "rec":
            return self.pf_rec
        return self.pf if self.nseg == 1 else self.pf_plain

    def next_block(self):
        """(adjoint?, TBItem) for this job's next block, or None when done."""
        if self.done:
            return None
        kind, a, b = self.phases[self.ip]
        item = _native.TBItem()
        if kind in ("fwd", "rec"):
            stop = b
            if kind == "fwd" and self.gradient and self.nseg > 1:
                stop = min(b, (self.n // self.C + 1) * self.C)      # end at the next save point
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

    def after_block(self, stream):
        kind, a, b = self.phases[self.ip]
        self.parity ^= 1
        if kind in ("fwd", "rec"):
            self.n += self._ns
            boundary = (kind == "fwd" and self.gradient and self.nseg > 1 and self.n % self.C == 0
                        and self.n < b)
            if self.n >= b or boundary:
                check(lib.fgpr_tb_settle(ctypes.byref(self._prob), self.parity, stream), "settle")
                self.parity = 0
            if boundary:
                self.states.save(self.n // self.C, self.fb)
                return
            if self.n < b:
                return
            if kind == "fwd":
                self._forward_done()
        else:
            self.n -= self._ns
            if self.n >= b:
                return
            check(lib.fgpr_tb_settle(ctypes.byref(self.pa), self.parity, stream), "settle")
            check(lib.fgpr_reduce_shots(ctypes.byref(self.pa), stream), "reduce")
        self.ip += 1
        if self.done:
            self._finish(stream)
        else:
            self._enter_phase()

    def _forward_done(self):
        if self.gradient:
            torch.sub(self.fb.traces, self.obs, out=self.resid)     # dL/dtraces of 0.5||d - obs||^2

    def _finish(self, stream):
        if self.gradient:
            out = {}
            NI, NJ = self.fb.NI, self.fb.NJ
            for name, g in (("er", self.ge), ("se", self.gs)):
                if name in self.want:
                    o = torch.empty((NI, NJ, 1), dtype=torch.float32, device=self.G.device)
                    check(lib.fgpr_normalise(ctypes.c_void_p(g.data_ptr()), NI * NJ, 1.0,
                                             ctypes.c_void_p(o.data_ptr()), stream), "normalise")
                    out[name] = o
            out["traces"] = self.fb.traces
            self.result = out
        else:
            self.result = {"traces": self.fb.traces}
        self.done_event = torch.cuda.Event()
        self.done_event.record()

    def cell_updates(self) -> int:
        return (2 if self.gradient else 1) * self.fb.S * self.fb.NI * self.fb.NJ * self.nt


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

    def add(self, job: Job):
        load = [0] * len(self.streams)
        for j in self.active():
            load[j.stream_idx] += j.cell_updates()
        job.stream_idx = min(range(len(self.streams)), key=lambda i: load[i])
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.device))
        self.streams[job.stream_idx].wait_event(ready)
        self.jobs.append(job)

    def active(self):
        return [j for j in self.jobs if not j.done]

    def tick(self) -> int:
        """One launch per group; returns the number of blocks advanced."""
        groups: dict = {}
        for j in self.active():
            blk = j.next_block()
            if blk is None:
                continue
            adj, item = blk
            groups.setdefault((j.stream_idx, j.tb, adj), []).append((j, item))
        for (si, tb, adj), members in groups.items():
            st = self.streams[si]
            with torch.cuda.stream(st):
                stream = ctypes.c_void_p(st.cuda_stream)
                n = len(members)
                items = (_native.TBItem * n)(*[m[1] for m in members])
                nbytes = lib.fgpr_tb_table_bytes(n)
                host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
                total = ctypes.c_int()
                check(lib.fgpr_tb_pack(items, n, 1 if adj else 0,
                                       ctypes.c_void_p(host.data_ptr()), nbytes,
                                       ctypes.byref(total)), "pack")
                dev_table = host.to(self.device, non_blocking=True)
                check(lib.fgpr_tb_launch_table(ctypes.c_void_p(dev_table.data_ptr()), n,
                                               total.value, tb[0], tb[1], tb[2],
                                               1 if adj else 0, stream), "launch")
                self.launches += 1
                for j, _ in members:
                    j.after_block(stream)
        return sum(len(m) for m in groups.values())

    def join(self):
        """The calling stream waits for the work queued on every stream."""
        main = torch.cuda.current_stream(self.device)
        for st in self.streams:
            ev = torch.cuda.Event()
            ev.record(st)
            main.wait_event(ev)

    def run(self):
        while self.active():
            self.tick()
        self.join()
