"""Serving loop: gradient (and forward-modelling) requests from concurrent
inversions, continuously batched on one GPU.

* Requests arrive on a thread-safe queue (``submit``) with their own grid,
  shots, acquisition, current model and observed traces; ``cancel`` withdraws
  one, waiting or running.
* Device memory is one arena the server allocates at start
  (``batching.Arena``); every job's large buffers are carved from one block
  of it, so admission is exact: a request is admitted when a free block
  holds its footprint (``batching.job_bytes``), and nothing else competes
  for that memory.
* The footprint adapts to the memory free at admission.  By default
  (``histories="strict"``) a gradient keeps its whole snapshot history when
  that fits in the largest free block; one whose history would fit the
  arena but not now waits for memory while another gradient runs rather
  than recompute, and past ``max_hold_s`` holds back every younger request
  until it fits.  Only a gradient whose history exceeds the arena takes a
  snapshot ring: the largest that fits in the largest free block
  (``checkpoint.History``: only the steps before the ring are recomputed).
  ``histories="whole"`` takes a ring once ``max_hold_s`` has passed instead,
  and ``"share"`` never waits for a whole history; both size a ring to half
  of the largest free block while other requests run.  A request whose
  smallest footprint does not fit waits; one that waited longer than
  ``max_hold_s`` holds back every younger request until it is admitted, so
  a stream of small requests cannot starve a large one.
* Admitted requests become ``batching.Job``s; every tick one multi-problem
  launch per (stream, variant, direction) advances every active job by one
  block, so requests that arrive while others run join at the next tick and
  their work overlaps on the device.  Models, CPML state, snapshot rings and
  saved states stay resident until the request completes.
* A request that cannot be served -- invalid input, larger than the arena,
  a non-finite trace or gradient -- fails alone: its result is
  ``{"error": ...}`` and every other request goes on.
* A request is complete when its final event has passed on the device; the
  loop records its latency (arrival to completion) and the cell updates it
  did.  Results do not depend on the mix: each job owns its buffers and the
  kernels do the same arithmetic alone or batched.

No counterpart in the reference (one model per call, compute.py:33-50); its
only guards are the NaN prints of compute.py:888-889 and the adjoint loop's
stop at a NaN field (compute.py:1049-1054).
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import batching
from . import trace as _trace
from .checkpoint import History
from .solver import to_device
from .workloads import Workload


WAIT = "wait"      # _plan: admit later, with the whole history


@dataclass
class Request:
    rid: int
    wl: Workload                   # grid, acquisition, time stepping
    model: tuple                   # (er, se) numpy (NI, NJ): the model to differentiate at
    obs: np.ndarray | None         # observed traces (S, nt, nrx); None = forward only
    arrival: float = 0.0           # perf_counter() at submission
    done_at: float | None = None
    admitted_at: float | None = None   # perf_counter() when its job was built
    result: dict | None = None
    cell_updates: int = 0
    tb: tuple | None = None
    history: History | None = None
    recompute_steps: int = 0
    _sizes: tuple | None = None        # (variant, smallest ring, its bytes, whole-history bytes)
    _held_at: int = -1                 # arena version of the last admission that did not fit
    _waited: bool = False              # put off once to keep its whole history


@dataclass
class Stats:
    completed: int = 0
    cell_updates: int = 0
    recompute_cell_updates: int = 0
    latencies: list = field(default_factory=list)
    started: float = 0.0
    finished: float = 0.0
    launches: int = 0
    peak_bytes: int = 0
    max_active: int = 0
    ticks: int = 0
    tick_host_s: float = 0.0        # host time spent packing and launching ticks
    admit_host_s: float = 0.0       # host time spent admitting (planning, building jobs)
    holds: int = 0                  # admissions put off because no free block held the request
    refused: int = 0                # requests that failed (invalid, too large, non-finite)
    cancelled: int = 0
    partial_histories: int = 0      # gradients admitted with a snapshot ring shorter than nt
    history_waits: int = 0          # requests put off to keep a whole history ("whole", "strict")

    def summary(self) -> dict:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        wall = max(self.finished - self.started, 1e-9)
        return {"completed": self.completed,
                "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)),
                "sustained_gcell_updates_per_s": self.cell_updates / wall / 1e9,
                "recompute_share": self.recompute_cell_updates / max(
                    self.cell_updates + self.recompute_cell_updates, 1),
                "wall_s": wall, "launches": self.launches, "ticks": self.ticks,
                "host_ms_per_tick": 1e3 * self.tick_host_s / max(self.ticks, 1),
                "admit_host_s": self.admit_host_s,
                "holds": self.holds, "refused": self.refused, "cancelled": self.cancelled,
                "partial_histories": self.partial_histories,
                "history_waits": self.history_waits,
                "peak_admitted_bytes": self.peak_bytes, "max_concurrent_requests": self.max_active}


def _host_view(a: np.ndarray) -> torch.Tensor:
    """A tensor view of a host array that is only read (``to_device`` copies
    it into pinned memory): a read-only array (a client's shared buffer) is
    taken without torch's warning about non-writable arrays."""
    import warnings
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", message="The given NumPy array is not writable")
        return torch.as_tensor(a)


class Server:
    """Continuous-batching gradient server for one device.

    ``hbm_budget``: bytes of the arena (default 85 % of the free memory);
    ``variant_for(wl)``: the time-blocking variant of a request without one.
    ``histories``: "share" -- a gradient keeps its whole snapshot history
    when it fits in half of the largest free block (all of it when nothing
    else runs), otherwise the largest ring that does; "whole" -- it keeps
    its whole history when that fits in the largest free block, and one whose
    history could fit the arena but not now waits for memory while another
    gradient runs (for at most ``max_hold_s``) instead of recomputing with a
    ring; "strict" -- as "whole", but past ``max_hold_s`` it holds back every
    younger request until its whole history fits, and only a gradient whose
    history exceeds the arena takes a ring, sized to the largest free block.
    ``scheduler`` and ``arena`` may be given (tests)."""

    def __init__(self, device="cuda:0", hbm_budget: int | None = None, variant_for=None,
                 max_inflight_ticks: int = 64, max_hold_s: float = 2.0, scheduler=None,
                 arena=None, histories: str = "strict"):
        if histories not in ("share", "whole", "strict"):
            raise ValueError(f"histories must be 'share', 'whole' or 'strict', not {histories!r}")
        self.histories = histories
        self.device = torch.device(device)
        if arena is None:
            if hbm_budget is None:
                free, _ = torch.cuda.mem_get_info(self.device)
                hbm_budget = int(0.85 * free)
            arena = batching.Arena(self.device, hbm_budget)
        self.arena = arena
        self.budget = arena.size
        self.variant_for = variant_for or (lambda wl: (64, 1024, 6))
        self.pending: "queue.Queue[Request]" = queue.Queue()
        self.waiting: list = []
        self.active: list = []          # (request, job, arena offset, bytes)
        self.in_use = 0
        self.sched = scheduler if scheduler is not None else batching.Scheduler(self.device)
        self.stats = Stats()
        self.max_inflight = max_inflight_ticks
        self.max_hold_s = max_hold_s
        self._tick_events: list = []
        self._cancel: set = set()
        self._lock = threading.Lock()
        self._arena_version = 0         # advances whenever a block is released

    # --- client side -----------------------------------------------------------
    def submit(self, req: Request) -> Request:
        req.arrival = time.perf_counter()
        self.pending.put(req)
        return req

    def cancel(self, rid: int):
        """Withdraw request ``rid``: a waiting one is dropped, a running one
        stops at its next block (its memory is released once its queued work
        has finished).  Its result is {"error": "cancelled"}."""
        with self._lock:
            self._cancel.add(rid)

    # --- scheduler side ----------------------------------------------------------
    def _fail(self, req: Request, msg: str, cancelled=False):
        req.result = {"error": msg}
        req.done_at = time.perf_counter()
        if cancelled:
            self.stats.cancelled += 1
        else:
            self.stats.refused += 1

    def _plan(self, req: Request):
        """(variant, history or None, bytes, smallest bytes) of a request
        given the memory free now."""
        wl = req.wl
        tb = tuple(req.tb or self.variant_for(wl))
        if tb not in batching.MULTI_VARIANTS:
            raise ValueError(f"time-blocking variant {tb} has no multi-problem kernel")
        spec = (wl.NI, wl.NJ, wl.S, wl.nt, int(wl.rx.shape[1]), wl.P)
        if req.obs is None:
            nbytes = batching.job_bytes(*spec, wl.nt, False, tb)
            return tb, None, nbytes, nbytes
        if req._sizes is None or req._sizes[0] != tb:      # the request's own sizes, once
            low = _smallest_history(spec, tb)
            req._sizes = (tb, low, batching.job_bytes(*spec, low, True, tb),
                          batching.job_bytes(*spec, wl.nt, True, tb))
        _, low, smallest, full = req._sizes
        largest = self.arena.largest_free()
        if self.histories in ("whole", "strict"):
            if full <= largest:
                return tb, History(wl.nt, wl.nt), full, smallest
            waited = (self.histories == "whole"
                      and time.perf_counter() - req.arrival > self.max_hold_s)
            if (full <= self.arena.size and not waited
                    and any(r.obs is not None for r, _, _, _ in self.active)):
                return tb, WAIT, full, smallest
        # "strict": a history larger than the arena takes the whole largest
        # free block (the fewer steps recomputed); otherwise half when others run
        share = largest if (not self.active or self.histories == "strict") else largest // 2
        if full <= share:
            return tb, History(wl.nt, wl.nt), full, smallest
        try:
            hist = batching.job_history(*spec, max(share, smallest), tb)
        except MemoryError:
            hist = low
        return tb, hist, batching.job_bytes(*spec, hist, True, tb), smallest

    def _admit(self):
        while True:
            try:
                self.waiting.append(self.pending.get_nowait())
            except queue.Empty:
                break
        with self._lock:
            cancel, self._cancel = self._cancel, set()
        still = []
        now = time.perf_counter()
        blocked = False                 # a long-held request blocks the younger ones
        for req in self.waiting:
            if req.rid in cancel:
                self._fail(req, "cancelled", cancelled=True)
                continue
            if blocked:
                still.append(req)
                continue
            if req._held_at == self._arena_version:
                still.append(req)          # nothing was released since it did not fit
                if now - req.arrival > self.max_hold_s:
                    blocked = True
                continue
            try:
                tb, hist, nbytes, smallest = self._plan(req)
            except Exception as e:                       # invalid request: it fails alone
                self._fail(req, f"invalid request: {type(e).__name__}: {e}")
                continue
            if smallest > self.arena.size:
                self._fail(req, f"request needs at least {smallest} bytes, arena {self.arena.size}")
                continue
            if hist is WAIT:                 # its whole history will fit once memory frees
                if not req._waited:
                    req._waited = True
                    self.stats.history_waits += 1
                still.append(req)
                if now - req.arrival > self.max_hold_s:
                    blocked = True           # "strict": memory drains for it
                continue
            off = self.arena.alloc(nbytes)
            if off is None:
                self.stats.holds += 1
                req._held_at = self._arena_version
                still.append(req)
                if now - req.arrival > self.max_hold_s:
                    blocked = True
                continue
            try:
                with _trace.span(f"fgpr admit request {req.rid}"):
                    job = self._make_job(req, tb, hist, self.arena.region(off))
            except Exception as e:
                self.arena.release(off)
                self._arena_version += 1
                self._fail(req, f"could not build the job: {type(e).__name__}: {e}")
                continue
            req.tb, req.history = tb, hist
            req.admitted_at = time.perf_counter()
            if hist is not None and not hist.full:
                self.stats.partial_histories += 1
                req.recompute_steps = hist.recompute_steps()
            self.sched.add(job)
            self.active.append((req, job, off, nbytes))
            self.in_use += nbytes
            self.stats.peak_bytes = max(self.stats.peak_bytes, self.in_use)
        self.waiting = still
        for req, job, _, _ in self.active:
            if req.rid in cancel and not job.done:
                job.cancel()
                req.result = {"error": "cancelled"}
                self.sched.mark_done(job)            # its event follows the job's queued work
        self.stats.max_active = max(self.stats.max_active, len(self.active))

    def _make_job(self, req: Request, tb, hist, region):
        import fastgpr
        wl = req.wl
        dev = self.device
        # pinned, asynchronous copies: a pageable copy would first wait for
        # every queued tick and leave the device idle while this job is set up
        t = lambda a: to_device(torch.from_numpy(np.ascontiguousarray(a, np.float32))[..., None], dev)
        er = torch.clamp(t(req.model[0]), min=1)
        se = torch.clamp(t(req.model[1]), min=0)
        mr = t(wl.mr)
        G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                   None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                   er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                                   False, False, None, torch.tensor(wl.src), torch.tensor(wl.rx))
        obs = None if req.obs is None else to_device(_host_view(req.obs), dev)
        job = batching.Job(G, er, se, mr, obs=obs, tb=tb, history=hist, tag=req.rid, alloc=region,
                           keep=lambda shape, dt: torch.empty(shape, dtype=dt, device=dev))
        job.finite_check = True
        return job

    def _reap(self):
        keep = []
        for req, job, off, nbytes in self.active:
            if job.done and job.done_event is not None and job.done_event.query():
                self.arena.release(off)
                self._arena_version += 1
                self.in_use -= nbytes
                self.sched.jobs.remove(job)
                req.done_at = time.perf_counter()
                if job.cancelled:
                    self.stats.cancelled += 1
                    continue
                res = job.results()
                if not job.finite():
                    self._fail(req, "non-finite traces or gradient (the reference prints "
                                    "'contains NaN' here)")
                    continue
                req.result = res
                req.cell_updates = job.cell_updates()
                self.stats.completed += 1
                self.stats.cell_updates += req.cell_updates
                self.stats.recompute_cell_updates += (req.recompute_steps * req.wl.S * req.wl.NI
                                                      * req.wl.NJ)
                self.stats.latencies.append(req.done_at - req.arrival)
            else:
                keep.append((req, job, off, nbytes))
        self.active = keep

    def step(self) -> bool:
        """Admit, advance one tick, reap.  Returns False when idle."""
        if not self.stats.started:
            self.stats.started = time.perf_counter()
        t = time.perf_counter()
        self._admit()
        self.stats.admit_host_s += time.perf_counter() - t
        busy = bool(self.sched.active())
        if busy:
            before = self.sched.launches
            t = time.perf_counter()
            self.sched.tick()
            self.stats.tick_host_s += time.perf_counter() - t
            self.stats.ticks += 1
            self.stats.launches += self.sched.launches - before
            self._tick_events.append(self.sched.tick_events())
            if len(self._tick_events) > self.max_inflight:
                for ev in self._tick_events.pop(0):        # bound how far the host runs ahead
                    ev.synchronize()
        self._reap()
        return busy or bool(self.active) or bool(self.waiting) or not self.pending.empty()

    def run(self, stop_when_idle=True, until: float | None = None):
        while True:
            busy = self.step()
            now = time.perf_counter()
            if until is not None and now >= until and not self.active:
                break
            if not busy:
                if stop_when_idle:
                    break
                time.sleep(0.0005)
        self.sched.join()
        self.stats.finished = time.perf_counter()
        return self.stats.summary()


def _smallest_history(spec, tb) -> History:
    """The ring that minimises a gradient's snapshot footprint: about
    sqrt(nt * state / slot) slots."""
    NI, NJ, S, nt, nrx, P = spec
    slot, state = batching._slot_state(NI, NJ, S, P, tb)
    R = max(2, min(nt, int(round(math.sqrt(nt * state / max(slot, 1))))))
    best = min((History(nt, r) for r in range(max(2, R // 2), min(nt, 2 * R) + 1)),
               key=lambda h: h.nbytes(slot, state))
    return best


def poisson_trace(rate_hz: float, duration_s: float, seed: int = 4) -> np.ndarray:
    """Arrival offsets (seconds) of a Poisson process."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= duration_s:
            return np.array(out)
        out.append(t)


def serving_request(rid: int, rng, grids=((128, 128), (256, 128), (512, 256), (1024, 512)),
                    shots=(1, 32), steps=(1500, 5000), gradient=True, P=20,
                    pool: dict | None = None) -> Request:
    """One request of the serving mix: grid drawn uniformly from ``grids``,
    shots uniform in ``shots``, step count interpolated in ``steps`` by the
    grid diagonal (smallest grid -> steps[0], largest -> steps[1]), seeded
    layered + blob model.  The observed traces are seeded noise (the
    gradient's cost does not depend on them).

    ``pool`` (a dict the caller keeps) makes a long stream of requests cheap
    to generate: models are built once per (grid, seed mod 8) and the
    observed traces are a read-only slice of one noise block per (steps,
    receivers).  The noise is then not drawn from ``rng``, so a pooled
    stream's requests differ from the pool-less stream's after the first."""
    from .grid import cfl_dt
    from .workloads import _cells, _smooth, layered_blob
    NI, NJ = grids[int(rng.integers(0, len(grids)))]
    S = int(rng.integers(shots[0], shots[1] + 1))
    diags = [math.hypot(*g) for g in grids]
    span = max(diags) - min(diags)
    frac = 0.0 if span == 0 else (math.hypot(NI, NJ) - min(diags)) / span
    nt = int(round(steps[0] + (steps[1] - steps[0]) * frac))
    seed = int(rng.integers(0, 2**31))
    mkey = ("model", NI, NJ, P, seed % 8)
    if pool is not None and mkey in pool:
        er, se, er0, se0 = pool[mkey]
    else:
        er, se = layered_blob(NI, NJ, seed=seed, nlayers=6, nblobs=2, P=P)
        er0, se0 = _smooth(er, 5).astype(np.float32), _smooth(se, 5).astype(np.float32)
        er, se = er.astype(np.float32), se.astype(np.float32)
        if pool is not None:
            pool[mkey] = (er, se, er0, se0)
    dx = 0.025
    dt = 0.9 * float(cfl_dt(dx))
    si = np.round(np.linspace(P + 5, NI - P - 6, S)).astype(int)
    nrx = int(max(1, min(64, NI - 2 * P - 10)))
    ri = np.round(np.linspace(P + 5, NI - P - 6, nrx)).astype(int)
    sj, rj = P + 5, NJ - P - 6
    wl = Workload(f"req{rid}_{NI}x{NJ}_{S}shots", er, se, np.ones((NI, NJ), np.float32), er0, se0,
                  dx, dt, nt, P, 2e8,
                  _cells([[(int(i), sj)] for i in si]), _cells([[(int(i), rj) for i in ri]] * S))
    obs = None
    if gradient and pool is None:
        obs = rng.standard_normal((S, nt, nrx)).astype(np.float32) * 1e-3
    elif gradient:
        nkey = ("noise", nt, nrx, shots[1])
        if nkey not in pool:
            block = np.random.default_rng(nt * 1009 + nrx).standard_normal(
                (shots[1], nt, nrx), dtype=np.float32) * np.float32(1e-3)
            block.flags.writeable = False
            pool[nkey] = block
        obs = pool[nkey][:S]
    return Request(rid, wl, (wl.er0, wl.se0), obs)
