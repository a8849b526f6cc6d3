"""Serving loop: gradient (and forward-modelling) requests from concurrent
inversions, continuously batched on one GPU.

* Requests arrive on a thread-safe queue (``submit``) with their own grid,
  shots, acquisition, current model and observed traces.
* Admission is by device bytes: a request is admitted when its estimated
  footprint (``batching.job_bytes``) fits in what the budget has left;
  histories that would take more than a quarter of the budget are
  checkpointed (interval ~ sqrt(3 nt)).
* Admitted requests become ``batching.Job``s; every tick one multi-problem
  launch per (variant, direction) advances every active job by one block, so
  requests that arrive while others run join at the next tick and their work
  overlaps on the device.  Models, CPML state, snapshot rings and
  checkpoints stay resident until the request completes.
* A request is complete when its final event has passed on the device; the
  loop records its latency (arrival to completion) and the cell updates it
  did.  Results do not depend on the mix: each job owns its buffers and the
  kernels do the same arithmetic alone or batched.

No counterpart in the reference (one model per call, compute.py:33-50).
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
from .solver import to_device
from .workloads import Workload


@dataclass
class Request:
    rid: int
    wl: Workload                   # grid, acquisition, time stepping
    model: tuple                   # (er, se) numpy (NI, NJ): the model to differentiate at
    obs: np.ndarray | None         # observed traces (S, nt, nrx); None = forward only
    arrival: float = 0.0           # perf_counter() at submission
    done_at: float | None = None
    result: dict | None = None
    cell_updates: int = 0
    tb: tuple | None = None
    checkpoint: int | None = None


@dataclass
class Stats:
    completed: int = 0
    cell_updates: int = 0
    latencies: list = field(default_factory=list)
    started: float = 0.0
    finished: float = 0.0
    launches: int = 0
    peak_bytes: int = 0
    max_active: int = 0
    ticks: int = 0
    tick_host_s: float = 0.0        # host time spent packing and launching ticks
    cache_releases: int = 0         # admissions that fitted only after releasing torch's cache
    oom_retries: int = 0            # admissions put back because the device ran out of memory
    refused: int = 0                # requests refused (larger than the budget, or no memory)

    def summary(self) -> dict:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        wall = max(self.finished - self.started, 1e-9)
        return {"completed": self.completed,
                "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)),
                "sustained_gcell_updates_per_s": self.cell_updates / wall / 1e9,
                "wall_s": wall, "launches": self.launches, "ticks": self.ticks,
                "host_ms_per_tick": 1e3 * self.tick_host_s / max(self.ticks, 1),
                "cache_releases": self.cache_releases, "oom_retries": self.oom_retries,
                "refused": self.refused,
                "peak_admitted_bytes": self.peak_bytes, "max_concurrent_requests": self.max_active}


class Server:
    """Continuous-batching gradient server for one device."""

    def __init__(self, device="cuda:0", hbm_budget: int | None = None, variant_for=None,
                 max_inflight_ticks: int = 64):
        self.device = torch.device(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.7 * free)
        self.budget = hbm_budget
        self.variant_for = variant_for or (lambda wl: (64, 1024, 6))
        self.pending: "queue.Queue[Request]" = queue.Queue()
        self.waiting: list = []
        self.active: list = []          # (request, job, bytes)
        self.in_use = 0
        self.sched = batching.Scheduler(self.device)
        self.stats = Stats()
        self.max_inflight = max_inflight_ticks
        self._tick_events: list = []
        self._lock = threading.Lock()
        self._hold = None               # (completed count, bytes) of the last request that did not fit

    # --- client side -----------------------------------------------------------
    def submit(self, req: Request) -> Request:
        req.arrival = time.perf_counter()
        self.pending.put(req)
        return req

    # --- scheduler side ----------------------------------------------------------
    def _plan(self, req: Request):
        wl = req.wl
        tb = req.tb or self.variant_for(wl)
        K = tb[2]
        full = batching.job_bytes(wl.NI, wl.NJ, wl.S, wl.nt, wl.rx.shape[1], wl.P, wl.nt,
                                  req.obs is not None, tb)
        C = wl.nt
        if req.obs is not None and full > self.budget // 4:
            c = int(math.ceil(math.sqrt(3.0 * wl.nt)))
            C = max(2 * K, ((c + 2 * K - 1) // (2 * K)) * 2 * K)
        nbytes = batching.job_bytes(wl.NI, wl.NJ, wl.S, wl.nt, wl.rx.shape[1], wl.P, C,
                                    req.obs is not None, tb)
        return tb, (None if C >= wl.nt else C), nbytes

    def _admit(self):
        while True:
            try:
                self.waiting.append(self.pending.get_nowait())
            except queue.Empty:
                break
        if self._hold is not None and self.stats.completed != self._hold[0]:
            self._hold = None              # a request has finished: memory was freed
        still = []
        for k, req in enumerate(self.waiting):
            tb, C, nbytes = self._plan(req)
            if self._hold is not None and nbytes >= self._hold[1]:
                still.append(req)          # no larger than what just failed to fit may pass
                continue
            if nbytes > self.budget:
                req.result = {"error": f"request needs {nbytes} bytes, budget {self.budget}"}
                req.done_at = time.perf_counter()
                self.stats.refused += 1
                continue
            if self.in_use + nbytes > self.budget:
                still.append(req)          # FIFO order kept; later small ones may pass
                continue
            job = None
            for attempt in range(2):
                try:
                    job = self._make_job(req, tb, C)
                    break
                except torch.OutOfMemoryError:
                    # The device holds less than the byte accounting says:
                    # memory cached by torch in pieces too small for this
                    # request.  Release the unused cache and try once more.
                    torch.cuda.empty_cache()
            if job is not None and attempt == 1:
                self.stats.cache_releases += 1
            if job is None:
                # Still no room: this request, and any not smaller, wait until
                # a running request has finished; smaller ones may still be
                # admitted.  With nothing running, nothing will free memory:
                # refuse the request.
                if not self.active:
                    req.result = {"error": "device out of memory with no request running"}
                    req.done_at = time.perf_counter()
                    self.stats.refused += 1
                    continue
                self.stats.oom_retries += 1
                self._hold = (self.stats.completed, nbytes)
                still.append(req)
                continue
            self.sched.add(job)
            self.active.append((req, job, nbytes))
            self.in_use += nbytes
            self.stats.peak_bytes = max(self.stats.peak_bytes, self.in_use)
        self.waiting = still
        self.stats.max_active = max(self.stats.max_active, len(self.active))

    def _make_job(self, req: Request, tb, C):
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
        obs = None if req.obs is None else to_device(torch.as_tensor(req.obs), dev)
        req.tb, req.checkpoint = tb, C
        return batching.Job(G, er, se, mr, obs=obs, tb=tb, checkpoint=C, tag=req.rid)

    def _reap(self):
        keep = []
        for req, job, nbytes in self.active:
            if job.done and job.done_event is not None and job.done_event.query():
                req.done_at = time.perf_counter()
                req.result = job.result
                req.cell_updates = job.cell_updates()
                self.stats.completed += 1
                self.stats.cell_updates += req.cell_updates
                self.stats.latencies.append(req.done_at - req.arrival)
                self.in_use -= nbytes
                self.sched.jobs.remove(job)
            else:
                keep.append((req, job, nbytes))
        self.active = keep

    def step(self) -> bool:
        """Admit, advance one tick, reap.  Returns False when idle."""
        if not self.stats.started:
            self.stats.started = time.perf_counter()
        self._admit()
        busy = bool(self.sched.active())
        if busy:
            before = self.sched.launches
            t = time.perf_counter()
            self.sched.tick()
            self.stats.tick_host_s += time.perf_counter() - t
            self.stats.ticks += 1
            self.stats.launches += self.sched.launches - before
            evs = []
            for st in self.sched.streams:
                ev = torch.cuda.Event()
                ev.record(st)
                evs.append(ev)
            self._tick_events.append(evs)
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
                    shots=(1, 32), steps=(1500, 5000), gradient=True, P=20) -> Request:
    """One request of the serving mix: grid drawn uniformly from ``grids``,
    shots uniform in ``shots``, step count interpolated in ``steps`` by the
    grid diagonal (smallest grid -> steps[0], largest -> steps[1]), seeded
    layered + blob model.  The observed traces are seeded noise (the
    gradient's cost does not depend on them)."""
    from .grid import cfl_dt
    from .workloads import _cells, _smooth, layered_blob
    NI, NJ = grids[int(rng.integers(0, len(grids)))]
    S = int(rng.integers(shots[0], shots[1] + 1))
    diags = [math.hypot(*g) for g in grids]
    span = max(diags) - min(diags)
    frac = 0.0 if span == 0 else (math.hypot(NI, NJ) - min(diags)) / span
    nt = int(round(steps[0] + (steps[1] - steps[0]) * frac))
    er, se = layered_blob(NI, NJ, seed=int(rng.integers(0, 2**31)), nlayers=6, nblobs=2, P=P)
    dx = 0.025
    dt = 0.9 * float(cfl_dt(dx))
    si = np.round(np.linspace(P + 5, NI - P - 6, S)).astype(int)
    nrx = int(max(1, min(64, NI - 2 * P - 10)))
    ri = np.round(np.linspace(P + 5, NI - P - 6, nrx)).astype(int)
    sj, rj = P + 5, NJ - P - 6
    wl = Workload(f"req{rid}_{NI}x{NJ}_{S}shots", er.astype(np.float32), se.astype(np.float32),
                  np.ones((NI, NJ), np.float32), _smooth(er, 5).astype(np.float32),
                  _smooth(se, 5).astype(np.float32), dx, dt, nt, P, 2e8,
                  _cells([[(int(i), sj)] for i in si]), _cells([[(int(i), rj) for i in ri]] * S))
    obs = None
    if gradient:
        obs = rng.standard_normal((S, nt, nrx)).astype(np.float32) * 1e-3
    return Request(rid, wl, (wl.er0, wl.se0), obs)
