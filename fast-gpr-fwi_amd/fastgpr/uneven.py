"""Uneven batches: the gradients of several models -- different grids, shot
counts and step counts -- computed together on one device from one captured
HIP graph.

``UnevenBatch`` takes one problem per model (grid description ``G`` from
``fastgpr.initialization``, model tensors, observed traces, time-blocking
variant) and

1. groups them into waves that fit the device memory with whole snapshot
   histories (first fit, largest first).  A model whose history does not fit
   even alone runs in a wave of its own and keeps the largest snapshot ring
   that fits (``checkpoint.History``: only the steps before the ring are
   recomputed);
2. carves every wave's buffers from one device allocation, each wave from
   its start (a wave reuses the memory of the one before; results live in
   separate tensors);
3. plans the whole schedule on the host -- one multi-problem launch per
   (stream, direction) per tick for all jobs of a wave, the shots of every
   model spread over the side streams -- uploads all launch tables at once
   and captures the schedule in one HIP graph (``batching.Scheduler.capture``).

``run`` replays the graph: every gradient is recomputed from the current
model tensors and observed traces, with no host work per tick.  Each result
is bitwise the one ``compute`` gives for that model alone.

No counterpart in the reference (one model per call, compute.py:33-50).
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import batching as _b
from .checkpoint import History


@dataclass
class Problem:
    """One model's gradient request: ``G`` from ``initialization``, model
    tensors (NI, NJ, 1), observed traces (S, nt, nrx), variant ``tb`` (a
    multi-problem variant, or "auto": the autotuner's fastest multi-problem
    variant for this grid and shot count, one batch on one stream)."""
    G: object
    er: torch.Tensor
    se: torch.Tensor
    mr: torch.Tensor
    obs: torch.Tensor
    tb: tuple = (64, 1024, 6)

    def spec(self):
        G = self.G
        return (G.NI, G.NJ, G.S, G.iterations, int(G.receiver.shape[1]), G.pmlthick,
                tuple(self.tb))


def tune(p: Problem) -> tuple:
    """The autotuner's fastest multi-problem variant for a problem's grid and
    shots, timed as one batch on one stream (how a job runs in a wave)."""
    import dataclasses
    from . import autotune
    G = dataclasses.replace(p.G, options=dict(p.G.options, shot_streams=1))
    return tuple(autotune.choose(G, p.er, p.se, p.mr, candidates=_b.MULTI_VARIANTS))


def plan_waves(specs, budget: int):
    """Waves of job indices and their histories: first fit, largest job
    first, whole histories; a job that does not fit an empty wave runs alone
    with the largest ring the budget holds."""
    full = [_b.job_bytes(*sp[:6], sp[3], True, sp[6]) for sp in specs]
    order = sorted(range(len(specs)), key=lambda k: -full[k])
    waves, used, hist = [], [], [None] * len(specs)
    for k in order:
        sp = specs[k]
        if full[k] > budget:
            hist[k] = _b.job_history(*sp[:6], budget, sp[6])
            waves.append([k])
            used.append(budget)
            continue
        for w in range(len(waves)):
            if used[w] + full[k] <= budget and all(hist[i].full for i in waves[w]):
                waves[w].append(k)
                used[w] += full[k]
                break
        else:
            waves.append([k])
            used.append(full[k])
        hist[k] = History(sp[3], sp[3])
    return waves, hist


class UnevenBatch:
    """Gradients of several models from one captured graph (see module
    docstring).  ``budget``: device bytes for the jobs' working buffers
    (default: 90 % of the free memory, counting what torch caches unused,
    less 8 GiB)."""

    def __init__(self, device, problems, budget: int | None = None, streams: int | None = None):
        self.device = torch.device(device)
        self.problems = list(problems)
        for p in self.problems:
            if p.tb == "auto":
                p.tb = tune(p)
        if budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            free += torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            budget = int(0.9 * free) - (8 << 30)
        specs = [p.spec() for p in self.problems]
        self.waves, self.histories = plan_waves(specs, budget)
        need = [sum(_b.job_bytes(*specs[k][:6], self.histories[k], True, specs[k][6])
                    for k in w) for w in self.waves]
        keep_bytes = sum(4 * (sp[2] * sp[3] * sp[4] + 2 * sp[0] * sp[1]) + 4 * _b.ALIGN
                         for sp in specs)
        torch.cuda.empty_cache()
        self.keep_arena = _b.Arena(self.device, keep_bytes)
        keep = self.keep_arena.region(self.keep_arena.alloc(keep_bytes))
        self.arena = _b.Arena(self.device, max(need))
        off = self.arena.alloc(max(need))
        self.sched = _b.Scheduler(self.device, streams)
        self.jobs = [None] * len(self.problems)
        for w in self.waves:
            region = _b.Region(self.arena.buf, off, max(need))
            for k in w:
                p = self.problems[k]
                self.jobs[k] = _b.Job(p.G, p.er, p.se, p.mr, obs=p.obs, tb=p.tb,
                                      history=self.histories[k], alloc=region, keep=keep, tag=k)
            # each wave's jobs spread over the streams by their own work
            before = list(self.sched.jobs)
            self.sched.jobs = []
            for k in sorted(w, key=lambda k: -self.jobs[k].cell_updates()):
                self.sched.add(self.jobs[k], start=False)
            self.sched.jobs = before + self.sched.jobs
        self.captured = self.sched.capture([[self.jobs[k] for k in w] for w in self.waves])

    def run(self) -> list:
        """Replay the batch on the calling stream; returns every problem's
        results (normalised gradients "er", "se" and "traces"), valid until
        the next run."""
        self.captured.replay()
        return [j.results() for j in self.jobs]

    def cell_updates(self) -> int:
        return sum(j.cell_updates() for j in self.jobs)

    def recompute_steps(self) -> list:
        return [h.recompute_steps() for h in self.histories]

    def summary(self) -> dict:
        return {"waves": [[self.problems[k].G.NI for k in w] for w in self.waves],
                "rings": [h.R for h in self.histories],
                "recompute_steps": self.recompute_steps(),
                "ticks": self.captured.ticks, "launches": self.captured.launches,
                "arena_gb": round(self.arena.size / 1e9, 2)}
