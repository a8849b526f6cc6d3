"""Uneven batches: the gradients of several models -- different grids, shot
counts and step counts -- computed together on one device from one captured
HIP graph.

``UnevenBatch`` takes one problem per model (grid description ``G`` from
``fastgpr.initialization``, model tensors, observed traces, time-blocking
variant) and

1. splits every model whose whole snapshot history fits the device into
   shot-group jobs (up to three, consecutive shots), as ``compute`` splits a
   batch over its shot streams: the groups share the model's traces and
   per-shot partial gradients, reduced once over all shots in shot order at
   the end of the graph;
2. groups the jobs into waves that fit the device memory with whole snapshot
   histories (first fit, largest first), by default one model per wave (its
   shot groups already fill the device; packing several models into a wave
   is an option).  A model whose history does not fit even alone runs as one
   job in a wave of its own and keeps the largest snapshot ring that fits
   (``checkpoint.History``: only the steps before the ring are recomputed);
3. carves every wave's buffers from one device allocation, each wave from
   its start (a wave reuses the memory of the one before; results and shot
   groups' shared buffers live in a separate allocation);
4. plans the whole schedule on the host -- one multi-problem launch per
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
from .checkpoint import History, device_budget


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


def tune(p: Problem, groups: int = 1) -> tuple:
    """The autotuner's fastest multi-problem variant for a problem's grid and
    shots, timed as the model runs in its wave: ``groups`` shot groups on
    their own streams (1: one batch on one stream)."""
    import dataclasses
    from . import autotune
    G = dataclasses.replace(p.G, options=dict(p.G.options, shot_streams=groups))
    return tuple(autotune.choose(G, p.er, p.se, p.mr, candidates=_b.MULTI_VARIANTS))


def plan_waves(specs, budget: int, owners=None):
    """Waves of job indices and their histories: first fit, largest job
    first, whole histories; a job that does not fit an empty wave runs alone
    with the largest ring the budget holds.  ``owners`` (one key per job):
    a job only joins a wave whose jobs have its key (None: any wave)."""
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
            if owners is not None and any(owners[i] != owners[k] for i in waves[w]):
                continue
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
    (default: ``checkpoint.device_budget``, the rule ``compute`` keeps a
    whole history by, so a model keeps its whole history here when it
    would there).  ``shot_groups``: a model whose whole snapshot history fits
    the budget runs as up to this many jobs of consecutive shots (spread over
    the streams like ``compute``'s shot groups; one reduction over all shots
    at the end keeps its results bitwise those of one job).  ``pack``: jobs of
    different models share a wave when they fit together; by default every
    model's jobs form waves of their own, one model after another, each
    model's shot groups filling the device (on the four-model, 64-shot
    batch: 0.727 s against 0.741 s packed)."""

    def __init__(self, device, problems, budget: int | None = None, streams: int | None = None,
                 shot_groups: int = 3, pack: bool = False):
        self.device = torch.device(device)
        self.problems = list(problems)
        if budget is None:
            budget = device_budget(self.device)

        def groups_of(p, tb):
            # shot groups of a model: split when its whole history fits
            NI, NJ, S, nt, nrx, P = p.spec()[:6]
            n = max(1, min(int(shot_groups), S))
            return n if n > 1 and _b.job_bytes(NI, NJ, S, nt, nrx, P, nt, True, tb) <= budget else 1

        for p in self.problems:
            if p.tb == "auto":
                # tuned split when some candidate's whole history fits
                p.tb = tune(p, max(groups_of(p, tb) for tb in _b.MULTI_VARIANTS))
        # jobs: (problem index, shots (lo, hi)); models whose history fits are split
        self.parts = []
        for k, p in enumerate(self.problems):
            S = p.spec()[2]
            n = groups_of(p, p.tb)
            if n > 1:
                cuts = [round(i * S / n) for i in range(n + 1)]
                self.parts += [(k, (cuts[i], cuts[i + 1])) for i in range(n)]
            else:
                self.parts.append((k, (0, S)))
        specs = []
        for k, (lo, hi) in self.parts:
            sp = self.problems[k].spec()
            specs.append((sp[0], sp[1], hi - lo) + sp[3:])
        owners = None if pack else [k for k, _ in self.parts]
        self.waves, self.histories = plan_waves(specs, budget, owners)
        need = [sum(_b.job_bytes(*specs[j][:6], self.histories[j], True, specs[j][6])
                    for j in w) for w in self.waves]
        split = {k for k, (lo, hi) in self.parts if hi - lo < self.problems[k].spec()[2]}
        keep_bytes = 0
        for k, p in enumerate(self.problems):
            sp = p.spec()
            if k in split:
                keep_bytes += _b.ShotGroup.bytes(p.G)
            else:
                keep_bytes += 4 * (sp[2] * sp[3] * sp[4] + 2 * sp[0] * sp[1]) + 4 * _b.ALIGN
        torch.cuda.empty_cache()
        self.keep_arena = _b.Arena(self.device, keep_bytes)
        keep = self.keep_arena.region(self.keep_arena.alloc(keep_bytes))
        self.groups = {k: _b.ShotGroup(self.problems[k].G, keep) for k in sorted(split)}
        self.arena = _b.Arena(self.device, max(need))
        off = self.arena.alloc(max(need))
        self.sched = _b.Scheduler(self.device, streams)
        self.jobs = [None] * len(self.parts)
        for w in self.waves:
            region = _b.Region(self.arena.buf, off, max(need))
            for j in w:
                k, shots = self.parts[j]
                p = self.problems[k]
                grp = self.groups.get(k)
                self.jobs[j] = _b.Job(p.G, p.er, p.se, p.mr, obs=p.obs, tb=p.tb,
                                      history=self.histories[j], alloc=region, keep=keep, tag=k,
                                      shots=shots if grp is not None else None, group=grp)
                if grp is not None:
                    grp.adopt(self.jobs[j])
            # each wave's jobs spread over the streams by their own work
            before = list(self.sched.jobs)
            self.sched.jobs = []
            for j in sorted(w, key=lambda j: -self.jobs[j].cell_updates()):
                self.sched.add(self.jobs[j], start=False)
            self.sched.jobs = before + self.sched.jobs
        final = [op for k in sorted(self.groups) for op in self.groups[k].finish_ops()]
        self.captured = self.sched.capture([[self.jobs[j] for j in w] for w in self.waves],
                                           final_ops=final)

    def run(self) -> list:
        """Replay the batch on the calling stream; returns every problem's
        results (normalised gradients "er", "se" and "traces"), valid until
        the next run."""
        self.captured.replay()
        return self.results()

    def results(self) -> list:
        out = []
        for k in range(len(self.problems)):
            if k in self.groups:
                out.append(self.groups[k].results())
            else:
                j = next(i for i, (kk, _) in enumerate(self.parts) if kk == k)
                out.append(self.jobs[j].results())
        return out

    def cell_updates(self) -> int:
        return sum(j.cell_updates() for j in self.jobs)

    def recompute_steps(self) -> list:
        return [h.recompute_steps() for h in self.histories]

    def summary(self) -> dict:
        return {"waves": [[f"{self.problems[self.parts[j][0]].G.NI}"
                           + (f"[{self.parts[j][1][0]}:{self.parts[j][1][1]}]"
                              if self.parts[j][0] in self.groups else "") for j in w]
                          for w in self.waves],
                "rings": [h.R for h in self.histories],
                "recompute_steps": self.recompute_steps(),
                "ticks": self.captured.ticks, "launches": self.captured.launches,
                "arena_gb": round(self.arena.size / 1e9, 2)}
