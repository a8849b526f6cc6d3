"""Autotuner for the time-blocked kernels.

For one grid (NI, NJ, P) and batch size S it times every compiled variant
(region rows x threads x steps per launch) on a short forward + adjoint run
on scratch buffers and keeps the fastest.  Results are cached per process
and, when FASTGPR_TUNE_CACHE names a file, in that JSON file.
"""
from __future__ import annotations

import dataclasses
import json
import os

import torch

from . import _native
from . import solver as _solver

CANDIDATES = ([(32, 256, k) for k in (1, 2, 3, 4, 6)]
              + [(64, 512, k) for k in (1, 2, 3, 4, 6, 8)]
              + [(64, 256, k) for k in (2, 4, 8)]
              + [(32, 512, k) for k in (2, 4, 6)]
              + [(64, 1024, k) for k in (2, 4, 6, 8)])

_cache: dict = {}


def _key(G) -> str:
    name = torch.cuda.get_device_name(G.device)
    groups = max(1, min(int(G.options.get("shot_streams", 3) or 1), G.S, _solver.max_shot_groups()))
    return f"{name}|{G.NI}x{G.NJ}|P{G.pmlthick}|S{G.S}|rx{G.nrx}|g{groups}"


def _load_file():
    path = os.environ.get("FASTGPR_TUNE_CACHE")
    if path and os.path.isfile(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                _cache.setdefault(k, tuple(v))


def _save_file():
    path = os.environ.get("FASTGPR_TUNE_CACHE")
    if path:
        with open(path, "w") as f:
            json.dump({k: list(v) for k, v in _cache.items()}, f, indent=1)


def time_variant(G, er, se, mr, tb, steps: int, reps: int = 2) -> float:
    """Seconds per time step (forward + adjoint) of one variant."""
    g = dataclasses.replace(G, iterations=steps, srcwaves=G.srcwaves[:, :steps].contiguous(),
                            time_block=tb, exec_mode="eager")
    best = float("inf")
    for _ in range(reps + 1):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        _, b = _solver.forward(g, er, se, mr, store=True)
        resid = torch.zeros((b.S, b.nt, b.nrx), dtype=torch.float32, device=g.device)
        _solver.adjoint(g, b, er, se, mr, resid)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) / 1e3)
    return best / (2 * steps)


def choose(G, er, se, mr, candidates=None, steps: int = 48) -> tuple:
    """Fastest time-blocked variant for this grid and batch (cached)."""
    _load_file()
    key = _key(G)
    if key in _cache:
        return _cache[key]
    results = {}
    for tb in candidates or CANDIDATES:
        try:
            _native.tb_tiles(G.NI, G.NJ, *tb)
            results[tb] = time_variant(G, er, se, mr, tb, steps)
        except (RuntimeError, ValueError):
            continue
    best = min(results, key=results.get)
    _cache[key] = best
    _save_file()
    return best
