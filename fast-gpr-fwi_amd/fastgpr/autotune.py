"""Autotuner for the time-blocked kernels.

For one grid (NI, NJ, P) and batch size S it times every compiled variant
(region rows x threads x steps per launch): repeated forward and adjoint
launches of one block on scratch buffers, split into the shot groups the
sweep will use and replayed back to back from captured graphs as the sweeps
replay theirs, and keeps the fastest per step.
Results are cached per process and,
when FASTGPR_TUNE_CACHE names a file, in that JSON file (written atomically).
"""
from __future__ import annotations

import ctypes
import dataclasses
import json
import os

import torch

from . import _native
from . import solver as _solver
from ._native import lib

CANDIDATES = ([(32, 256, k) for k in (1, 2, 3, 4, 6)]
              + [(64, 512, k) for k in (1, 2, 3, 4, 6, 8)]
              + [(64, 256, k) for k in (2, 4, 8)]
              + [(32, 512, k) for k in (2, 4, 6)]
              + [(64, 1024, k) for k in (2, 4, 6, 8)])

_cache: dict = {}


class NoVariant(RuntimeError):
    """No compiled time-blocked variant runs this grid and acquisition (for
    example more receivers in one region than a region's lists hold)."""


def _key(G, candidates=None) -> str:
    name = torch.cuda.get_device_name(G.device)
    groups = max(1, min(int(G.options.get("shot_streams", 3) or 1), G.S, _solver.max_shot_groups()))
    cand = "all" if candidates is None else ";".join(",".join(map(str, c)) for c in sorted(candidates))
    return f"{name}|{G.NI}x{G.NJ}|P{G.pmlthick}|S{G.S}|rx{G.nrx}|g{groups}|{cand}"


def _load_file():
    path = os.environ.get("FASTGPR_TUNE_CACHE")
    if path and os.path.isfile(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                _cache.setdefault(k, tuple(v))


def _save_file():
    path = os.environ.get("FASTGPR_TUNE_CACHE")
    if path:
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump({k: list(v) for k, v in _cache.items()}, f, indent=1)
        os.replace(tmp, path)                 # readers never see a partial file


def time_variant(G, er, se, mr, tb, reps: int = 96, graph: bool = True) -> float:
    """Seconds per time step (forward + adjoint launch of one block, divided
    by its steps) of one variant, as the sweep runs it: the batch split into
    the same shot groups on their streams, each replaying its block ``reps``
    times back to back from a captured graph (fgpr_profile_kernel with
    FGPR_PROFILE_GRAPH), as the sweeps replay theirs; a first, untimed
    replay captures the graphs.  HIP events on the calling stream bracket
    the fan-out and the join.  Buffers are allocated before the timed
    region.  ``graph=False`` issues the launches one by one instead (for
    comparison: on small grids, where a block takes a few microseconds, it
    charges each launch a gap the sweeps do not have and picked a variant
    6 % slower at 200 x 100 and 800 x 400 with 4 shots).

    ``reps`` is sized for a stable ranking: on the 400 x 200, 13-shot
    cross-hole batch, 12 reps put the 6- and 8-step variants of the
    64-row, 1024-thread region within noise of each other (12.2-12.8 and
    12.4-13.0 us per step over six trials); 96 reps separate them
    (11.25-11.56 and 12.27-12.36), as full sweeps do (the 8-step variant's
    gradient is 9 % slower).  ``choose`` scales it down for larger blocks
    (``timing_reps``)."""
    K = int(tb[2])
    nt = 4 * K
    waves = torch.zeros((G.srcwaves.shape[0], nt), dtype=G.srcwaves.dtype)
    m = min(nt, G.srcwaves.shape[1])
    waves[:, :m] = G.srcwaves[:, :m]          # zero past the grid's own steps: nt columns always
    g = dataclasses.replace(G, iterations=nt, srcwaves=waves, time_block=tb)
    dev = g.device
    se_f = torch.clamp(se.detach(), min=0, max=1e6)
    coef = _solver.cell_coeffs(er, se_f, mr, g.dt32, g.dx32, False, dev)
    n = _solver.shot_groups(g)
    if n > 1:
        sg = _solver.ShotGroups(g, True, tb, n)
        sg.gshot_buffer()
        groups = sg.groups
    else:
        groups = [_solver.Batch(g, store=True, tb=tb)]
        groups[0].gshot_buffer()
    resid = torch.zeros((g.S, nt, int(g.receiver.shape[1])), dtype=torch.float32, device=dev)
    ge = torch.zeros((g.NI, g.NJ), dtype=torch.float64, device=dev)
    gs = torch.zeros_like(ge)
    probs = [b.problem(coef, srcscale=_solver.emitting_sources(g, b), resid=resid[b.lo:b.hi],
                       grad_eps=ge, grad_sig=gs) for b in groups]
    for b in groups:
        b.snaps.zero_()
    main = torch.cuda.current_stream(dev)
    flag = _native.PROFILE_GRAPH if graph else 0
    total = 0.0
    for kid, step in ((_native.KERNEL_TB_FWD, K), (_native.KERNEL_TB_ADJ, 2 * K)):
        def run(b, st, r):
            _native.check(lib.fgpr_profile_kernel(ctypes.byref(probs[groups.index(b)]),
                                                  kid | flag, step, r, st),
                          "profile")

        def launches(r):
            if n > 1:
                _solver._fan_out(dev, groups, lambda b, st: run(b, st, r))
            else:
                run(groups[0], ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream), r)
        launches(reps)                        # warm-up; captures each group's graph
        torch.cuda.synchronize(dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(main)
        launches(reps)
        e1.record(main)
        e1.synchronize()
        total += e0.elapsed_time(e1) / 1e3 / reps
    return total / K


def timing_reps(G) -> int:
    """Block launches timed per variant and direction: 96 for a block of
    1.04 M cell-shots (400 x 200, 13 shots), fewer in proportion for larger
    blocks, whose launches are longer than the launch-to-launch jitter that
    96 reps average out on the small one, and never fewer than 12.  The timed
    window stays near the small grid's (about 3 ms per variant and
    direction) instead of growing with the grid: 1600 x 800 with 32 shots
    takes 12 reps (0.8 s of tuning rather than 6 s)."""
    cells = max(1, (int(G.NI) + 1) * (int(G.NJ) + 1) * int(G.S))
    return max(12, min(96, round(96 * 401 * 201 * 13 / cells)))


def choose(G, er, se, mr, candidates=None) -> tuple:
    """Fastest time-blocked variant for this grid and batch (cached)."""
    _load_file()
    key = _key(G, candidates)
    if key in _cache:
        return _cache[key]
    results, errors = {}, []
    reps = timing_reps(G)
    for tb in candidates or CANDIDATES:
        try:
            _native.tb_tiles(G.NI, G.NJ, *tb)
            results[tb] = time_variant(G, er, se, mr, tb, reps=reps)
        except (RuntimeError, ValueError) as e:
            errors.append(f"{tb}: {e}")
            continue
    if not results:
        raise NoVariant(f"no time-blocked variant runs on a {G.NI} x {G.NJ} grid with "
                        f"{G.S} shots: " + "; ".join(errors[:4]))
    best = min(results, key=results.get)
    _cache[key] = best
    _save_file()
    return best
