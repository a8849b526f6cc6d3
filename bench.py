"""Gradient throughput benchmark (one MI355X).

A step is one full gradient -- forward sweep with snapshot storage, L2
residual, adjoint sweep with FP64 gradient accumulation and normalisation --
of the cross-hole batch: 400 x 200 grid, dx 0.025 m, 20-cell CPML, 3000
steps, Ricker 200 MHz, 13 shots in one launch, 80 receivers; seeded
synthetic layered model.  Inputs are resident on the device before timing.

Prints one JSON line (metric, value in Gcell-updates/s, roofline of the
dominant kernel measured with HIP events on the launch stream, and the CPU
oracle timed on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
for _p in (str(REPO), str(REPO / "fast-gpr-fwi_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "gradient Gcell-updates/s (fwd+adjoint) + p50/p99 request latency; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--exec", dest="exec_mode", default="graph", choices=["graph", "eager"])
    ap.add_argument("--chunk", type=int, default=None,
                    help="steps per captured graph (default: the library's choice)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--time-block", default="auto",
                    help="'auto' (autotuned), 'none' (per-step kernels) or rows,threads,steps")
    ap.add_argument("--plain", action="store_true",
                    help="also time the plain one-shot-per-launch eager path")
    ap.add_argument("--cpu-steps", type=int, default=3000,
                    help="time steps of the CPU oracle sample (forward + adjoint)")
    ap.add_argument("--profile-reps", type=int, default=200)
    ap.add_argument("--shot-streams", type=int, default=3,
                    help="shot groups swept on separate HIP streams (1 = one batch)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only launch a 256 MiB copy and the dominant kernels a few times "
                         "(the program profiled by the FETCH_SIZE / WRITE_SIZE counter runs)")
    ap.add_argument("--uneven", action="store_true",
                    help="time the uneven batch (4 models, 64 shots) through the multi-problem "
                         "scheduler against one model at a time; prints its own JSON line")
    ap.add_argument("--serve", type=float, default=0.0, metavar="SECONDS",
                    help="run the serving loop under Poisson arrivals (20 requests/s) for this "
                         "many seconds of arrivals; prints its own JSON line")
    ap.add_argument("--rate", type=float, default=20.0, help="serving arrival rate (requests/s)")
    ap.add_argument("--kernel-busy", nargs=3, metavar=("DB", "T0_NS", "T1_NS"),
                    help="read a rocprofv3 kernel-trace database and print how much of the "
                         "monotonic window [T0, T1] (the serving line's window_ns) had a kernel "
                         "running, with per-kernel totals; uses no GPU")
    ap.add_argument("--traffic", default=str(Path(__file__).resolve().parent / "profiles" /
                                             "traffic_c2_crosshole.json"),
                    help="measured HBM bytes per launch (from the counter runs), if present")
    return ap.parse_args(argv)


def step_bytes(S: int, NI: int, NJ: int, P: int) -> float:
    """Algorithmic HBM bytes of one time step of the batch: 40 B per
    cell-update (Ez/Hx/Hy read+write 24 B, three
    coefficients 12 B, snapshot write or read 4 B) plus 16 B per CPML cell
    and active direction, Npsi = 2P*NJ + 2P*NI (corners counted twice)."""
    return S * (40.0 * NI * NJ + 16.0 * (2 * P * NJ + 2 * P * NI))


def kernel_bytes(kernel: int, S: int, NI: int, NJ: int, P: int, nrx: int, k: int = 1) -> float:
    """Algorithmic bytes of one launch of each solver kernel.

    Time-blocked launches advance k steps: k x step_bytes.  The per-step
    kernels split one step in two half-steps; each is charged the fields it
    must read and write (4 B per shot-cell per field access), its coefficient
    planes (4 B per model cell, shared by the S shots), the snapshot (4 B),
    its CPML memory variables (8 B per slab cell) and, for the adjoint E
    kernel, the FP64 gradient pair (32 B per model cell)."""
    cells = S * NI * NJ
    model = NI * NJ
    psi_cells = S * (2 * P * NJ + 2 * P * NI)
    if kernel in (5, 6):
        return k * step_bytes(S, NI, NJ, P)
    if kernel in (0, 4):
        b = 20 * cells + 8 * model + 8 * psi_cells
        return b + (8 * S * nrx if kernel == 0 else 0)
    if kernel == 1:
        return 20 * cells + 12 * model + 8 * psi_cells
    if kernel == 3:
        return 24 * cells + 12 * model + 8 * psi_cells + 32 * model
    raise ValueError(kernel)


def kernel_busy(db: str, t0: int, t1: int, top: int = 10) -> dict:
    """Union of the kernel intervals of a rocprofv3 trace inside [t0, t1]
    (CLOCK_MONOTONIC ns, the clock of rocprofv3 timestamps and of
    time.monotonic_ns) and per-kernel launch counts and total durations."""
    import re
    import sqlite3
    con = sqlite3.connect(db)
    iv, tot, cnt = [], {}, {}
    for name, s, e in con.execute("select name, start, end from kernels"):
        if s < t0 or e > t1:
            continue
        iv.append((s, e))
        short = re.sub(r"\(.*", "", name.replace("(anonymous namespace)::", "")).replace("void ", "")
        tot[short] = tot.get(short, 0) + (e - s)
        cnt[short] = cnt.get(short, 0) + 1
    iv.sort()
    busy, cs, ce = 0, None, None
    for s, e in iv:
        if ce is None or s > ce:
            if ce is not None:
                busy += ce - cs
            cs, ce = s, e
        else:
            ce = max(ce, e)
    if ce is not None:
        busy += ce - cs
    kern = sorted(tot, key=tot.get, reverse=True)[:top]
    return {"window_s": (t1 - t0) / 1e9, "busy_s": busy / 1e9,
            "busy_fraction": busy / max(t1 - t0, 1), "launches": len(iv),
            "kernels": {k: {"launches": cnt[k], "total_s": round(tot[k] / 1e9, 4),
                            "avg_us": round(tot[k] / cnt[k] / 1e3, 2)} for k in kern}}


def main():
    args = parse()
    if args.kernel_busy:
        db, t0, t1 = args.kernel_busy
        print(json.dumps(kernel_busy(db, int(t0), int(t1))))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the shot batch runs on one MI355X; multi-GPU sharding is design-only).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import fastgpr
    from fastgpr import _native
    from fastgpr import solver as S_
    from fastgpr.workloads import c2_crosshole

    dev = torch.device("cuda:0")
    wl = c2_crosshole()
    if args.pmc_probe:
        pmc_probe(wl, dev, args)
        return
    if args.uneven:
        uneven_bench(dev)
        return
    if args.serve > 0:
        serve_bench(dev, args.serve, args.rate)
        return
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
    er_true, se_true, mr = t(wl.er), t(wl.se), t(wl.mr)
    er0, se0 = t(wl.er0), t(wl.se0)
    src = torch.tensor(wl.src)
    rx = torch.tensor(wl.rx)
    tb = (None if args.time_block == "none" else
          "auto" if args.time_block == "auto" else tuple(int(v) for v in args.time_block.split(",")))
    kw = dict(time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P, tv=False,
              customsource=src, customreceiver=rx, exec_mode=args.exec_mode,
              graph_chunk=args.chunk, time_block=tb, shot_streams=args.shot_streams)
    common = (dev, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'], ['ricker', 1, wl.freq], wl.S)

    with torch.no_grad():
        obs, _ = fastgpr.compute(*common, er_true, se_true, mr, 1, wl.rx.shape[1], **kw)
    obs = obs.detach().clone()
    er = er0.clone().requires_grad_(True)
    se = se0.clone().requires_grad_(True)

    def gradient_step():
        er.grad = None
        se.grad = None
        pred, _ = fastgpr.compute(*common, er, se, mr, 1, wl.rx.shape[1], **kw)
        loss = 0.5 * ((pred - obs) ** 2).sum()
        loss.backward()
        return loss

    for _ in range(args.warmup):
        gradient_step()
    torch.cuda.synchronize()
    fastgpr.solver.record_sweeps(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gradient_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sweeps = fastgpr.solver.record_sweeps(False)
    assert torch.isfinite(er.grad).all() and torch.isfinite(se.grad).all()

    cu = wl.cell_updates_per_gradient()
    ms = 1e3 * elapsed / args.steps
    value = cu / (elapsed / args.steps) / 1e9

    # ---- dominant-kernel roofline: launch durations with HIP events on the
    # launch stream, on scratch buffers (same shapes as the timed run)
    if tb == "auto":
        from fastgpr.autotune import _cache, _key
        Gk = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                    None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                    er.detach(), se.detach(), mr, 1, wl.rx.shape[1], wl.dt,
                                    'HORIPML', None, None, False, False, None, src, rx,
                                    time_block="auto", shot_streams=args.shot_streams)
        tb_used = _cache[_key(Gk)]
    else:
        tb_used = tb
    roofline, per_kernel = profile_kernels(wl, er.detach(), se.detach(), mr, src, rx, tb_used,
                                           args.profile_reps)
    roofline["traffic"] = measured_traffic(args.traffic, roofline["kernel"])
    # Sweeps of the timed region as run (shot groups on their streams): HIP
    # events on the calling stream before each fan-out and after its join.
    if tb_used and sweeps:
        blocks = -(-wl.nt // tb_used[2])
        per = {}
        for kind, a, b in sweeps:
            per.setdefault(kind, []).append(a.elapsed_time(b) * 1e3 / blocks)
        us = {k: sum(v) / len(v) for k, v in per.items()}
        kind = "adjoint" if ",true>" in roofline["kernel"] else "forward"
        blk = us.get(kind)
        if blk:
            gbs = roofline["bytes_per_launch"] / blk / 1e3
            roofline["effective"] = {
                "what": f"timed-region {kind} sweeps (the dominant kernel's direction): wall "
                        "time per whole-batch block, all shot groups",
                "groups": max(1, min(args.shot_streams, wl.S, fastgpr.solver.max_shot_groups())),
                "us_per_block": round(blk, 2), "achieved": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4),
                "forward_us_per_block": round(us.get("forward", 0.0), 2),
                "adjoint_us_per_block": round(us.get("adjoint", 0.0), 2)}

    # ---- CPU baseline: the numpy FP64 oracle on a bounded sample
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, f64 gradient accumulation",
        "data": "synthetic (seeded layered model, inputs resident in HBM)",
        "config": {"workload": wl.name, "grid": [wl.NI, wl.NJ], "dx_m": wl.dx, "cpml": wl.P,
                   "nt": wl.nt, "shots": wl.S, "receivers": int(wl.rx.shape[1]),
                   "cell_updates_per_gradient": cu, "exec": args.exec_mode,
                   "graph_chunk": (args.chunk if args.chunk else
                                   fastgpr.solver.TB_GRAPH_LAUNCHES * tb_used[2] if tb_used else 32),
                   "time_block": list(tb_used) if tb_used else None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_us": {k: round(v["us"], 2) for k, v in per_kernel.items()},

    }
    if args.plain:
        out["plain_path"] = plain_path(wl, er.detach(), se.detach(), mr, obs, src, rx, cu)
        out["plain_path"]["speedup"] = round(value / out["plain_path"]["value"], 2)
    print(json.dumps(out), flush=True)


KERNEL_NAMES = {
    0: "k_h_phase<true> (forward H half-step)",
    1: "k_e_phase<true,true,false> (forward E half-step)",
    3: "k_e_phase<false,false,true> (adjoint E half-step + gradient)",
    4: "k_h_phase<false> (adjoint H half-step)",
    5: "k_tb<{ri},{nt},{k},false> (time-blocked forward, {k} steps)",
    6: "k_tb<{ri},{nt},{k},true> (time-blocked adjoint + gradient, {k} steps)",
}


def profile_kernels(wl, er, se, mr, src, rx, tb, reps):
    """Average launch duration of each kernel of the path (HIP events on the
    stream the kernels run on), its algorithmic bytes and GB/s; the dominant
    kernel is the one with the largest share of a gradient's time.  The
    roofline is that of one launch over the whole batch on one stream (the
    shot groups' concurrent sweeps are timed in the timed region itself)."""
    import ctypes
    import torch
    import fastgpr
    from fastgpr import _native
    from fastgpr import solver as S_
    dev = er.device
    G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                               None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                               er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                               False, False, None, src, rx, time_block=tb)
    b = S_.Batch(G, store=True, tb=tb)
    coef = S_.cell_coeffs(er, se, mr, G.dt32, G.dx32, False, dev)
    scale = S_.source_scale(G).to(dev)
    ge = torch.zeros((wl.NI, wl.NJ), dtype=torch.float64, device=dev)
    gs = torch.zeros_like(ge)
    resid = torch.zeros((wl.S, b.nt, b.nrx), dtype=torch.float32, device=dev)
    if tb:
        b.gshot_buffer()
    p = b.problem(coef, srcscale=scale, resid=resid, grad_eps=ge, grad_sig=gs)
    b.snaps.zero_()
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    if tb:
        k = tb[2]
        kids = {5: (100, b.nt / k), 6: (100 + k, b.nt / k)}
    else:
        k = 1
        kids = {0: (10, b.nt), 1: (10, b.nt), 3: (10, b.nt - 1), 4: (10, b.nt - 1)}
    per = {}
    best = None
    for kid, (n, launches) in kids.items():
        _native.check(_native.lib.fgpr_profile_kernel(ctypes.byref(p), kid, n, 4, sp))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _native.check(_native.lib.fgpr_profile_kernel(ctypes.byref(p), kid, n, reps, sp))
        e1.record(stream)
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        byts = kernel_bytes(kid, wl.S, wl.NI, wl.NJ, wl.P, b.nrx, k)
        name = KERNEL_NAMES[kid].format(ri=tb[0] if tb else 0, nt=tb[1] if tb else 0, k=k)
        per[name] = dict(us=us, bytes=byts, gbs=byts / us / 1e3, share=us * launches)
        if best is None or per[name]["share"] > per[best]["share"]:
            best = name
    d = per[best]
    roofline = {"bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": best, "avg_launch_us": round(d["us"], 2),
                "bytes_per_launch": int(d["bytes"])}
    return roofline, per


def measured_traffic(path, kernel_name):
    """HBM bytes per launch of ``kernel_name`` from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), or None when no measurement of this kernel exists."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel_name.split(" ")[0])
    return None if k is None else k.get("bytes_per_launch")


def pmc_probe(wl, dev, args):
    """Program for the FETCH_SIZE / WRITE_SIZE counter runs: a 256 MiB
    device copy (the calibration: 256 MiB read + 256 MiB written) and the
    autotuned variant's forward and adjoint kernels, 5 launches each."""
    import numpy as np
    import torch
    import fastgpr
    from fastgpr.autotune import _cache, _key
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t = lambda x: torch.tensor(np.asarray(x, np.float32))[..., None].to(dev)
    er, se, mr = t(wl.er0), t(wl.se0), t(wl.mr)
    src, rx = torch.tensor(wl.src), torch.tensor(wl.rx)
    tb = (None if args.time_block == "none" else
          "auto" if args.time_block == "auto" else tuple(int(v) for v in args.time_block.split(",")))
    if tb == "auto":
        Gk = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                    None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                    er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                                    False, False, None, src, rx)
        from fastgpr.autotune import choose
        tb = choose(Gk, er, se, mr)
    profile_kernels(wl, er, se, mr, src, rx, tb, 5)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "done", "time_block": list(tb) if tb else None}), flush=True)


def plain_path(wl, er, se, mr, obs, src, rx, cu):
    """The plain GPU path: one shot per launch sequence, per-step kernels
    launched eagerly from the host loop, no graph, no batching, no time
    blocking.  One full gradient of the same workload, shot by shot."""
    import torch
    import fastgpr
    dev = er.device
    er = er.clone().requires_grad_(True)
    se = se.clone().requires_grad_(True)
    common = (dev, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'], ['ricker', 1, wl.freq], 1)

    def one_gradient():
        for s in range(wl.S):
            pred, _ = fastgpr.compute(*common, er, se, mr, 1, wl.rx.shape[1],
                                      time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P,
                                      tv=False, customsource=src[s:s + 1],
                                      customreceiver=rx[s:s + 1], exec_mode="eager")
            (0.5 * ((pred - obs[s:s + 1]) ** 2).sum()).backward()

    one_gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    one_gradient()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"value": round(cu / el / 1e9, 3), "unit": "Gcell-updates/s",
            "ms_per_gradient": round(el * 1e3, 1),
            "what": "per-step eager launches, one shot per launch sequence, no graph"}


def _job_for(wl, dev, tb, checkpoint):
    """A gradient job of workload ``wl`` (initial model, zero observed data),
    checkpointed when ``checkpoint`` is true."""
    import math
    import numpy as np
    import torch
    import fastgpr
    from fastgpr.batching import Job
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
    er, se, mr = torch.clamp(t(wl.er0), min=1), torch.clamp(t(wl.se0), min=0), t(wl.mr)
    G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                               None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                               er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                               False, False, None, torch.tensor(wl.src), torch.tensor(wl.rx))
    C = None
    if checkpoint:
        K = tb[2]
        c = int(math.ceil(math.sqrt(3.0 * wl.nt)))
        C = ((c + 2 * K - 1) // (2 * K)) * 2 * K
    obs = torch.zeros((wl.S, wl.nt, wl.rx.shape[1]), dtype=torch.float32, device=dev)
    return Job(G, er, se, mr, obs=obs, tb=tb, checkpoint=C), C


def uneven_bench(dev):
    """Uneven batch: 4 models of different sizes and step counts, 64 shots,
    advanced together by multi-problem launches (one per direction and
    variant per tick), each model with the time-blocking variant the
    autotuner picks for its grid among those the multi-problem kernel has,
    against the same gradients run one model at a time through ``compute``
    (autotuned time-blocked kernels, captured graphs, shot streams).
    Checkpoint recompute is not counted."""
    import numpy as np
    import torch
    import fastgpr
    from fastgpr import autotune
    from fastgpr.batching import MULTI_VARIANTS, Scheduler, job_bytes
    from fastgpr.workloads import c3_uneven
    wls = c3_uneven()
    # Per grid, the variant with the best throughput when the device is full
    # (in a multi-problem launch other jobs fill a job's launch tail): tuned
    # on the grid's acquisition repeated to at least 64 shots.
    tbs = []
    for w in wls:
        t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
        er, se, mr = torch.clamp(t(w.er0), min=1), torch.clamp(t(w.se0), min=0), t(w.mr)
        rep = -(-64 // w.S)
        G = fastgpr.initialization(dev, w.dx, w.NJ, w.NI, 1, w.time_window, ['hertzian', 'z'],
                                   None, None, None, None, w.P, ['ricker', 1, w.freq], w.S * rep,
                                   er, se, mr, 1, w.rx.shape[1], w.dt, 'HORIPML', None, None,
                                   False, False, None, torch.tensor(np.tile(w.src, (rep, 1, 1))),
                                   torch.tensor(np.tile(w.rx, (rep, 1, 1))), shot_streams=1)
        tbs.append(autotune.choose(G, er, se, mr, candidates=MULTI_VARIANTS))
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info(dev)
    # Checkpoint until every job fits in 80 % of the free memory, choosing
    # first the jobs whose recompute (one more forward sweep) costs the
    # fewest cell updates per byte the checkpointing frees.
    full = [job_bytes(w.NI, w.NJ, w.S, w.nt, w.rx.shape[1], w.P, w.nt, tb=tb)
            for w, tb in zip(wls, tbs)]
    small = [job_bytes(w.NI, w.NJ, w.S, w.nt, w.rx.shape[1], w.P, 144, tb=tb)
             for w, tb in zip(wls, tbs)]
    ck = [False] * len(wls)
    order = sorted(range(len(wls)),
                   key=lambda k: wls[k].S * wls[k].NI * wls[k].NJ * wls[k].nt
                   / max(full[k] - small[k], 1))
    for k in order:
        if sum(small[i] if ck[i] else full[i] for i in range(len(wls))) <= 0.8 * free:
            break
        ck[k] = True
    cu = sum(w.cell_updates_per_gradient() for w in wls)

    def batched():
        jobs = [_job_for(w, dev, tb, c) for w, tb, c in zip(wls, tbs, ck)]
        sch = Scheduler(dev)
        for j, _ in jobs:
            sch.add(j)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sch.run()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, sch.launches, [c for _, c in jobs]

    batched()                                    # warm-up (graphs, allocator)
    el_b, launches, ckpts = batched()

    def one_at_a_time():
        t0 = time.perf_counter()
        for w in wls:
            t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
            er = torch.clamp(t(w.er0), min=1).requires_grad_(True)
            se = torch.clamp(t(w.se0), min=0).requires_grad_(True)
            pred, _ = fastgpr.compute(dev, w.dx, w.NJ, w.NI, 1, ['hertzian', 'z'],
                                      ['ricker', 1, w.freq], w.S, er, se, t(w.mr), 1,
                                      w.rx.shape[1], time_windows=w.time_window, dt=w.dt,
                                      pmlthick=w.P, tv=False, customsource=torch.tensor(w.src),
                                      customreceiver=torch.tensor(w.rx), time_block="auto",
                                      checkpoint="auto")
            (0.5 * (pred ** 2).sum()).backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    one_at_a_time()
    el_s = one_at_a_time()
    print(json.dumps({
        "metric": "uneven batch gradient Gcell-updates/s (fwd+adjoint)",
        "value": round(cu / el_b / 1e9, 3), "unit": "Gcell-updates/s",
        "seconds": round(el_b, 3), "launches": launches,
        "one_model_at_a_time": {"value": round(cu / el_s / 1e9, 3), "seconds": round(el_s, 3)},
        "speedup": round(el_s / el_b, 3),
        "config": {"workload": "4 models, 64 shots: " + ", ".join(w.name + f"/{w.nt}" for w in wls),
                   "cell_updates": cu, "time_block": [list(tb) for tb in tbs],
                   "checkpoint_interval": ckpts}}), flush=True)


def serve_bench(dev, seconds, rate):
    """Serving loop under Poisson arrivals: requests (grid uniform over
    128x128, 256x128, 512x256, 1024x512; 1-32 shots; 1500-5000 steps by grid
    diagonal) are submitted from a thread at their arrival times, admitted by
    device bytes and continuously batched.  Reports p50/p99 latency, sustained
    throughput and the largest difference between a request served in the mix
    and the same request served alone (8 sampled).  Up to 200 requests are
    generated beforehand; longer runs generate each request in the submitting
    thread just before its arrival (same sequence) and keep the inputs and
    results of the 8 sampled requests only, so memory stays bounded."""
    import threading
    import numpy as np
    import torch
    from fastgpr.serving import Request, Server, poisson_trace, serving_request
    rng = np.random.default_rng(4)
    arrivals = poisson_trace(rate, seconds, seed=4)
    lazy = len(arrivals) > 200
    sampled = set(k * len(arrivals) // 8 for k in range(8))
    reqs = [] if lazy else [serving_request(i, rng) for i in range(len(arrivals))]
    offered_cu = [0]
    # warm-up: one small request of every grid (graph and allocator set-up)
    warm = Server(dev)
    for g in ((128, 128), (256, 128), (512, 256), (1024, 512)):
        warm.submit(serving_request(-1, np.random.default_rng(5), grids=(g,), shots=(2, 2),
                                    steps=(60, 60)))
    warm.run()
    torch.cuda.synchronize()

    del warm
    torch.cuda.empty_cache()
    srv = Server(dev)
    t0 = time.perf_counter()

    def submitter():
        for i, at in enumerate(arrivals):
            if lazy:
                reqs.append(serving_request(i, rng))
            r = reqs[i]
            offered_cu[0] += r.wl.cell_updates_per_gradient()
            d = t0 + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(r)

    th = threading.Thread(target=submitter, daemon=True)
    m0 = time.monotonic_ns()
    th.start()
    last = time.perf_counter()
    while True:
        busy = srv.step()
        if not busy:
            if not th.is_alive() and srv.pending.empty():
                break
            time.sleep(0.0005)
        now = time.perf_counter()
        if now - last > 20:
            print(f"serving: {srv.stats.completed}/{len(arrivals)} done, {len(srv.active)} active, "
                  f"{len(srv.waiting)} waiting", file=sys.stderr, flush=True)
            last = now
            if lazy:                     # completed requests outside the sample: drop their data
                for r in reqs[:]:
                    if r.done_at is not None and r.rid not in sampled and r.result is not None:
                        r.result, r.obs, r.model = None, None, None
    torch.cuda.synchronize()
    srv.stats.finished = time.perf_counter()
    m1 = time.monotonic_ns()
    summ = srv.stats.summary()
    # alone vs in-mix on a sample of the completed requests
    worst = 0.0
    done = [r for r in reqs if r.result and "er" in r.result]
    picks = [r for r in done if r.rid in sampled] if lazy else done[:: max(1, len(done) // 8)][:8]
    budget = srv.budget
    del srv
    torch.cuda.empty_cache()
    for r in picks:
        a = Server(dev, hbm_budget=budget)
        ra = a.submit(Request(r.rid, r.wl, r.model, r.obs, tb=r.tb))
        a.run()
        torch.cuda.synchronize()
        if "er" not in ra.result:
            raise RuntimeError(f"request {r.rid} served alone failed: {ra.result}")
        for k in ("er", "se", "traces"):
            x, y = r.result[k].double(), ra.result[k].double()
            worst = max(worst, float((x - y).abs().max() / max(y.abs().max(), 1e-30)))
    print(json.dumps({
        "metric": "serving: p50/p99 request latency, sustained gradient Gcell-updates/s",
        "p50_latency_s": round(summ["p50_latency_s"], 3),
        "p99_latency_s": round(summ["p99_latency_s"], 3),
        "sustained_gcell_updates_per_s": round(summ["sustained_gcell_updates_per_s"], 3),
        "completed": summ["completed"], "requests": len(reqs),
        "arrival_window_s": seconds, "rate_per_s": rate, "wall_s": round(summ["wall_s"], 2),
        "offered_gcell_updates_per_s": round(offered_cu[0] / seconds / 1e9, 3),
        "max_concurrent_requests": summ["max_concurrent_requests"],
        "peak_admitted_gb": round(summ["peak_admitted_bytes"] / 1e9, 2),
        "refused": summ["refused"], "cache_releases": summ["cache_releases"],
        "oom_retries": summ["oom_retries"],
        "launches": summ["launches"], "ticks": summ["ticks"],
        "host_ms_per_tick": round(summ["host_ms_per_tick"], 3),
        "window_ns": [m0, m1],
        "alone_vs_mix_max_rel_diff": worst}), flush=True)


def cpu_baseline(wl, nsteps: int):
    """The FP64 oracle's forward + adjoint over the first ``nsteps`` time
    steps of the same batch (all shots): the C/OpenMP restatement on every
    host thread OpenMP is given, or the numpy restatement on one core when
    the C library is not built."""
    import numpy as np
    from oracle import c_oracle as CO
    from oracle import fdtd_oracle as O
    sys.path.insert(0, str(REPO / "tests"))
    from helpers import oracle_problem
    pb = oracle_problem(wl)
    pb.nt = nsteps
    pb.wavelet = pb.wavelet[:, :nsteps]
    resid = np.random.default_rng(0).standard_normal((wl.S, nsteps, wl.rx.shape[1]))
    c, pml = O.setup(pb, np.float64)
    cadj, _ = O.setup(pb, np.float64, adjoint=True)
    scale = O.source_scale(pb.wavelet, pb.dx, np.float64)
    if CO.available():
        impl, cores = "C/OpenMP FP64 oracle", CO.threads()
        fwd, adj = CO.forward, CO.adjoint_reference
        t0 = time.perf_counter()
        tr, snaps = fwd(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, nsteps, True)
    else:
        impl, cores = "numpy FP64 oracle", 1
        adj = O.adjoint_reference
        t0 = time.perf_counter()
        tr, snaps = O.forward(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, nsteps, np.float64, True)
    adj(cadj, pml, pb.P, pb.dx, pb.rx, resid, snaps)
    el = time.perf_counter() - t0
    cu = 2 * wl.S * wl.NI * wl.NJ * nsteps
    return {"value": round(cu / el / 1e9, 5), "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"{impl}, {wl.name}, all {wl.S} shots, first {nsteps} of "
                      f"{wl.nt} steps forward + adjoint ({el:.1f} s, {cores} threads)"}


if __name__ == "__main__":
    main()
