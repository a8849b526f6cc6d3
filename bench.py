"""Gradient throughput benchmark (one MI355X).

A step is one pass of the hot path over one batch: the gradients of the
uneven batch -- four seeded models, 64 shots: 300x150 / 2000 steps,
500x250 / 3000, 800x400 / 4500, 1200x600 / 6000, dx 0.025 m, 20-cell CPML,
Ricker 200 MHz -- each a forward sweep with snapshot storage, L2 residual,
adjoint sweep with FP64 gradient accumulation and normalisation, computed
together by ``fastgpr.uneven.UnevenBatch`` from one captured HIP graph.
Inputs are resident on the device before timing.

Prints one JSON line: the metric in Gcell-updates/s, the roofline of the
dominant kernel measured with HIP events on the launch stream, the CPU
oracle timed on a bounded sample of the same workload, and beside them the
cross-hole batch (400x200, 13 shots, 3000 steps) through ``compute``, the
same four models one at a time through ``compute``, the plain
one-shot-per-launch path and a short run of the serving loop (p50 / p99).
"""
from __future__ import annotations

import argparse
import gc
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


TRAFFIC_FILE = REPO / "profiles" / "traffic.json"


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=["c3", "c2"],
                    help="c3: the uneven batch (default line); c2: the cross-hole batch alone")
    ap.add_argument("--quick", action="store_true",
                    help="the C3 line only (no one-at-a-time, C2, plain-path or serving keys)")
    ap.add_argument("--c3-variant", default="auto",
                    help="time-blocking variant of the uneven batch's jobs: 'auto' (the autotuner's "
                         "pick per model) or rows,threads,steps for all")
    ap.add_argument("--c3-shot-groups", type=int, default=3,
                    help="shot-group jobs per model of the uneven batch whose history fits "
                         "(1 = one job per model)")
    ap.add_argument("--exec", dest="exec_mode", default="graph", choices=["graph", "eager"])
    ap.add_argument("--chunk", type=int, default=None,
                    help="C2: steps per captured graph (default: the library's choice)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--time-block", default="auto",
                    help="C2: 'auto' (autotuned), 'none' (per-step kernels) or rows,threads,steps")
    ap.add_argument("--plain", action="store_true",
                    help="with --workload c2: also time the plain one-shot-per-launch eager path")
    ap.add_argument("--cpu-steps", type=int, default=3000,
                    help="C2: time steps of the CPU oracle sample (forward + adjoint)")
    ap.add_argument("--cpu-steps-c3", type=int, default=240,
                    help="C3: time steps of the CPU oracle sample on the largest model")
    ap.add_argument("--profile-reps", type=int, default=200)
    ap.add_argument("--shot-streams", type=int, default=3,
                    help="C2: shot groups swept on separate HIP streams (1 = one batch)")
    ap.add_argument("--serve-check", type=float, default=20.0, metavar="SECONDS",
                    help="seconds of Poisson arrivals of the serving run in the default line "
                         "(0 = none)")
    ap.add_argument("--serve-rate", type=float, default=4.0,
                    help="arrival rate of that serving run (requests/s)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only launch a 256 MiB copy and the dominant kernels a few times "
                         "(the program profiled by the FETCH_SIZE / WRITE_SIZE counter runs)")
    ap.add_argument("--serve", type=float, default=0.0, metavar="SECONDS",
                    help="only run the serving loop under Poisson arrivals for this many "
                         "seconds of arrivals; prints its own JSON line")
    ap.add_argument("--rate", type=float, default=20.0, help="--serve arrival rate (requests/s)")
    ap.add_argument("--serve-histories", default="strict", choices=["share", "whole", "strict"],
                    help="the server's snapshot-history admission rule")
    ap.add_argument("--serve-max-hold", type=float, default=2.0, metavar="SECONDS",
                    help="the server's max_hold_s: how long a request waits before it holds back "
                         "younger ones")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs of --pmc-probe (FETCH_SIZE, then "
                         "WRITE_SIZE; CSV output) into profiles/traffic.json; uses no GPU")
    ap.add_argument("--kernel-stats", metavar="DB",
                    help="per (kernel, grid size) dispatch counts and durations of a rocprofv3 "
                         "kernel-trace database; uses no GPU")
    ap.add_argument("--section", choices=["c3", "one", "c2", "plain", "serve", "notebook", "c4"],
                    default=None,
                    help="run one section of the default line and print its JSON (the default "
                         "run starts one process per section)")
    ap.add_argument("--ref-value", type=float, default=None,
                    help="--section plain: the headline value the plain path is compared with")
    ap.add_argument("--envelope", action="store_true",
                    help="autotuner envelope: on cross-hole grids 200x100..1600x800 with 1..32 "
                         "shots, time every variant graph-replayed (the tuner) and launched one "
                         "by one, and the full gradient with each of the fastest three pinned; "
                         "one JSON line per grid")
    ap.add_argument("--kernel-busy", nargs=3, metavar=("DB", "T0_NS", "T1_NS"),
                    help="read a rocprofv3 kernel-trace database and print how much of the "
                         "monotonic window [T0, T1] (the serving line's window_ns) had a kernel "
                         "running, with per-kernel totals; uses no GPU")
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


def kernel_stats(db: str, top: int = 12) -> list:
    """Dispatch count, total / average / min / max duration (us) per
    (kernel, grid size) of a rocprofv3 kernel-trace database, largest total
    first."""
    import re
    import sqlite3
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    gx = next((c for c in ("grid_size", "grid_size_x", "grid_x") if c in cols), None)
    q = f"select name, start, end{', ' + gx if gx else ''} from kernels"
    acc = {}
    for row in con.execute(q):
        name = re.sub(r"\(.*", "", row[0].replace("(anonymous namespace)::", "")).replace("void ", "")
        key = (name, int(row[3]) if gx else 0)
        acc.setdefault(key, []).append((row[2] - row[1]) / 1e3)
    out = []
    for (name, grid), d in sorted(acc.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append({"kernel": name, "grid": grid, "calls": len(d), "total_ms": round(sum(d) / 1e3, 3),
                    "avg_us": round(sum(d) / len(d), 2), "min_us": round(min(d), 2),
                    "max_us": round(max(d), 2)})
    return out


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
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary), indent=1))
        return
    if args.kernel_stats:
        print(json.dumps(kernel_stats(args.kernel_stats), indent=1))
        return
    if args.kernel_busy:
        db, t0, t1 = args.kernel_busy
        print(json.dumps(kernel_busy(db, int(t0), int(t1))))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the batch runs on one MI355X; multi-GPU sharding is design-only).",
              file=sys.stderr)
        sys.exit(2)

    if args.section is None and args.workload == "c3" and not args.quick and not args.pmc_probe \
            and args.serve <= 0 and not args.envelope:
        sys.exit(run_sections(args))

    import torch

    dev = torch.device("cuda:0")
    if args.pmc_probe:
        pmc_probe(dev, args)
        return
    if args.envelope:
        tuner_envelope(dev)
        return
    if args.serve > 0:
        print(json.dumps(serve_bench(dev, args.serve, args.rate, args.serve_histories,
                                     args.serve_max_hold)), flush=True)
        return
    if args.workload == "c2" and args.section is None:
        out = c2_line(dev, args)
        if args.plain:
            out["plain_path"] = plain_path(dev, out["value"])
        print(json.dumps(out), flush=True)
        return
    sec = args.section or "c3"
    if sec == "c3":
        out = c3_line(dev, args)
    elif sec == "one":
        out = one_at_a_time(dev)
    elif sec == "c2":
        c2 = c2_line(dev, args, extras=False)
        out = {k: c2[k] for k in ("value", "ms_per_step", "config", "roofline")}
    elif sec == "plain":
        out = plain_path(dev, args.ref_value)
    elif sec == "notebook":
        out = notebook_inversion(dev)
    elif sec == "c4":
        out = c4_memory_bound(dev)
    else:
        out = serve_bench(dev, args.serve_check, args.serve_rate, args.serve_histories,
                          args.serve_max_hold)
    print(json.dumps(out), flush=True)


# Sections of the default line, each run in a process of its own (in this
# order): the headline, then the extra keys.  A section that fails leaves an
# "error" in its key and the others still run.
SECTIONS = (("c3", None), ("one", "one_model_at_a_time"), ("c2", "c2_crosshole"),
            ("plain", "plain_path"), ("serve", "serving"), ("notebook", "notebook_inversion"),
            ("c4", "c4_memory_bound"))
SECTION_TIMEOUT_S = {"c3": 420, "one": 300, "c2": 240, "plain": 240, "serve": 300,
                     "notebook": 240, "c4": 300}


def run_sections(args) -> int:
    """The default line: every section in a child process of this script
    (this process does no GPU work, so each child has the whole device),
    merged into one JSON line.  Returns the exit status: that of the
    headline section."""
    import subprocess
    here = str(Path(__file__).resolve())
    out, status = None, 0
    for sec, key in SECTIONS:
        if sec == "serve" and args.serve_check <= 0:
            continue
        argv = [sys.executable, here, *sys.argv[1:], "--section", sec]
        if sec == "plain":
            if out is None or out.get("value") is None:
                continue
            argv += ["--ref-value", str(out["value"])]
        log(f"section {sec}")
        try:
            p = subprocess.run(argv, stdout=subprocess.PIPE, text=True,
                               timeout=SECTION_TIMEOUT_S[sec])
            lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
            res = json.loads(lines[-1]) if lines else None
            err = None if res is not None else f"section {sec} exited with status {p.returncode}"
            if res is not None and p.returncode != 0:
                # printed its result, then failed on the way out: keep the
                # figures and say so
                log(f"section {sec} printed its result and exited with status {p.returncode}")
                res["exit_status"] = p.returncode
        except subprocess.TimeoutExpired:
            res, err = None, f"section {sec} exceeded {SECTION_TIMEOUT_S[sec]} s"
        if err:
            log(err)
        if sec == "c3":
            if res is None:
                status = 1
                res = {"metric": METRIC, "value": None, "unit": "Gcell-updates/s", "n_gpus": 1,
                       "steps": args.steps, "warmup": args.warmup, "error": err}
            out = res
            continue
        out[key] = res if res is not None else {"error": err}
        if sec == "one" and res is not None and out.get("value"):
            res["batched_speedup"] = round(out["value"] / res["value"], 3)
    print(json.dumps(out), flush=True)
    return status


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _tensors(wl, dev, model):
    import numpy as np
    import torch
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
    return t(model[0]), t(model[1]), t(wl.mr)


def _compute(wl, dev, er, se, mr, **kw):
    import torch
    import fastgpr
    return fastgpr.compute(dev, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'], ['ricker', 1, wl.freq],
                           wl.S, er, se, mr, wl.src.shape[1], wl.rx.shape[1],
                           time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P, tv=False,
                           customsource=torch.tensor(wl.src), customreceiver=torch.tensor(wl.rx),
                           **kw)[0]


def c3_problems(dev, tbs):
    """The uneven batch's problems at the smoothed initial models, with the
    true models' traces as observed data."""
    import torch
    import fastgpr
    from fastgpr.uneven import Problem
    from fastgpr.workloads import c3_uneven
    wls = c3_uneven()
    probs = []
    for wl, tb in zip(wls, tbs):
        with torch.no_grad():
            obs = _compute(wl, dev, *_tensors(wl, dev, (wl.er, wl.se)),
                           time_block=(64, 1024, 6) if tb == "auto" else tb).detach()
        er, se, mr = _tensors(wl, dev, (wl.er0, wl.se0))
        er, se = torch.clamp(er, min=1), torch.clamp(se, min=0)
        G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                   None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                   er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                                   False, False, None, torch.tensor(wl.src), torch.tensor(wl.rx))
        probs.append(Problem(G, er, se, mr, obs, tb))
    return wls, probs


def c3_line(dev, args):
    """The headline: the uneven batch from one captured graph."""
    import torch
    from fastgpr.uneven import UnevenBatch
    tb = "auto" if args.c3_variant == "auto" else tuple(int(v) for v in args.c3_variant.split(","))
    log("C3: observed traces and problems")
    wls, probs = c3_problems(dev, [tb] * 4)
    t0 = time.perf_counter()
    ub = UnevenBatch(dev, probs, shot_groups=args.c3_shot_groups)
    build_s = time.perf_counter() - t0
    tbs = [p.tb for p in ub.problems]
    log(f"C3: built and captured in {build_s:.1f} s: {ub.summary()}")
    for _ in range(args.warmup):
        ub.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = ub.run()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for r in res:
        assert torch.isfinite(r["er"]).all() and torch.isfinite(r["se"]).all()
    cu = ub.cell_updates()
    value = cu / (elapsed / args.steps) / 1e9
    log(f"C3: {value:.1f} Gcell-updates/s")
    roofline, kernels = multi_roofline(ub, args.profile_reps)
    summ = ub.summary()
    del ub, res, probs
    gc.collect()                       # jobs hold their operations' closures (cycles)
    torch.cuda.empty_cache()
    cpu = None if args.no_cpu_baseline else cpu_baseline(wls[3], args.cpu_steps_c3)
    return {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, f64 gradient accumulation",
        "data": "synthetic (seeded layered + blob models, inputs resident in HBM)",
        "config": {"workload": "c3_uneven: 4 models, 64 shots: " +
                               ", ".join(f"{w.NI}x{w.NJ}/{w.nt} steps/{w.S} shots" for w in wls),
                   "dx_m": wls[0].dx, "cpml": wls[0].P,
                   "cell_updates_per_step": cu, "time_block": [list(t) for t in tbs],
                   "exec": "one captured HIP graph per step (all ticks, all waves)",
                   "shot_groups": args.c3_shot_groups,
                   "waves": summ["waves"], "snapshot_rings": summ["rings"],
                   "recompute_steps": summ["recompute_steps"],
                   "ticks": summ["ticks"], "launches": summ["launches"],
                   "build_and_capture_s": round(build_s, 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_us": kernels,
    }


def multi_roofline(ub, reps):
    """Average launch duration of the multi-problem kernels on the largest
    model's block (one launch holding the items of all its shot-group jobs:
    every shot once, the workgroups of one tick of its wave), timed with HIP
    events on the launch stream, against the algorithmic bytes of a block and, when
    the counter runs measured this exact kernel and workload, the HBM bytes
    it moved."""
    import ctypes
    import torch
    from fastgpr import _native
    from fastgpr._native import check, lib
    # the largest model; its shot-group jobs together are one block of all its shots
    work = {}
    for j, (kk, _) in zip(ub.jobs, ub.parts):
        work[kk] = work.get(kk, 0) + j.cell_updates()
    k = max(work, key=work.get)
    jobs = [j for j, (kk, _) in zip(ub.jobs, ub.parts) if kk == k]
    tb = jobs[0].tb
    G = ub.problems[k].G
    name = f"c3_model{k}_{G.NI}x{G.NJ}_{G.S}shots"
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    per = {}
    K = tb[2]
    n = len(jobs)
    for adj, step0 in ((False, G.iterations - K), (True, G.iterations - 1)):
        # the forward as the sweep runs it, with its snapshot stores
        probs = [(j.pa if adj else (j.pf if j.gradient else j.pf_plain)) for j in jobs]
        items = (_native.TBItem * n)()
        for it, prob in zip(items, probs):
            it.prob = ctypes.pointer(prob)
            it.n0, it.nsteps, it.parity = step0, K, 0
        nbytes = lib.fgpr_tb_table_bytes(n)
        host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        total = ctypes.c_int()
        check(lib.fgpr_tb_pack(items, n, 1 if adj else 0,
                               ctypes.c_void_p(host.data_ptr()), nbytes, ctypes.byref(total)))
        table = host.to(G.device)

        def launch(r):
            for _ in range(r):
                check(lib.fgpr_tb_launch_table(ctypes.c_void_p(table.data_ptr()), n, total.value,
                                               tb[0], tb[1], tb[2], 1 if adj else 0, sp))
        launch(4)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        launch(reps)
        e1.record(st)
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        kname = f"k_tb_multi<{tb[0]},{tb[1]},{tb[2]},{'true' if adj else 'false'}>"
        per[kname] = {"us": us, "bytes": K * step_bytes(G.S, G.NI, G.NJ, G.pmlthick),
                      "share": us * (G.iterations / K)}
    best = max(per, key=lambda n: per[n]["share"])
    d = per[best]
    gbs = d["bytes"] / d["us"] / 1e3
    roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "basis": "achieved = algorithmic bytes of the block (40 B per cell-update + 16 B per "
                     "CPML cell and direction, x steps) / launch time; time blocking keeps the "
                     "fields on chip, so this is an equivalent rate -- physical is what HBM moved",
            "kernel": f"{best} on {name} (the largest model: one launch over all its shots, "
                      f"{n} shot-group item(s))",
            "workload_key": f"{name}|{best}",
            "avg_launch_us": round(d["us"], 2), "bytes_per_launch": int(d["bytes"]),
            "launches_per_step": int(2 * G.iterations / K)}
    t = measured_traffic(TRAFFIC_FILE, roof["workload_key"])
    if t is not None:
        roof["traffic"] = t
        roof["physical"] = {"achieved": round(t / d["us"] / 1e3, 1),
                            "frac": round(t / d["us"] / 1e3 / HBM_PEAK_GBS, 4),
                            "what": "rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch of this "
                                    "kernel on this workload, copy-calibrated, / launch time"}
        # the measured-bytes figures beside the algorithmic ones, at the same level
        roof["achieved_physical"] = roof["physical"]["achieved"]
        roof["frac_physical"] = roof["physical"]["frac"]
    return roof, {n: round(v["us"], 2) for n, v in per.items()}


def one_at_a_time(dev):
    """The same four gradients one model after another through ``compute``
    (autotuned variant, captured graphs, shot streams, the history kept as
    far as it fits)."""
    import torch
    gc.collect()
    torch.cuda.empty_cache()
    from fastgpr.workloads import c3_uneven
    wls = c3_uneven()
    obs = []
    with torch.no_grad():
        for wl in wls:
            obs.append(_compute(wl, dev, *_tensors(wl, dev, (wl.er, wl.se)),
                                time_block="auto").detach())

    def run():
        t0 = time.perf_counter()
        for wl, o in zip(wls, obs):
            er, se, mr = _tensors(wl, dev, (wl.er0, wl.se0))
            er = torch.clamp(er, min=1).requires_grad_(True)
            se = torch.clamp(se, min=0).requires_grad_(True)
            pred = _compute(wl, dev, er, se, mr, time_block="auto", checkpoint="auto")
            (0.5 * ((pred - o) ** 2).sum()).backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    run()
    el = min(run() for _ in range(2))
    cu = sum(w.cell_updates_per_gradient() for w in wls)
    return {"value": round(cu / el / 1e9, 3), "seconds": round(el, 3),
            "what": "compute() per model: autotuned variant, captured graphs, three shot "
                    "streams, checkpoint='auto' (best of 2 after a warm-up)"}


def c2_line(dev, args, extras=True):
    """The cross-hole batch through ``compute`` (autotuned, three shot
    streams, captured graphs)."""
    import torch
    import fastgpr
    from fastgpr.workloads import c2_crosshole
    wl = c2_crosshole()
    er_true, se_true, mr = _tensors(wl, dev, (wl.er, wl.se))
    er0, se0, _ = _tensors(wl, dev, (wl.er0, wl.se0))
    tb = (None if args.time_block == "none" else
          "auto" if args.time_block == "auto" else tuple(int(v) for v in args.time_block.split(",")))
    kw = dict(exec_mode=args.exec_mode, graph_chunk=args.chunk, time_block=tb,
              shot_streams=args.shot_streams)
    with torch.no_grad():
        obs = _compute(wl, dev, er_true, se_true, mr, **kw).detach().clone()
    er = er0.clone().requires_grad_(True)
    se = se0.clone().requires_grad_(True)

    def gradient_step():
        er.grad = None
        se.grad = None
        pred = _compute(wl, dev, er, se, mr, **kw)
        (0.5 * ((pred - obs) ** 2).sum()).backward()

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
    value = cu / (elapsed / args.steps) / 1e9
    if tb == "auto":
        from fastgpr.autotune import _cache, _key
        Gk = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                    None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                    er.detach(), se.detach(), mr, 1, wl.rx.shape[1], wl.dt,
                                    'HORIPML', None, None, False, False, None,
                                    torch.tensor(wl.src), torch.tensor(wl.rx),
                                    time_block="auto", shot_streams=args.shot_streams)
        tb_used = _cache[_key(Gk)]
    else:
        tb_used = tb
    src, rx = torch.tensor(wl.src), torch.tensor(wl.rx)
    roofline, per_kernel = profile_kernels(wl, er.detach(), se.detach(), mr, src, rx, tb_used,
                                           args.profile_reps)
    t = measured_traffic(TRAFFIC_FILE, f"{wl.name}|{roofline['kernel'].split(' ')[0]}")
    if t is not None:
        roofline["traffic"] = t
        us = roofline["avg_launch_us"]
        roofline["physical"] = {"achieved": round(t / us / 1e3, 1),
                                "frac": round(t / us / 1e3 / HBM_PEAK_GBS, 4)}
        roofline["achieved_physical"] = roofline["physical"]["achieved"]
        roofline["frac_physical"] = roofline["physical"]["frac"]
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
                "us_per_block": round(blk, 2), "achieved": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 fields, f64 gradient accumulation",
        "data": "synthetic (seeded layered model, inputs resident in HBM)",
        "config": {"workload": wl.name, "grid": [wl.NI, wl.NJ], "dx_m": wl.dx, "cpml": wl.P,
                   "nt": wl.nt, "shots": wl.S, "receivers": int(wl.rx.shape[1]),
                   "cell_updates_per_gradient": cu, "exec": args.exec_mode,
                   "time_block": list(tb_used) if tb_used else None},
        "roofline": roofline,
        "kernels_us": {k: round(v["us"], 2) for k, v in per_kernel.items()},
    }
    if extras and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps)
    return out


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


def measured_traffic(path, key):
    """HBM bytes per launch of the kernel and workload ``key``
    ("workload|kernel") from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE per dispatch, copy-calibrated), or None when the counter runs
    did not measure this exact kernel on this exact workload."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("measured", {}).get(key)
    return None if k is None else k.get("bytes_per_launch")


# Dispatches the counter summary attributes to a workload: kernel name
# prefix and grid size (work-items) of the probe's launches, for every
# variant the probe may run.
def _pmc_targets() -> dict:
    from fastgpr import _native
    from fastgpr.batching import MULTI_VARIANTS
    out = {}
    for name, (NI, NJ, S), kern, variants in (
            ("c3_model3_1200x600_15shots", (1200, 600, 15), "k_tb_multi", MULTI_VARIANTS),
            ("c2_crosshole_400x200_13shots", (400, 200, 13), "k_tb", [(64, 1024, 6)])):
        for ri, nt, k in variants:
            _, _, ti, tj = _native.tb_tiles(NI, NJ, ri, nt, k)
            for adj in ("false", "true"):
                out[(f"{kern}<{ri}, {nt}, {k}, {adj}>", S * ti * tj * nt)] = \
                    f"{name}|{kern}<{ri},{nt},{k},{adj}>"
    return out


COPY_BYTES = 3 * 256 * 1024 * 1024       # the probe's three 256 MiB copies, read and written


def _pmc_rows(root):
    """(kernel name, grid size, counter, value) of every dispatch in the
    counter_collection CSV files under ``root``."""
    import csv
    out = []
    for f in Path(root).rglob("*counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                name = r["Kernel_Name"].replace("(anonymous namespace)::", "")
                name = name[5:] if name.startswith("void ") else name
                out.append((name, int(r["Grid_Size"]), r["Counter_Name"], float(r["Counter_Value"]),
                            int(r.get("Dispatch_Id") or 0)))
    return sorted(out, key=lambda t: t[4])


def pmc_summary(fetch_dir, write_dir, out=None):
    """HBM bytes per launch of the dominant kernels from two counter runs of
    the probe: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by
    what the same counter reads for the probe's device copies of known size
    (3 x 256 MiB read and written)."""
    res = {"calibration": {}, "measured": {}}
    per = {}
    for which, root in (("fetch", fetch_dir), ("write", write_dir)):
        rows = _pmc_rows(root)
        # the calibration copies are the copy dispatches before the first solver kernel
        first = next((i for i, r in enumerate(rows) if r[0].startswith("k_")), len(rows))
        copy = [v for n, g, c, v, _ in rows[:first] if "copyBuffer" in n]
        if not copy:
            raise SystemExit(f"no copy-kernel dispatches under {root}")
        kib = sum(copy)
        res["calibration"][which] = {"copy_dispatches": len(copy), "kib_total": kib,
                                     "factor": COPY_BYTES / (kib * 1024)}
        for (prefix, grid), key in _pmc_targets().items():
            vals = [v for n, g, c, v, _ in rows if n.startswith(prefix + "(") and g == grid]
            if vals:
                per.setdefault(key, {})[which] = (len(vals), sum(vals) / len(vals))
    for key, d in per.items():
        if "fetch" not in d or "write" not in d:
            continue
        rd = d["fetch"][1] * 1024 * res["calibration"]["fetch"]["factor"]
        wr = d["write"][1] * 1024 * res["calibration"]["write"]["factor"]
        res["measured"][key] = {"dispatches": [d["fetch"][0], d["write"][0]],
                                "fetch_kib_per_dispatch": round(d["fetch"][1], 1),
                                "write_kib_per_dispatch": round(d["write"][1], 1),
                                "read_bytes": int(rd), "write_bytes": int(wr),
                                "bytes_per_launch": int(rd + wr)}
    with open(out or TRAFFIC_FILE, "w") as f:
        json.dump(res, f, indent=1)
    return res


def pmc_probe(dev, args):
    """Program for the FETCH_SIZE / WRITE_SIZE counter runs: a 256 MiB
    device copy (the calibration: 256 MiB read + 256 MiB written), then the
    headline's multi-problem kernels on the uneven batch's largest model and
    the cross-hole batch's time-blocked kernels, 5 launches each; prints the
    order of the dispatches."""
    import torch
    from fastgpr.uneven import UnevenBatch
    from fastgpr.workloads import c2_crosshole
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    tb = "auto" if args.c3_variant == "auto" else tuple(int(v) for v in args.c3_variant.split(","))
    wls, probs = c3_problems(dev, [tb] * 4)
    ub = UnevenBatch(dev, probs)
    ub.run()
    torch.cuda.synchronize()
    roof, _ = multi_roofline(ub, 5)
    torch.cuda.synchronize()
    del ub, probs
    torch.cuda.empty_cache()
    wl = c2_crosshole()
    er, se, mr = _tensors(wl, dev, (wl.er0, wl.se0))
    profile_kernels(wl, er, se, mr, torch.tensor(wl.src), torch.tensor(wl.rx), (64, 1024, 6), 5)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "done", "c3": roof["workload_key"],
                      "c2": f"{wl.name}|k_tb<64,1024,6,true>"}), flush=True)


# Grids and shot counts of the autotuner envelope: (NI, NJ, S, steps).  The
# 1600 x 800 32-shot batch is left out: its history (320 GB) does not fit.
ENVELOPE = [(NI, NJ, S, nt) for NI, NJ, nt in ((200, 100, 3000), (400, 200, 3000),
                                              (800, 400, 1500), (1600, 800, 1500))
            for S in (1, 4, 13, 32) if (NI, S) != (1600, 32)]


def tuner_envelope(dev, cases=ENVELOPE, grad_reps=3):
    """For each grid: every variant's time per step as the tuner takes it
    (graph-replayed blocks) and launched one by one, the picks of both, and
    the gradient rate (compute + backward, two timings of ``grad_reps``)
    with each of the three fastest variants of either timing pinned."""
    import torch
    import fastgpr
    from fastgpr import autotune
    from fastgpr.workloads import crosshole
    rows = []
    for NI, NJ, S, nt in cases:
        wl = crosshole(NI, NJ, S, nt)
        er_t, se_t, mr = _tensors(wl, dev, (wl.er, wl.se))
        er0, se0, _ = _tensors(wl, dev, (wl.er0, wl.se0))
        with torch.no_grad():
            obs = _compute(wl, dev, er_t, se_t, mr, time_block=(64, 1024, 6)).detach().clone()
        G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                   None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                   er0, se0, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                                   False, False, None, torch.tensor(wl.src), torch.tensor(wl.rx),
                                   time_block="auto", shot_streams=3)
        reps = autotune.timing_reps(G)
        times = {"graph": {}, "eager": {}}
        for mode in times:
            for tb in autotune.CANDIDATES:
                try:
                    _native_tiles(G, tb)
                except (RuntimeError, ValueError):
                    continue
                times[mode][tb] = 1e6 * autotune.time_variant(G, er0, se0, mr, tb, reps=reps,
                                                              graph=mode == "graph")
        top = set()
        for t in times.values():
            top |= set(sorted(t, key=t.get)[:3])
        er = er0.clone().requires_grad_(True)
        se = se0.clone().requires_grad_(True)

        def rate(tb):
            def step():
                er.grad = None
                se.grad = None
                pred = _compute(wl, dev, er, se, mr, time_block=tb, shot_streams=3)
                (0.5 * ((pred - obs) ** 2).sum()).backward()
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(grad_reps):
                step()
            torch.cuda.synchronize()
            return wl.cell_updates_per_gradient() / ((time.perf_counter() - t0) / grad_reps) / 1e9
        rates = {tb: [round(rate(tb), 1) for _ in range(2)] for tb in sorted(top)}
        best = max(rates, key=lambda tb: sum(rates[tb]))
        picks = {m: min(t, key=t.get) for m, t in times.items()}
        row = {"grid": f"{NI}x{NJ}", "shots": S, "steps": nt,
               "pick": list(picks["graph"]), "eager_pick": list(picks["eager"]),
               "best_pinned": list(best),
               "pick_vs_best": round(sum(rates[picks["graph"]]) / sum(rates[best]), 4)
               if picks["graph"] in rates else None,
               "variants": {",".join(map(str, tb)): {
                   "graph_us_per_step": round(times["graph"][tb], 3),
                   "eager_us_per_step": round(times["eager"][tb], 3),
                   "gradient_gcell_updates_per_s": rates[tb]} for tb in sorted(top)}}
        print(json.dumps(row), flush=True)
        rows.append(row)
        fastgpr.solver.release_graphs()
    return rows


def _native_tiles(G, tb):
    from fastgpr import _native
    _native.tb_tiles(G.NI, G.NJ, *tb)


def c4_memory_bound(dev, reps=2):
    """The 2000 x 1000, 4-shot, 8000-step gradient three ways: the
    reference gradient with its whole snapshot history in HBM (256 GB:
    ``checkpoint="auto"`` keeps it on a 288 GB device), the reference
    gradient with a checkpoint every 90 steps (recompute), and the exact
    gradient (the finite-difference-checked mode) with a checkpoint every 96
    steps.  Seconds per gradient after one warm-up, best of ``reps``."""
    import torch
    import fastgpr
    from fastgpr import checkpoint as ck
    from fastgpr.workloads import c4_memory
    wl = c4_memory()
    er_t, se_t, mr = _tensors(wl, dev, (wl.er, wl.se))
    er0, se0, _ = _tensors(wl, dev, (wl.er0, wl.se0))
    with torch.no_grad():
        obs = _compute(wl, dev, er_t, se_t, mr, time_block="auto").detach().clone()
    del er_t, se_t
    cu = wl.cell_updates_per_gradient()
    out = {"workload": wl.name, "grid": [wl.NI, wl.NJ], "shots": wl.S, "steps": wl.nt,
           "cell_updates_per_gradient": cu}
    modes = (("reference_whole_history", dict(checkpoint="auto")),
             ("reference_checkpoint_90", dict(checkpoint=90)),
             ("exact_checkpoint_96", dict(gradient="exact", checkpoint=96, graph_chunk=48)))
    for name, kw in modes:
        er = er0.clone().requires_grad_(True)
        se = se0.clone().requires_grad_(True)
        times = []
        for r in range(reps + 1):
            er.grad = None
            se.grad = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pred = _compute(wl, dev, er, se, mr, time_block="auto", **kw)
            (0.5 * ((pred - obs) ** 2).sum()).backward()
            torch.cuda.synchronize()
            if r:
                times.append(time.perf_counter() - t0)
        assert torch.isfinite(er.grad).all() and torch.isfinite(se.grad).all()
        best = min(times)
        rec = {"seconds": round(best, 4), "gcell_updates_per_s": round(cu / best / 1e9, 1)}
        if kw.get("checkpoint") == "auto":
            G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window,
                                       ['hertzian', 'z'], None, None, None, None, wl.P,
                                       ['ricker', 1, wl.freq], wl.S, er.detach(), se.detach(), mr,
                                       1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None, False,
                                       False, None, torch.tensor(wl.src), torch.tensor(wl.rx))
            rec["whole_history_kept"] = ck.history(G, "auto") is None
            rec["history_gb"] = round(wl.S * wl.nt * wl.NI * wl.NJ * 4 / 1e9, 1)
        out[name] = rec
        del er, se, pred
        torch.cuda.empty_cache()
    return out


# The reference's only recorded timings: its example notebook's epoch wall
# times (unstated NVIDIA GPU; example.ipynb:145, 148 for phase 1, :295-748 for
# phase 2) and 262.99 s for all 202 epochs (:748).
NOTEBOOK_REFERENCE = {"phase1_epoch_median_s": 1.15, "phase2_epoch_median_s": 1.33,
                      "total_202_epochs_s": 262.99, "hardware": "unstated NVIDIA GPU",
                      "source": "example.ipynb:145-748 (recorded stdout)"}


def notebook_inversion(dev, phase1=50, phase2=152):
    """The reference notebook's inversion (example.ipynb cells 0-1) through
    ``fastgpr.inversion``: 220 x 120 padded models rebuilt from the
    reference's newmodel.npy (tests/golden/notebook_models.npz), 40 shots,
    200 fixed receivers, 10-cell CPML, dt 1e-10 s, 1001 steps, TV on, Adam.
    ``phase1`` epochs on epsilon, then ``phase2`` on epsilon and sigma (the
    notebook's 50 + 152); median epoch wall time per phase and the total
    beside the notebook's recorded ones, and the loss of every epoch against
    the loss the notebook printed for it."""
    import numpy as np
    import torch
    from fastgpr.compute import compute
    from fastgpr.inversion import InversionConfig, invert
    d = np.load(Path(__file__).resolve().parent / "tests" / "golden" / "notebook_models.npz",
                allow_pickle=False)
    t = lambda k: torch.tensor(d[k], device=dev).unsqueeze(-1)          # float64, as the notebook
    ep_t, se_t, ep_init, se_init = t("ep_true"), t("se_true"), t("ep_init"), t("se_init")
    mr = torch.ones_like(ep_t)
    dx, nx, ny, nz = torch.tensor(0.05), torch.tensor(220), torch.tensor(120), torch.tensor(1)
    kw = dict(nsrc=1, nrx=200, time_windows=torch.tensor(1e-7), dt=1e-10,
              source_location=torch.tensor([0.5, 0.50, 0]), source_step=torch.tensor([0.25, 0, 0]),
              receiver_location=torch.tensor([0.5, 5.50, 0]),
              receiver_step=torch.tensor([0.05, 0, 0]), pmlthick=10, Fixedreceiver=True,
              time_block="auto")
    args = (dev, dx, ny, nx, nz, ['hertzian', 'z'], ['ricker', 1, 1e8], 40)
    with torch.no_grad():
        observed, _ = compute(*args, ep_t, se_t, mr, 1, 200,
                              **{k: v for k, v in kw.items() if k not in ("nsrc", "nrx")})
    torch.cuda.synchronize()
    cfg = InversionConfig(n_epochs=phase1 + phase2, phase1_epochs=phase1, pmlthick=10)
    out = invert(args, kw, ep_init, se_init, mr, observed, cfg)
    sec = out["epoch_seconds"]
    p1, p2 = float(np.median(sec[:phase1])), float(np.median(sec[phase1:]))
    cu = 2 * 40 * 220 * 120 * 1001            # cell updates of one gradient (forward + adjoint)
    ref = NOTEBOOK_REFERENCE
    got = np.asarray(out["losses"], np.float64)
    rec = np.asarray(d["losses"][:len(got)], np.float64)
    rel = np.abs(got - rec) / rec
    res = {"what": "the reference notebook's inversion (forward + adjoint + TV + Adam per epoch, "
                   "epsilon only in phase 1, epsilon and sigma in phase 2), wall time per epoch",
           "epochs": [phase1, phase2], "first_epoch_s": round(sec[0], 4),
           "phase1_epoch_median_s": round(p1, 4), "phase2_epoch_median_s": round(p2, 4),
           "total_s": round(float(np.sum(sec)), 3),
           "phase1_gcell_updates_per_s": round(cu / p1 / 1e9, 1),
           "reference": ref,
           "speedup_vs_notebook_epochs": [round(ref["phase1_epoch_median_s"] / p1, 1),
                                          round(ref["phase2_epoch_median_s"] / p2, 1)],
           "loss_vs_notebook": {
               "epochs_compared": int(len(got)),
               "rel_diff": {str(e): float(f"{rel[e]:.2e}") for e in (0, 1, 2, 10, 49, 50, 100, 201)
                            if e < len(got)},
               "max_rel_diff_phase1": float(f"{rel[:phase1].max():.2e}"),
               "first_epoch_over_1e-3": (int(np.argmax(rel > 1e-3)) if (rel > 1e-3).any()
                                         else None),
               "last": [round(float(got[-1]), 6), round(float(rec[-1]), 6)],
               "mean_last_50": [round(float(got[-50:].mean()), 4),
                                round(float(rec[-50:].mean()), 4)]}}
    if phase1 + phase2 == 202:
        res["speedup_vs_notebook_total"] = round(ref["total_202_epochs_s"] / float(np.sum(sec)), 1)
    return res


def plain_path(dev, value):
    """The plain GPU path on the cross-hole batch: one shot per launch
    sequence, per-step kernels launched eagerly from the host loop, no
    graph, no batching, no time blocking; one full gradient, shot by shot.
    ``value`` (the headline) over its rate is the speedup of the batched
    path."""
    import torch
    import fastgpr
    from fastgpr.workloads import c2_crosshole
    wl = c2_crosshole()
    er, se, mr = _tensors(wl, dev, (wl.er0, wl.se0))
    with torch.no_grad():
        obs = _compute(wl, dev, *_tensors(wl, dev, (wl.er, wl.se)), time_block="auto")
    er = er.requires_grad_(True)
    se = se.requires_grad_(True)
    src, rx = torch.tensor(wl.src), torch.tensor(wl.rx)
    common = (dev, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'], ['ricker', 1, wl.freq], 1)

    def one_gradient():
        for s in range(wl.S):
            pred, _ = fastgpr.compute(*common, er, se, mr, 1, wl.rx.shape[1],
                                      time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P,
                                      tv=False, customsource=src[s:s + 1],
                                      customreceiver=rx[s:s + 1], exec_mode="eager",
                                      time_block=None)
            (0.5 * ((pred - obs[s:s + 1]) ** 2).sum()).backward()

    one_gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    one_gradient()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rate = wl.cell_updates_per_gradient() / el / 1e9
    return {"value": round(rate, 3), "unit": "Gcell-updates/s", "workload": wl.name,
            "ms_per_gradient": round(el * 1e3, 1), "speedup": round(value / rate, 2),
            "what": "per-step eager launches, one shot per launch sequence, no graph"}


def serve_bench(dev, seconds, rate, histories="strict", max_hold_s=2.0):
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
    gc.collect()
    torch.cuda.empty_cache()
    from fastgpr.serving import Request, Server, poisson_trace, serving_request
    rng = np.random.default_rng(4)
    arrivals = poisson_trace(rate, seconds, seed=4)
    lazy = len(arrivals) > 200
    sampled = set(k * len(arrivals) // 8 for k in range(8))
    pool = {}        # models and noise shared between requests: cheap generation, no GIL stalls
    reqs = [] if lazy else [serving_request(i, rng, pool=pool) for i in range(len(arrivals))]
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
    srv = Server(dev, histories=histories, max_hold_s=max_hold_s)
    t0 = time.perf_counter()

    def submitter():
        for i, at in enumerate(arrivals):
            if lazy:
                reqs.append(serving_request(i, rng, pool=pool))
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
            print(f"[bench] serving: {srv.stats.completed}/{len(arrivals)} done, {len(srv.active)} active, "
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
    del srv
    gc.collect()
    torch.cuda.empty_cache()
    for r in picks:
        gc.collect()                     # release the last arena before sizing the next
        torch.cuda.empty_cache()
        a = Server(dev)
        ra = a.submit(Request(r.rid, r.wl, r.model, r.obs, tb=r.tb))
        a.run()
        torch.cuda.synchronize()
        del a
        if "er" not in ra.result:
            raise RuntimeError(f"request {r.rid} served alone failed: {ra.result}")
        for k in ("er", "se", "traces"):
            x, y = r.result[k].double(), ra.result[k].double()
            worst = max(worst, float((x - y).abs().max() / max(y.abs().max(), 1e-30)))
    by_grid = {}
    for r in reqs:
        if r.done_at is None or r.admitted_at is None:
            continue
        by_grid.setdefault(f"{r.wl.NI}x{r.wl.NJ}", []).append(
            (r.done_at - r.arrival, r.admitted_at - r.arrival, r.wl.cell_updates_per_gradient()))
    grids = {}
    for g, v in sorted(by_grid.items(), key=lambda kv: int(kv[0].split("x")[0])):
        lat = np.array([x[0] for x in v])
        wait = np.array([x[1] for x in v])
        grids[g] = {"n": len(v), "p50_s": round(float(np.percentile(lat, 50)), 3),
                    "p99_s": round(float(np.percentile(lat, 99)), 3),
                    "queue_p99_s": round(float(np.percentile(wait, 99)), 3),
                    "mean_gcell_updates": round(float(np.mean([x[2] for x in v])) / 1e9, 2)}
    return {
        "metric": "serving: p50/p99 request latency, sustained gradient Gcell-updates/s",
        "p50_latency_s": round(summ["p50_latency_s"], 3),
        "p99_latency_s": round(summ["p99_latency_s"], 3),
        "sustained_gcell_updates_per_s": round(summ["sustained_gcell_updates_per_s"], 3),
        "completed": summ["completed"], "requests": len(reqs),
        "arrival_window_s": seconds, "rate_per_s": rate, "wall_s": round(summ["wall_s"], 2),
        "offered_gcell_updates_per_s": round(offered_cu[0] / seconds / 1e9, 3),
        "max_concurrent_requests": summ["max_concurrent_requests"],
        "peak_admitted_gb": round(summ["peak_admitted_bytes"] / 1e9, 2),
        "refused": summ["refused"], "holds": summ["holds"],
        "partial_histories": summ["partial_histories"],
        "histories": histories, "history_waits": summ["history_waits"],
        "recompute_share": round(summ["recompute_share"], 4),
        "launches": summ["launches"], "ticks": summ["ticks"],
        "host_ms_per_tick": round(summ["host_ms_per_tick"], 3),
        "admit_host_s": round(summ["admit_host_s"], 3),
        "window_ns": [m0, m1],
        "by_grid": grids,
        "alone_vs_mix_max_rel_diff": worst}


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
