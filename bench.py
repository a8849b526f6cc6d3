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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--exec", dest="exec_mode", default="graph", choices=["graph", "eager"])
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=60,
                    help="time steps of the CPU oracle sample (forward + adjoint)")
    ap.add_argument("--profile-reps", type=int, default=200)
    return ap.parse_args()


def kernel_bytes(kernel: int, S: int, NI: int, NJ: int, P: int, nrx: int) -> float:
    """Algorithmic HBM bytes of one launch of each solver kernel.

    Unit figures (bytes per shot-cell unless noted): a field read or write is
    4 B; a coefficient plane is 4 B per model cell, read once per launch and
    shared by the S shots; the snapshot is 4 B; the CPML memory variable of a
    slab cell is read and written (8 B) in the half-step that owns it; the
    FP64 gradient pair is read and written once per model cell (32 B)."""
    cells = S * NI * NJ
    model = NI * NJ
    psi_cells = S * (2 * P * NJ + 2 * P * NI)
    if kernel in (0, 4):     # H: read Ez, Hx, Hy; write Hx, Hy; uH1, uH4; H psi
        b = 20 * cells + 8 * model + 8 * psi_cells
        if kernel == 0:
            b += 8 * S * nrx
        return b
    if kernel == 1:          # forward E: read Ez, Hx, Hy; write Ez, snapshot; uE0, uE1, uE4; E psi
        return 20 * cells + 12 * model + 8 * psi_cells
    if kernel == 3:          # adjoint E: + two snapshot reads, FP64 gradient pair
        return 24 * cells + 12 * model + 8 * psi_cells + 32 * model
    raise ValueError(kernel)


def main():
    args = parse()
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
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
    er_true, se_true, mr = t(wl.er), t(wl.se), t(wl.mr)
    er0, se0 = t(wl.er0), t(wl.se0)
    src = torch.tensor(wl.src)
    rx = torch.tensor(wl.rx)
    kw = dict(time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P, tv=False,
              customsource=src, customreceiver=rx, exec_mode=args.exec_mode,
              graph_chunk=args.chunk)
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
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gradient_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert torch.isfinite(er.grad).all() and torch.isfinite(se.grad).all()

    cu = wl.cell_updates_per_gradient()
    ms = 1e3 * elapsed / args.steps
    value = cu / (elapsed / args.steps) / 1e9

    # ---- dominant-kernel roofline: per-kernel launch durations with HIP events
    Gx = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                er.detach(), se.detach(), mr, 1, wl.rx.shape[1], wl.dt,
                                'HORIPML', None, None, False, False, None, src, rx)
    b = S_.Batch(Gx, store=True)
    coef = S_.cell_coeffs(er.detach(), se.detach(), mr, Gx.dt32, Gx.dx32, False, dev)
    scale = S_.source_scale(Gx).to(dev)
    ge = torch.zeros((wl.NI, wl.NJ), dtype=torch.float64, device=dev)
    gs = torch.zeros_like(ge)
    resid = torch.zeros((wl.S, b.nt, b.nrx), dtype=torch.float32, device=dev)
    p = b.problem(coef, srcscale=scale, resid=resid, grad_eps=ge, grad_sig=gs)
    b.snaps.zero_()
    stream = torch.cuda.current_stream(dev)
    launches_per_grad = {0: b.nt, 1: b.nt, 3: b.nt - 1, 4: b.nt - 1}
    per_kernel = {}
    import ctypes
    for kid in (0, 1, 3, 4):
        reps = args.profile_reps
        _native.check(_native.lib.fgpr_profile_kernel(ctypes.byref(p), kid, 10, 5,
                                                       ctypes.c_void_p(stream.cuda_stream)))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _native.check(_native.lib.fgpr_profile_kernel(ctypes.byref(p), kid, 10, reps,
                                                       ctypes.c_void_p(stream.cuda_stream)))
        e1.record(stream)
        e1.synchronize()
        dur_us = e0.elapsed_time(e1) * 1e3 / reps
        byts = kernel_bytes(kid, wl.S, wl.NI, wl.NJ, wl.P, b.nrx)
        per_kernel[kid] = dict(us=dur_us, bytes=byts, gbs=byts / dur_us / 1e3,
                               share_us=dur_us * launches_per_grad[kid])
    names = {0: "k_h_phase<true> (forward H)", 1: "k_e_phase<true,true,false> (forward E)",
             3: "k_e_phase<false,false,true> (adjoint E + gradient)", 4: "k_h_phase<false> (adjoint H)"}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["share_us"])
    d = per_kernel[dom]
    roofline = {"bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": names[dom], "avg_launch_us": round(d["us"], 2),
                "bytes_per_launch": int(d["bytes"])}

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
                   "graph_chunk": args.chunk},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_us": {names[k]: round(v["us"], 2) for k, v in per_kernel.items()},
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(wl, nsteps: int):
    """Oracle FP64 forward + adjoint over the first ``nsteps`` time steps of
    the same batch (all shots), on one host core."""
    import numpy as np
    from oracle import fdtd_oracle as O
    sys.path.insert(0, str(REPO / "tests"))
    from helpers import oracle_problem
    pb = oracle_problem(wl)
    pb.nt = nsteps
    pb.wavelet = pb.wavelet[:, :nsteps]
    t0 = time.perf_counter()
    tr, snaps = O.run_forward(pb, np.float64, store=True)
    O.run_gradient(pb, tr, snaps, np.float64)
    el = time.perf_counter() - t0
    cu = 2 * wl.S * wl.NI * wl.NJ * nsteps
    cores = 1
    return {"value": round(cu / el / 1e9, 5), "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"numpy FP64 oracle, {wl.name}, all {wl.S} shots, first {nsteps} of "
                      f"{wl.nt} steps forward + adjoint ({el:.1f} s)"}


if __name__ == "__main__":
    main()
