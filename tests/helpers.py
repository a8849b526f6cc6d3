"""Shared test plumbing: run a Workload through the product (fastgpr on the
GPU) and through the oracle (CPU), on identical inputs."""
from __future__ import annotations

import numpy as np

from fastgpr.workloads import Workload
from oracle import fdtd_oracle as O


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def host_grid(wl: Workload):
    """The product's host-side set-up (wavelet, CPML profile, geometry)
    evaluated on the CPU; used to hand the oracle identical inputs."""
    import torch
    from fastgpr import grid as g
    nt = g.n_iterations(wl.time_window, wl.dt)
    w = g.ricker(nt, wl.dt, 1, wl.freq, torch.tensor(wl.time_window))
    waves = np.zeros((wl.src.shape[1], nt), np.float32)
    waves[0] = w.numpy()
    return nt, waves


def oracle_problem(wl: Workload, model=None) -> O.Problem:
    nt, waves = host_grid(wl)
    er, se = (wl.er, wl.se) if model is None else model
    return O.Problem(er=np.asarray(er, np.float64), se=np.asarray(se, np.float64),
                     mr=wl.mr.astype(np.float64), dx=wl.dx, dt=wl.dt, nt=nt, P=wl.P,
                     src=wl.src[..., :2].astype(np.int64), rx=wl.rx[..., :2].astype(np.int64),
                     wavelet=waves.astype(np.float64))


def product_tensors(wl: Workload, device, er=None, se=None):
    import torch
    er = wl.er if er is None else er
    se = wl.se if se is None else se
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(device)
    return t(er), t(se), t(wl.mr)


def run_product(wl: Workload, device, exec_mode="graph", resid=None, want=("er", "se"),
                graph_chunk=32, tv=False, model=None, time_block=None, gradient="reference",
                checkpoint=None, shot_streams=2):
    """Forward (and, given ``resid``, backward) through fastgpr.compute.

    Returns traces (numpy) and a dict of normalised gradients."""
    import torch
    import fastgpr
    er, se, mr = product_tensors(wl, device, *(model or (None, None)))
    er.requires_grad_("er" in want and resid is not None)
    se.requires_grad_("se" in want and resid is not None)
    traces, dt = fastgpr.compute(device, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'],
                                 ['ricker', 1, wl.freq], wl.S, er, se, mr,
                                 wl.src.shape[1], wl.rx.shape[1],
                                 time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P, tv=tv,
                                 customsource=torch.tensor(wl.src),
                                 customreceiver=torch.tensor(wl.rx),
                                 exec_mode=exec_mode, graph_chunk=graph_chunk,
                                 time_block=time_block, gradient=gradient,
                                 checkpoint=checkpoint, shot_streams=shot_streams)
    grads = {}
    if resid is not None:
        traces.backward(torch.tensor(np.asarray(resid, np.float32)).to(device))
        if er.grad is not None:
            grads["er"] = er.grad.detach().cpu().numpy()[..., 0]
        if se.grad is not None:
            grads["se"] = se.grad.detach().cpu().numpy()[..., 0]
    torch.cuda.synchronize()
    return traces.detach().cpu().numpy(), grads


def run_oracle(wl: Workload, dtype=np.float64, resid=None, model=None):
    pb = oracle_problem(wl, model)
    tr, snaps = O.run_forward(pb, dtype, store=resid is not None)
    out = {}
    if resid is not None:
        ge, gs, ge_raw, gs_raw = O.run_gradient(pb, np.asarray(resid, np.float64), snaps, dtype)
        out = {"er": ge, "se": gs, "er_raw": ge_raw, "se_raw": gs_raw}
    return tr, out
