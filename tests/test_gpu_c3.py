"""The uneven batch at full size (GPU): four seeded models, 64 shots --
300x150/2000 steps, 500x250/3000, 800x400/4500, 1200x600/6000 -- computed
together by ``fastgpr.uneven.UnevenBatch`` from one captured graph.  Every
model's normalised gradients and traces must be bitwise those of
``fastgpr.compute`` on that model alone (same time-blocking variant); the
smallest model is also checked against the C oracle: traces within 1e-5
relative L2 of its FP32 mode, gradients within 1e-4 of its FP64 gradient."""
import time

import numpy as np
import pytest
import torch

from helpers import oracle_problem, rel_l2
from oracle import c_oracle as CO
from oracle import fdtd_oracle as O

pytestmark = pytest.mark.gpu

TB = (64, 1024, 6)


def _tensors(wl, dev, model):
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
    return t(model[0]), t(model[1]), t(wl.mr)


def _compute(wl, dev, er, se, mr, **kw):
    import fastgpr
    return fastgpr.compute(dev, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'], ['ricker', 1, wl.freq],
                           wl.S, er, se, mr, wl.src.shape[1], wl.rx.shape[1],
                           time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P, tv=False,
                           customsource=torch.tensor(wl.src), customreceiver=torch.tensor(wl.rx),
                           time_block=TB, **kw)[0]


def _problem(wl, dev, obs):
    import fastgpr
    from fastgpr.uneven import Problem
    er, se, mr = _tensors(wl, dev, (wl.er0, wl.se0))
    er, se = torch.clamp(er, min=1), torch.clamp(se, min=0)
    G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                               None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                               er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                               False, False, None, torch.tensor(wl.src), torch.tensor(wl.rx))
    return Problem(G, er, se, mr, obs, TB)


def test_c3_uneven_batch_equals_each_model_alone(hip_device):
    from fastgpr.uneven import UnevenBatch
    from fastgpr.workloads import c3_uneven
    dev = hip_device
    wls = c3_uneven()
    obs = []
    with torch.no_grad():
        for wl in wls:
            obs.append(_compute(wl, dev, *_tensors(wl, dev, (wl.er, wl.se))).detach().clone())
    t0 = time.perf_counter()
    ub = UnevenBatch(dev, [_problem(wl, dev, o) for wl, o in zip(wls, obs)])
    print("uneven batch built and captured in", round(time.perf_counter() - t0, 1), "s:",
          ub.summary(), flush=True)
    t0 = time.perf_counter()
    res = ub.run()
    torch.cuda.synchronize()
    print("uneven batch ran in", round(time.perf_counter() - t0, 2), "s", flush=True)
    got = [{k: v.detach().cpu().numpy() for k, v in r.items()} for r in res]
    # a second replay gives the same results (the graph restarts every job)
    res = ub.run()
    torch.cuda.synchronize()
    for g, r in zip(got, res):
        for k in g:
            assert np.array_equal(g[k], r[k].cpu().numpy()), ("replay", k)
    del ub, res
    torch.cuda.empty_cache()
    for wl, o, g in zip(wls, obs, got):
        print("alone:", wl.name, flush=True)
        er, se, mr = _tensors(wl, dev, (wl.er0, wl.se0))
        er = torch.clamp(er, min=1).detach().requires_grad_(True)
        se = torch.clamp(se, min=0).detach().requires_grad_(True)
        pred = _compute(wl, dev, er, se, mr, checkpoint="auto", shot_streams=3)
        (0.5 * ((pred - o) ** 2).sum()).backward()
        torch.cuda.synchronize()
        assert np.array_equal(pred.detach().cpu().numpy(), g["traces"]), (wl.name, "traces")
        assert np.array_equal(er.grad.cpu().numpy(), g["er"]), (wl.name, "er")
        assert np.array_equal(se.grad.cpu().numpy(), g["se"]), (wl.name, "se")
        del er, se, pred
        torch.cuda.empty_cache()

    # the smallest model against the C oracle
    wl, o, g = wls[0], obs[0].cpu().numpy(), got[0]
    if not (CO.available() and CO.available(np.float32)):
        pytest.skip("oracle/_build/liboracle*.so not built (make -C oracle)")
    t0 = time.perf_counter()
    f = np.float32
    pb = oracle_problem(wl)
    c, pml = O.setup(pb, f)
    tr32, _ = CO.forward(c, pml, pb.P, pb.dx, pb.src, O.source_scale(pb.wavelet, pb.dx, f),
                         pb.rx, pb.nt, store=False, dtype=f)
    e = rel_l2(o, tr32)
    print(f"{wl.name}: true-model traces rel-L2 vs FP32 oracle {e:.3e}", flush=True)
    assert e <= 1e-5
    pb0 = oracle_problem(wl, model=(wl.er0, wl.se0))
    c, pml = O.setup(pb0, f)
    tr0, _ = CO.forward(c, pml, pb0.P, pb0.dx, pb0.src, O.source_scale(pb0.wavelet, pb0.dx, f),
                        pb0.rx, pb0.nt, store=False, dtype=f)
    e = rel_l2(g["traces"], tr0)
    print(f"{wl.name}: initial-model traces rel-L2 vs FP32 oracle {e:.3e}", flush=True)
    assert e <= 1e-5
    resid = (g["traces"] - o).astype(np.float64)
    c, pml = O.setup(pb0, np.float64)
    _, snaps = CO.forward(c, pml, pb0.P, pb0.dx, pb0.src,
                          O.source_scale(pb0.wavelet, pb0.dx, np.float64), pb0.rx, pb0.nt, True)
    cadj, _ = O.setup(pb0, np.float64, adjoint=True)
    ge, gs = CO.adjoint_reference(cadj, pml, pb0.P, pb0.dx, pb0.rx, resid, snaps)
    del snaps
    for k, ref in (("er", O.normalise(ge)), ("se", O.normalise(gs))):
        e = rel_l2(g[k][..., 0], ref)
        print(f"{wl.name}: gradient {k} rel-L2 vs FP64 oracle {e:.3e}", flush=True)
        assert e <= 1e-4, k
    print("oracle checks", round(time.perf_counter() - t0, 1), "s", flush=True)
