"""Uneven batches (GPU): several jobs with different grids, shot counts,
step counts and checkpointing advanced together by multi-problem launches
give bitwise the results of each job alone and of fastgpr.compute."""
import numpy as np
import pytest
import torch

from helpers import run_product

pytestmark = pytest.mark.gpu


def _job_inputs(wl, dev, model=None):
    import fastgpr
    er_np, se_np = (wl.er, wl.se) if model is None else model
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
    er, se, mr = t(er_np), t(se_np), t(wl.mr)
    er = torch.clamp(er, min=1)
    se = torch.clamp(se, min=0)
    G = fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                               None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                               er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                               False, False, None, torch.tensor(wl.src), torch.tensor(wl.rx))
    return G, er, se, mr


def _cases():
    from fastgpr.workloads import small_case
    return [
        (small_case(NI=60, NJ=44, S=3, nt=157, seed=5), (64, 1024, 6), None),
        (small_case(NI=90, NJ=70, S=2, nt=200, seed=6), (32, 512, 4), 36),
        (small_case(NI=45, NJ=130, S=1, nt=96, seed=7), (64, 1024, 6), 24),
    ]


def test_mixed_jobs_equal_alone_and_compute(hip_device):
    from fastgpr.batching import Job, Scheduler
    dev = hip_device
    cases = _cases()
    obs = []
    for wl, tb, ck in cases:
        tr, _ = run_product(wl, dev, time_block=tb)
        obs.append(torch.tensor(tr).to(dev))
    inits = [(wl.er * 0.97, wl.se * 1.1) for wl, _, _ in cases]

    def make_jobs():
        jobs = []
        for (wl, tb, ck), o, init in zip(cases, obs, inits):
            G, er, se, mr = _job_inputs(wl, dev, init)
            jobs.append(Job(G, er, se, mr, obs=o, tb=tb, checkpoint=ck))
        return jobs

    mixed = make_jobs()
    sch = Scheduler(dev)
    for j in mixed:
        sch.add(j)
    sch.run()
    torch.cuda.synchronize()
    print("launches in the mix:", sch.launches, flush=True)
    for idx, ((wl, tb, ck), o, init) in enumerate(zip(cases, obs, inits)):
        alone = make_jobs()[idx]
        s1 = Scheduler(dev)
        s1.add(alone)
        s1.run()
        torch.cuda.synchronize()
        resid = (mixed[idx].result["traces"] - o).cpu().numpy()
        _, g = run_product(wl, dev, resid=resid, model=init, time_block=tb)
        for k in ("er", "se"):
            a = mixed[idx].result[k].cpu().numpy()[..., 0]
            b = alone.result[k].cpu().numpy()[..., 0]
            assert np.array_equal(a, b), (idx, k, "mix vs alone")
            assert np.array_equal(a, g[k]), (idx, k, "mix vs compute()")
        assert torch.equal(mixed[idx].result["traces"], alone.result["traces"])


def test_forward_only_jobs_in_the_mix(hip_device):
    from fastgpr.batching import Job, Scheduler
    dev = hip_device
    cases = _cases()
    sch = Scheduler(dev)
    jobs = []
    for wl, tb, ck in cases:
        G, er, se, mr = _job_inputs(wl, dev)
        jobs.append(Job(G, er, se, mr, obs=None, tb=tb))
        sch.add(jobs[-1])
    sch.run()
    torch.cuda.synchronize()
    for (wl, tb, ck), j in zip(cases, jobs):
        tr, _ = run_product(wl, dev, time_block=tb)
        assert np.array_equal(j.result["traces"].cpu().numpy(), tr)
