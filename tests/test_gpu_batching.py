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
            jobs.append(Job(G, er, se, mr, obs=o, tb=tb, history=ck))
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
        resid = (mixed[idx].results()["traces"] - o).cpu().numpy()
        _, g = run_product(wl, dev, resid=resid, model=init, time_block=tb)
        for k in ("er", "se"):
            a = mixed[idx].results()[k].cpu().numpy()[..., 0]
            b = alone.results()[k].cpu().numpy()[..., 0]
            assert np.array_equal(a, b), (idx, k, "mix vs alone")
            assert np.array_equal(a, g[k]), (idx, k, "mix vs compute()")
        assert torch.equal(mixed[idx].results()["traces"], alone.results()["traces"])


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
        assert np.array_equal(j.results()["traces"].cpu().numpy(), tr)


def test_captured_schedule_equals_eager_ticks(hip_device):
    """The same job set run tick by tick from the host and replayed from one
    captured graph (twice, with eager work between the replays) gives
    bitwise the same traces and gradients."""
    from fastgpr.batching import Job, Scheduler
    dev = hip_device
    cases = _cases()
    obs = [torch.tensor(run_product(wl, dev, time_block=tb)[0]).to(dev) for wl, tb, _ in cases]
    inits = [(wl.er * 0.97, wl.se * 1.1) for wl, _, _ in cases]

    def make_jobs():
        out = []
        for (wl, tb, ck), o, init in zip(cases, obs, inits):
            G, er, se, mr = _job_inputs(wl, dev, init)
            out.append(Job(G, er, se, mr, obs=o, tb=tb, history=ck))
        return out

    eager = make_jobs()
    sch = Scheduler(dev)
    for j in eager:
        sch.add(j)
    sch.run()
    torch.cuda.synchronize()
    ref = [{k: v.cpu().clone() for k, v in j.results().items()} for j in eager]

    cap_jobs = make_jobs()
    s2 = Scheduler(dev)
    for j in cap_jobs:
        s2.add(j, start=False)
    cb = s2.capture()
    print("captured ticks", cb.ticks, "launches", cb.launches, flush=True)
    for rep in range(2):
        cb.replay()
        torch.cuda.synchronize()
        torch.zeros(1 << 20, device=dev).sum().item()        # eager work between replays
        for j, r in zip(cap_jobs, ref):
            for k, v in j.results().items():
                assert torch.equal(v.cpu(), r[k]), (rep, k)


@pytest.mark.parametrize("poison", [False, True])
def test_uneven_batch_waves_reuse_memory_and_match_compute(hip_device, monkeypatch, poison):
    """A budget that holds one job's whole history but not all three: the
    jobs run in waves carved from one buffer (a wave reuses the memory of the
    one before) and a job too large alone keeps a shorter snapshot ring; every
    result is bitwise compute()'s for that model alone, on every replay --
    also with the carved buffers filled with NaN bytes first (``poison``), so
    no job reads memory it has not written."""
    from fastgpr import batching
    from fastgpr.batching import job_bytes
    from fastgpr.uneven import Problem, UnevenBatch, plan_waves
    dev = hip_device
    if poison:
        arena_init = batching.Arena.__init__

        def nan_arena(self, device, nbytes):
            arena_init(self, device, nbytes)
            self.buf.fill_(0xFF)
        monkeypatch.setattr(batching.Arena, "__init__", nan_arena)
    cases = _cases()
    obs = [torch.tensor(run_product(wl, dev, time_block=tb)[0]).to(dev) for wl, tb, _ in cases]
    inits = [(wl.er * 0.97, wl.se * 1.1) for wl, _, _ in cases]
    specs = [(wl.NI, wl.NJ, wl.S, wl.nt, wl.rx.shape[1], wl.P, tb) for wl, tb, _ in cases]
    full = [job_bytes(*sp[:6], sp[3], True, sp[6]) for sp in specs]
    budget = int(0.8 * max(full))              # the largest job cannot keep its whole history
    waves, hists = plan_waves(specs, budget)
    assert len(waves) >= 2 and any(not h.full for h in hists), (waves, hists)
    probs = []
    for (wl, tb, _), o, init in zip(cases, obs, inits):
        G, er, se, mr = _job_inputs(wl, dev, init)
        probs.append(Problem(G, er, se, mr, o, tb))
    ub = UnevenBatch(dev, probs, budget=budget)
    print(ub.summary(), flush=True)
    for rep in range(2):
        res = ub.run()
        torch.cuda.synchronize()
        for idx, ((wl, tb, _), o, init) in enumerate(zip(cases, obs, inits)):
            resid = (res[idx]["traces"] - o).cpu().numpy()
            tr, g = run_product(wl, dev, resid=resid, model=init, time_block=tb)
            assert np.array_equal(res[idx]["traces"].cpu().numpy(), tr), (rep, idx)
            for k in ("er", "se"):
                assert np.array_equal(res[idx][k].cpu().numpy()[..., 0], g[k]), (rep, idx, k)


def test_uneven_batch_with_tuned_and_eight_step_variants(hip_device):
    """Problems asking for "auto" get the autotuner's pick among the
    multi-problem variants, one asks for the 8-step 64-row variant; every
    result is bitwise compute()'s with the variant its job runs."""
    from fastgpr.batching import MULTI_VARIANTS
    from fastgpr.uneven import Problem, UnevenBatch
    dev = hip_device
    cases = _cases()
    asked = ["auto", (64, 1024, 8), "auto"]
    obs = [torch.tensor(run_product(wl, dev, time_block=(64, 1024, 6))[0]).to(dev)
           for wl, _, _ in cases]
    inits = [(wl.er * 0.97, wl.se * 1.1) for wl, _, _ in cases]
    probs = []
    for (wl, _, _), o, init, tb in zip(cases, obs, inits, asked):
        G, er, se, mr = _job_inputs(wl, dev, init)
        probs.append(Problem(G, er, se, mr, o, tb))
    ub = UnevenBatch(dev, probs)
    chosen = [p.tb for p in ub.problems]
    print("variants", chosen, flush=True)
    assert chosen[1] == (64, 1024, 8) and all(tuple(t) in MULTI_VARIANTS for t in chosen)
    res = ub.run()
    torch.cuda.synchronize()
    for idx, ((wl, _, _), o, init) in enumerate(zip(cases, obs, inits)):
        resid = (res[idx]["traces"] - o).cpu().numpy()
        tr, g = run_product(wl, dev, resid=resid, model=init, time_block=chosen[idx])
        assert np.array_equal(res[idx]["traces"].cpu().numpy(), tr), idx
        for k in ("er", "se"):
            assert np.array_equal(res[idx][k].cpu().numpy()[..., 0], g[k]), (idx, k)
