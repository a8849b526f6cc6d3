"""Serving loop (GPU): requests arriving while others run are admitted by
device bytes, continuously batched, and each returns bitwise what it returns
when served alone; requests larger than the budget are refused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRIDS = ((64, 48), (96, 64), (80, 120))


def _requests(n, seed=11, gradient=True):
    from fastgpr.serving import serving_request
    rng = np.random.default_rng(seed)
    return [serving_request(i, rng, grids=GRIDS, shots=(1, 5), steps=(120, 260),
                            gradient=gradient, P=8) for i in range(n)]


def _alone(req, dev):
    from fastgpr.serving import Request, Server
    srv = Server(dev, hbm_budget=1 << 34)
    r = srv.submit(Request(req.rid, req.wl, req.model, req.obs, tb=req.tb))
    srv.run()
    torch.cuda.synchronize()
    return r


def test_staggered_mix_equals_alone(hip_device):
    from fastgpr.batching import job_bytes
    from fastgpr.serving import Server
    dev = hip_device
    reqs = _requests(7)
    reqs[2].tb = (32, 512, 4)                  # a second variant in the mix
    sizes = [job_bytes(r.wl.NI, r.wl.NJ, r.wl.S, r.wl.nt, r.wl.rx.shape[1], r.wl.P, r.wl.nt)
             for r in reqs]
    budget = int(sum(sorted(sizes)[-3:]))      # not every request fits at once
    srv = Server(dev, hbm_budget=budget)
    for r in reqs[:3]:
        srv.submit(r)
    for _ in range(5):
        srv.step()
    for r in reqs[3:]:
        srv.submit(r)
        srv.step()
    summary = srv.run()
    torch.cuda.synchronize()
    print(summary, flush=True)
    assert summary["completed"] == len(reqs)
    assert summary["max_concurrent_requests"] >= 2
    assert summary["peak_admitted_bytes"] <= budget
    for r in reqs:
        a = _alone(r, dev)
        for k in ("er", "se", "traces"):
            assert torch.equal(r.result[k], a.result[k]), (r.rid, k)
        assert np.isfinite(r.result["er"].cpu().numpy()).all()


def test_forward_requests_and_refusal(hip_device):
    from fastgpr.serving import Server
    dev = hip_device
    reqs = _requests(3, seed=12, gradient=False)
    srv = Server(dev, hbm_budget=1 << 34)
    for r in reqs:
        srv.submit(r)
    srv.run()
    for r in reqs:
        a = _alone(r, dev)
        assert torch.equal(r.result["traces"], a.result["traces"])
    tiny = Server(dev, hbm_budget=1024)
    big = tiny.submit(_requests(1, seed=13)[0])
    summary = tiny.run()
    assert "error" in big.result and summary["completed"] == 0


def test_out_of_memory_at_admission_is_retried(hip_device, monkeypatch):
    """A job whose buffers do not fit (the device fuller than the byte
    accounting says) is built again after torch's cache is released; if it
    still does not fit, it goes back to waiting and is admitted once a running
    request has finished.  Results are bitwise those of each request alone.
    Failures are injected on the host, before anything is allocated: request
    1's job fails twice (held), request 2's once (fits after the release)."""
    from fastgpr.serving import Server
    dev = hip_device
    reqs = _requests(3, seed=13)
    srv = Server(dev, hbm_budget=1 << 36)
    real = srv._make_job
    to_fail = {1: 2, 2: 1}
    failed = []

    def flaky(req, tb, C):
        if to_fail.get(req.rid, 0) > 0:
            to_fail[req.rid] -= 1
            failed.append(req.rid)
            raise torch.OutOfMemoryError("injected")
        return real(req, tb, C)
    monkeypatch.setattr(srv, "_make_job", flaky)
    for r in reqs:
        srv.submit(r)
    stats = srv.run()
    torch.cuda.synchronize()
    assert failed == [1, 1, 2] and stats["completed"] == 3
    assert stats["oom_retries"] == 1 and stats["cache_releases"] == 1 and stats["refused"] == 0
    for r in reqs:
        assert "er" in r.result, r.result
        a = _alone(r, dev)
        for k in ("er", "se", "traces"):
            assert torch.equal(r.result[k], a.result[k]), (r.rid, k)
