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
    torch.cuda.empty_cache()
    srv = Server(dev)
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


def test_nan_request_and_cancellation_fail_alone(hip_device):
    """A request whose model holds a NaN fails with an error (its traces and
    gradient are not finite); a running request cancelled after a few ticks
    returns "cancelled"; the others complete bitwise as alone, and the arena
    is whole again at the end."""
    from fastgpr.serving import Server
    dev = hip_device
    reqs = _requests(4, seed=14)
    er = reqs[1].model[0].copy()
    er[er.shape[0] // 2, er.shape[1] // 2] = np.nan
    reqs[1].model = (er, reqs[1].model[1])
    srv = Server(dev, hbm_budget=1 << 34)
    for r in reqs:
        srv.submit(r)
    for _ in range(3):
        srv.step()
    srv.cancel(reqs[2].rid)
    stats = srv.run()
    torch.cuda.synchronize()
    print(stats, flush=True)
    assert "error" in reqs[1].result and "finite" in reqs[1].result["error"]
    assert reqs[2].result == {"error": "cancelled"}
    assert stats["refused"] == 1 and stats["cancelled"] == 1 and stats["completed"] == 2
    assert srv.arena.free_bytes() == srv.arena.size
    for r in (reqs[0], reqs[3]):
        a = _alone(r, dev)
        for k in ("er", "se", "traces"):
            assert torch.equal(r.result[k], a.result[k]), (r.rid, k)


def test_c5_grid_mix_alone_equals_mix(hip_device):
    """One request of each C5 grid (128x128, 256x128, 512x256, 1024x512)
    at the C5 step counts (1500 to 5000 by the grid diagonal), including a
    1-shot and a 32-shot request (whose history does not fit: it runs with a
    snapshot ring): served together, each result is bitwise the one it gets
    alone; the 128x128 request's traces and gradient agree with the oracle
    (FP32 traces <= 1e-5, FP64 gradient <= 1e-4)."""
    from fastgpr.serving import Server, serving_request
    from helpers import oracle_problem, rel_l2
    from oracle import c_oracle as CO
    from oracle import fdtd_oracle as O
    dev = hip_device
    grids = ((128, 128), (256, 128), (512, 256), (1024, 512))
    reqs = []
    for k, (g, S) in enumerate(zip(grids, (1, 7, 13, 32))):
        diag = (np.hypot(*g) - np.hypot(128, 128)) / (np.hypot(1024, 512) - np.hypot(128, 128))
        nt = int(round(1500 + 3500 * diag))
        r = serving_request(k, np.random.default_rng(40 + k), grids=(g,), shots=(S, S),
                            steps=(nt, nt))
        assert r.wl.nt == nt
        reqs.append(r)
    torch.cuda.empty_cache()
    srv = Server(dev)
    for r in reqs:
        srv.submit(r)
    stats = srv.run()
    torch.cuda.synchronize()
    print(stats, flush=True)
    assert stats["completed"] == 4, [r.result.get("error") for r in reqs]
    assert stats["max_concurrent_requests"] >= 2
    del srv
    torch.cuda.empty_cache()
    for r in reqs:
        a = _alone(r, dev)
        for k in ("er", "se", "traces"):
            assert torch.equal(r.result[k], a.result[k]), (r.rid, k)
    if not (CO.available() and CO.available(np.float32)):
        pytest.skip("oracle/_build/liboracle*.so not built (make -C oracle)")
    r = reqs[0]
    wl = r.wl
    pb = oracle_problem(wl, model=r.model)
    f = np.float32
    c, pml = O.setup(pb, f)
    tr32, _ = CO.forward(c, pml, pb.P, pb.dx, pb.src, O.source_scale(pb.wavelet, pb.dx, f),
                         pb.rx, pb.nt, store=False, dtype=f)
    e = rel_l2(r.result["traces"].cpu().numpy(), tr32)
    print(f"128x128 traces rel-L2 vs FP32 oracle {e:.3e}", flush=True)
    assert e <= 1e-5
    resid = (r.result["traces"].cpu().numpy() - r.obs).astype(np.float64)
    c, pml = O.setup(pb, np.float64)
    _, snaps = CO.forward(c, pml, pb.P, pb.dx, pb.src,
                          O.source_scale(pb.wavelet, pb.dx, np.float64), pb.rx, pb.nt, True)
    cadj, _ = O.setup(pb, np.float64, adjoint=True)
    ge, gs = CO.adjoint_reference(cadj, pml, pb.P, pb.dx, pb.rx, resid, snaps)
    for k, ref in (("er", O.normalise(ge)), ("se", O.normalise(gs))):
        e = rel_l2(r.result[k].cpu().numpy()[..., 0], ref)
        print(f"128x128 gradient {k} rel-L2 vs FP64 oracle {e:.3e}", flush=True)
        assert e <= 1e-4, k


def test_mix_in_a_nan_filled_arena_equals_alone(hip_device):
    """Arena memory is reused from request to request: with the whole arena
    filled with NaN bytes before the first request, every result of a
    staggered mix is finite and bitwise the one served alone, so no job
    reads memory it has not written."""
    from fastgpr.batching import Arena, job_bytes
    from fastgpr.serving import Server
    dev = hip_device
    reqs = _requests(5, seed=23)
    sizes = [job_bytes(r.wl.NI, r.wl.NJ, r.wl.S, r.wl.nt, r.wl.rx.shape[1], r.wl.P, r.wl.nt)
             for r in reqs]
    arena = Arena(dev, int(sum(sorted(sizes)[-2:])))   # blocks are reused as requests finish
    arena.buf.fill_(0xFF)                              # every float32 word a NaN
    srv = Server(dev, arena=arena)
    for r in reqs[:2]:
        srv.submit(r)
    for _ in range(3):
        srv.step()
    for r in reqs[2:]:
        srv.submit(r)
        srv.step()
    summary = srv.run()
    torch.cuda.synchronize()
    print(summary, flush=True)
    assert summary["completed"] == len(reqs) and summary["refused"] == 0
    for r in reqs:
        a = _alone(r, dev)
        for k in ("er", "se", "traces"):
            assert torch.isfinite(r.result[k]).all(), (r.rid, k)
            assert torch.equal(r.result[k], a.result[k]), (r.rid, k)
