"""Host-side set-up of the product (fastgpr.grid, CPU torch) against the
oracle and known answers; argument handling that needs no GPU."""
import numpy as np
import pytest
import torch

from fastgpr import grid as g
from oracle import fdtd_oracle as O


def test_dt_rule_and_step_count():
    assert g.cfl_dt(0.05) == pytest.approx(O.cfl_dt(0.05), rel=1e-15)
    dt_t = g.cfl_dt(torch.tensor(0.05))             # float32 tensor in, float32 tensor out
    assert isinstance(dt_t, torch.Tensor) and dt_t.dtype == torch.float32
    assert g.resolve_dt(0.05, 1e-10) == 1e-10
    assert g.resolve_dt(0.05, 1.0) == g.cfl_dt(0.05)
    assert g.n_iterations(torch.tensor(1e-7), 1e-10) == 1001      # example.ipynb
    assert g.n_iterations(1e-7, 1e-10) == O.n_iterations(1e-7, 1e-10)


@pytest.mark.parametrize("freq,dt", [(1e8, 1e-10), (2e8, 5.3e-11), (4e8, 2.1e-11)])
def test_ricker_matches_oracle_fp32(freq, dt):
    nt = 800
    tw = (nt - 1.5) * dt
    w = g.ricker(nt, dt, 1, freq, torch.tensor(tw)).numpy()
    o = O.ricker(nt, dt, freq, 1.0, t_stop=tw, dtype=np.float32)
    o64 = O.ricker(nt, dt, freq, 1.0, t_stop=tw)
    assert w.dtype == np.float32
    np.testing.assert_allclose(w, o, atol=2e-6, rtol=0)
    np.testing.assert_allclose(w, o64, atol=2e-6, rtol=0)
    assert w[-1] == 0.0          # outside [0, time_window]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("dx", [0.05, "tensor"])
def test_pml_profile_matches_oracle(dtype, dx):
    rng = np.random.default_rng(0)
    er = torch.tensor(rng.uniform(3, 9, (30, 24, 1)), dtype=dtype)
    mr = torch.ones((30, 24, 1), dtype=dtype)
    d = torch.tensor(0.05) if dx == "tensor" else 0.05
    pro = g.pml_profile(10, d, 1e-10, er, mr).numpy()
    o = O.cpml_profile(10, 0.05, 1e-10, er[0, :, 0].numpy(), np.ones(24))
    keys = ["ERA", "ERB", "ERE", "ERF", "HRA", "HRB", "HRE", "HRF"]
    assert pro.shape == (8, 10) and pro.dtype == np.float32
    for r, k in enumerate(keys):
        np.testing.assert_allclose(pro[r], o[k], rtol=2e-6, atol=1e-7, err_msg=k)


def test_geometry_notebook_known_answer():
    """example.ipynb acquisition: 40 shots, sources at (0.5+0.25 s, 0.5) m,
    200 fixed receivers at (0.5+0.05 r, 5.5) m, dx = 0.05 m."""
    src, rx = g.geometry(1, 200, 40, torch.tensor([0.5, 0.50, 0]), torch.tensor([0.25, 0, 0]),
                         torch.tensor([0.5, 5.50, 0]), torch.tensor([0.05, 0, 0]),
                         torch.tensor(0.05), Fixedreceiver=True)
    assert src.dtype == torch.int32 and src.shape == (40, 1, 3)
    np.testing.assert_array_equal(src[:, 0, 0].numpy(), 10 + 5 * np.arange(40))
    np.testing.assert_array_equal(src[:, 0, 1].numpy(), np.full(40, 10))
    np.testing.assert_array_equal(src[:, 0, 2].numpy(), np.zeros(40))
    assert rx.shape == (40, 200, 3)
    np.testing.assert_array_equal(rx[:, :, 0].numpy(), np.tile(10 + np.arange(200), (40, 1)))
    np.testing.assert_array_equal(rx[:, :, 1].numpy(), np.full((40, 200), 110))


def test_geometry_rounds_half_to_even():
    # positions 1.5, 2.5, 3.5 cells (exact in float32 with dx = 0.5) -> 2, 2, 4
    src, _ = g.geometry(3, 1, 1, torch.tensor([0.75, 0.0, 0.0]), torch.tensor([0.5, 0.0, 0.0]),
                        torch.zeros(3), torch.zeros(3), 0.5)
    np.testing.assert_array_equal(src[0, :, 0].numpy(), [2, 2, 4])


def test_device_cells_dead_plane_and_bounds():
    idx = torch.tensor([[[3, 4, 0], [5, 6, 1]]], dtype=torch.int32)
    out = g.device_cells(idx, 10, 10, "receiver")
    np.testing.assert_array_equal(out.numpy(), [[[3, 4], [-1, 6]]])
    with pytest.raises(ValueError):
        g.device_cells(torch.tensor([[[11, 0, 0]]]), 10, 10, "source")
    with pytest.raises(ValueError):
        g.device_cells(torch.tensor([[[0, -1, 0]]]), 10, 10, "source")


def test_source_scale_matches_oracle_bitwise():
    from fastgpr.solver import source_scale
    w = g.ricker(300, 1e-10, 1, 1e8, torch.tensor(3e-8))
    G = type("G", (), {})()
    G.dx32 = float(np.float32(0.05))
    G.srcwaves = w[None, :]
    s = source_scale(G).numpy()
    o = O.source_scale(w.numpy()[None, :], 0.05, np.float32)
    np.testing.assert_array_equal(s, o)


def test_compute_refuses_cpu_device():
    import fastgpr
    er = torch.full((30, 30, 1), 4.0)
    with pytest.raises(RuntimeError, match="HIP"):
        fastgpr.compute(torch.device("cpu"), 0.05, 30, 30, 1, ["hertzian", "z"],
                        ["ricker", 1, 1e8], 1, er, er * 0, er * 0 + 1, 1, 1,
                        time_windows=1e-8, pmlthick=5,
                        customsource=torch.tensor([[[15, 15, 0]]]),
                        customreceiver=torch.tensor([[[10, 10, 0]]]))


def test_shot_groups_are_held_to_spare_hardware_queues(monkeypatch):
    """Shot groups of a time-blocked sweep: the requested count, at most one
    per shot and one per hardware queue beside the calling stream's; one
    batch for the per-step kernels and the exact gradient."""
    from types import SimpleNamespace
    from fastgpr import solver
    G = lambda S=13, tb=(64, 1024, 6), **opt: SimpleNamespace(S=S, time_block=tb, options=opt)
    monkeypatch.delenv("GPU_MAX_HW_QUEUES", raising=False)
    assert solver.max_shot_groups() == 3
    assert solver.shot_groups(G()) == 3                       # default
    assert solver.shot_groups(G(shot_streams=2)) == 2
    assert solver.shot_groups(G(shot_streams=5)) == 3
    assert solver.shot_groups(G(S=2, shot_streams=3)) == 2
    assert solver.shot_groups(G(shot_streams=1)) == 1
    assert solver.shot_groups(G(tb=None, shot_streams=3)) == 1
    assert solver.shot_groups(G(gradient="exact", shot_streams=3)) == 1
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "2")
    assert solver.shot_groups(G(shot_streams=3)) == 1
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "8")
    assert solver.shot_groups(G(shot_streams=5)) == 5


def test_auto_checkpoint_counts_cached_memory_and_keeps_a_reserve(monkeypatch):
    """checkpoint="auto" stores the history when it fits in the free memory
    plus what torch's allocator caches unused, less the reserve; otherwise it
    keeps the largest snapshot ring that fits (device queries stubbed: no GPU
    here)."""
    from types import SimpleNamespace
    import torch
    from fastgpr import _native, checkpoint as ck
    GiB = 1 << 30
    monkeypatch.setattr(torch.cuda, "get_device_properties",
                        lambda d: SimpleNamespace(total_memory=288 * GiB))
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda d: (100 * GiB, 288 * GiB))
    monkeypatch.setattr(torch.cuda, "memory_allocated", lambda d: 10 * GiB)
    G = SimpleNamespace(iterations=1500, S=13, NI=1600, NJ=800, time_block=None, options={},
                        graph_chunk=32, device="cuda:0", pmlthick=20)
    slot = 13 * 1600 * _native.pitch(800) * 4
    full = 1500 * slot
    assert 72 * GiB < full < 100 * GiB
    monkeypatch.setattr(torch.cuda, "memory_reserved", lambda d: 150 * GiB)
    assert ck.history(G, "auto") is None                  # 100 + 140 cached - 14.4 reserve
    monkeypatch.setattr(torch.cuda, "memory_reserved", lambda d: 10 * GiB)
    h = ck.history(G, "auto")                              # 100 - 14.4 < full
    budget = int(100 * GiB - 0.05 * 288 * GiB)
    assert h is not None and not h.full
    assert h.nbytes(slot, ck.state_bytes(G)) <= budget < (h.R + 1) * slot
    assert h.recompute_steps() == h.h < 1500 // 2           # most of the history is kept
    G.NI = 400                                             # a quarter of the device: no query
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda d: (_ for _ in ()).throw(AssertionError))
    assert ck.history(G, "auto") is None


def _simulate_history(hist):
    """Replay a History's schedule on the host: ring contents per slot, the
    saved states, and every adjoint step with the snapshots it reads."""
    from fastgpr.checkpoint import forward_cuts
    nt, R = hist.nt, hist.R
    ring = [None] * R
    saved = {0}
    covered = []
    n = 0
    for n0, n1, save, store in forward_cuts(hist):
        assert n0 == n
        if save:
            saved.add(n0)
        for k in range(n0, n1):
            if store:
                ring[k % R] = k
        n = n1
    assert n == nt

    def adjoint(hi, lo):
        for k in range(hi, lo - 1, -1):
            assert ring[k % R] == k, (k, ring[k % R])
            if k >= 1:
                assert ring[(k - 1) % R] == k - 1
            covered.append(k)
    t = hist.tail()
    if t is not None:
        adjoint(*t)
    for s0, e in reversed(hist.segments()):
        assert s0 in saved and e - s0 <= hist.C
        if s0 >= 1:
            ring[(s0 - 1) % R] = s0 - 1                    # seeded from the saved Ez
        for k in range(s0, e):
            ring[k % R] = k
        adjoint(e - 1, max(s0, hist.lowest))
    assert covered == list(range(nt - 1, hist.lowest - 1, -1))
    assert sorted(saved - {0}) == hist.saves()


@pytest.mark.parametrize("nt,R,lowest", [(157, 38, 1), (157, 158, 1), (157, 157, 1), (100, 2, 1),
                                         (100, 3, 0), (1200, 65, 0), (8000, 97, 0), (6000, 4321, 1),
                                         (50, 50, 0), (50, 51, 0), (2, 2, 1), (1, 2, 0)])
def test_history_schedule_reads_the_right_snapshots(nt, R, lowest):
    """Every adjoint step runs once, in descending order, with its two
    snapshots in the ring; recompute segments restart from saved states."""
    from fastgpr.checkpoint import History
    hist = History(nt, R, lowest)
    _simulate_history(hist)
    assert hist.recompute_steps() <= max(nt - R + 1, 0)


def test_history_plan_takes_the_largest_ring_that_fits():
    from fastgpr.checkpoint import plan
    slot, state = 1000, 3300
    for nt in (10, 157, 3000):
        assert plan(nt, slot, state, nt * slot).full
        for budget in (nt * slot - 1, nt * slot // 2, nt * slot // 7, 40_000):
            try:
                h = plan(nt, slot, state, budget)
            except MemoryError:
                from fastgpr.checkpoint import History
                assert all(History(nt, R).nbytes(slot, state) > budget for R in range(2, nt + 1))
                continue
            assert h.nbytes(slot, state) <= budget
            if h.R < nt:
                from fastgpr.checkpoint import History
                bigger = History(nt, h.R + 1)
                assert bigger.nbytes(slot, state) > budget
            _simulate_history(h)
    with pytest.raises(MemoryError):
        plan(3000, slot, state, 3 * slot)


def test_envelope_workloads_are_well_formed():
    """The autotuner envelope's cross-hole grids (bench.py --envelope): every
    source and receiver outside the CPML, one source per shot, distinct
    receivers, the requested step count, and the envelope list itself."""
    import bench
    from fastgpr.workloads import crosshole
    assert len(bench.ENVELOPE) == 15 and (1600, 800, 32, 1500) not in bench.ENVELOPE
    for NI, NJ, S, nt in bench.ENVELOPE:
        wl = crosshole(NI, NJ, S, nt)
        assert (wl.NI, wl.NJ, wl.S, wl.nt, wl.P) == (NI, NJ, S, nt, 20)
        assert g.n_iterations(wl.time_window, wl.dt) == nt
        for cells in (wl.src, wl.rx):
            i, j = cells[..., 0], cells[..., 1]
            assert (i >= wl.P).all() and (i < NI - wl.P).all()
            assert (j >= wl.P).all() and (j < NJ - wl.P).all()
            assert (cells[..., 2] == 0).all()
        assert wl.src.shape == (S, 1, 3) and wl.rx.shape == (S, 80, 3)
        assert len({tuple(c) for c in wl.rx[0]}) == 80
        assert np.isfinite(wl.er).all() and (wl.er >= 1).all() and (wl.se > 0).all()


def test_trace_spans_nest_and_stay_off_by_default(monkeypatch):
    """fastgpr.trace: off unless FASTGPR_ROCTX=1; on, the ranges push and pop
    in pairs (roctx through torch.cuda.nvtx) and an exception still pops."""
    from fastgpr import trace
    calls = []
    monkeypatch.setattr(torch.cuda.nvtx, "range_push", lambda n: calls.append(("push", n)))
    monkeypatch.setattr(torch.cuda.nvtx, "range_pop", lambda: calls.append(("pop",)))
    monkeypatch.setattr(trace, "ENABLED", False)
    with trace.span("off"):
        pass
    assert calls == []
    monkeypatch.setattr(trace, "ENABLED", True)
    with trace.span("outer"):
        with trace.span("inner"):
            pass
    with pytest.raises(RuntimeError):
        with trace.span("fails"):
            raise RuntimeError("x")
    assert calls == [("push", "outer"), ("push", "inner"), ("pop",), ("pop",),
                     ("push", "fails"), ("pop",)]
