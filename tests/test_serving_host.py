"""Serving admission and failure handling on the host (no GPU): a CPU arena,
a stand-in scheduler that advances jobs one block per tick, and stand-in
jobs.  Covers exact admission by arena blocks, adaptive snapshot histories,
holds and their age bound, per-request failures, cancellation and the
finiteness guard."""
import numpy as np
import pytest
import torch

from fastgpr import batching
from fastgpr.serving import Request, Server, serving_request


class FakeEvent:
    def query(self):
        return True


class FakeJob:
    def __init__(self, req, ticks=3, finite=True):
        self.req, self.left, self._finite = req, ticks, finite
        self.done_event, self.cancelled, self.stream_idx = None, False, 0

    @property
    def done(self):
        return self.left <= 0

    def cancel(self):
        self.cancelled, self.left = True, 0

    def finite(self):
        return self._finite

    def results(self):
        return {"traces": self.req.rid}

    def cell_updates(self):
        return 1


class FakeScheduler:
    def __init__(self):
        self.jobs, self.launches, self.streams = [], 0, [None]

    def add(self, job):
        self.jobs.append(job)

    def active(self):
        return [j for j in self.jobs if not j.done]

    def tick(self):
        for j in self.active():
            j.left -= 1
            if j.done:
                j.done_event = FakeEvent()
        self.launches += 1

    def tick_events(self):
        return []

    def mark_done(self, job):
        job.done_event = FakeEvent()

    def join(self):
        pass


def _server(arena_bytes, ticks=None, bad=(), nonfinite=(), max_hold_s=2.0, histories="share"):
    arena = batching.Arena("cpu", arena_bytes)
    srv = Server("cpu", scheduler=FakeScheduler(), arena=arena, max_hold_s=max_hold_s,
                 histories=histories)
    built = []

    def make(req, tb, hist, region):
        if req.rid in bad:
            raise RuntimeError("job construction failed")
        region((max(1, region.end - region.base - 1024),), torch.uint8)   # the block is usable
        built.append((req.rid, hist))
        return FakeJob(req, (ticks or {}).get(req.rid, 3), req.rid not in nonfinite)
    srv._make_job = make
    return srv, built


def _reqs(n, seed=3, **kw):
    rng = np.random.default_rng(seed)
    kw.setdefault("grids", ((64, 48), (96, 64)))
    kw.setdefault("shots", (1, 4))
    kw.setdefault("steps", (100, 200))
    return [serving_request(i, rng, P=8, **kw) for i in range(n)]


def _bytes(r, R=None):
    w = r.wl
    return batching.job_bytes(w.NI, w.NJ, w.S, w.nt, w.rx.shape[1], w.P, R or w.nt, True)


def test_bad_requests_fail_alone():
    """An unsupported variant, a job that cannot be built and a non-finite
    result each fail their own request; the others complete, and the arena
    is whole again at the end."""
    reqs = _reqs(6)
    reqs[1].tb = (48, 999, 3)
    srv, _ = _server(1 << 30, bad={3}, nonfinite={4})
    for r in reqs:
        srv.submit(r)
    st = srv.run()
    assert st["completed"] == 3 and st["refused"] == 3
    for k in (1, 3, 4):
        assert "error" in reqs[k].result
    for k in (0, 2, 5):
        assert reqs[k].result == {"traces": k}
    assert not srv.waiting and not srv.active
    assert srv.arena.free_bytes() == srv.arena.size


def test_admission_is_exact_and_histories_adapt():
    """With room for one whole history at a time, a second request is
    admitted beside the first with a shorter snapshot ring (its recompute
    counted) instead of waiting."""
    reqs = _reqs(2, grids=((96, 64),), shots=(4, 4), steps=(2000, 2000))
    full = _bytes(reqs[0])
    srv, built = _server(int(1.5 * full))
    for r in reqs:
        srv.submit(r)
    srv.step()
    assert len(srv.active) == 2 and srv.in_use <= srv.arena.size
    assert built[0][1].full and not built[1][1].full
    assert srv.stats.partial_histories == 1
    st = srv.run()
    assert st["completed"] == 2 and st["recompute_share"] > 0


def test_whole_histories_wait_for_memory_instead_of_recomputing():
    """histories="whole": the second gradient, whose whole history fits the
    arena but not beside the first, waits for the first to finish and then
    keeps its whole history -- no ring, no recompute."""
    reqs = _reqs(2, grids=((96, 64),), shots=(4, 4), steps=(2000, 2000))
    full = _bytes(reqs[0])
    srv, built = _server(int(1.5 * full), ticks={0: 3, 1: 3}, histories="whole")
    for r in reqs:
        srv.submit(r)
    srv.step()
    assert len(srv.active) == 1 and srv.stats.history_waits >= 1
    st = srv.run()
    assert st["completed"] == 2 and st["recompute_share"] == 0
    assert [h.full for _, h in built] == [True, True]
    assert st["history_waits"] >= 1 and st["partial_histories"] == 0


def test_whole_histories_take_a_ring_after_the_age_bound_or_when_too_large():
    """histories="whole": a gradient that waited longer than max_hold_s is
    admitted with a ring beside the running one; one whose whole history
    exceeds the arena never waits for it."""
    reqs = _reqs(2, grids=((96, 64),), shots=(4, 4), steps=(2000, 2000))
    full = _bytes(reqs[0])
    srv, built = _server(int(1.5 * full), ticks={0: 50, 1: 3}, histories="whole",
                         max_hold_s=0.0)
    for r in reqs:
        srv.submit(r)
    srv.step()
    assert len(srv.active) == 2 and built[0][1].full and not built[1][1].full
    big = _reqs(1, grids=((96, 64),), shots=(4, 4), steps=(2000, 2000))[0]
    srv, built = _server(int(0.6 * _bytes(big)), histories="whole")
    srv.submit(big)
    st = srv.run()
    assert st["completed"] == 1 and not built[0][1].full and st["history_waits"] == 0
    with pytest.raises(ValueError):
        Server("cpu", scheduler=FakeScheduler(), arena=batching.Arena("cpu", 1 << 20),
               histories="most")


def test_strict_histories_hold_back_younger_requests_for_a_whole_history():
    """histories="strict": a gradient that waited past max_hold_s for its
    whole history does not take a ring; younger requests wait behind it
    until it fits, and nothing is recomputed."""
    big = _reqs(2, grids=((96, 64),), shots=(4, 4), steps=(2000, 2000))
    small = _reqs(1, seed=9, grids=((64, 48),), shots=(1, 1), steps=(100, 100))[0]
    small.rid = 2
    full = _bytes(big[0])
    srv, built = _server(int(1.5 * full), ticks={0: 4, 1: 3, 2: 1}, histories="strict",
                         max_hold_s=0.0)
    srv.submit(big[0])
    srv.submit(big[1])
    srv.step()
    assert [r.rid for r, _, _, _ in srv.active] == [0]
    srv.submit(small)
    srv.step()                                   # big[1] is past its age bound: small waits
    assert [r.rid for r, _, _, _ in srv.active] == [0] and len(srv.waiting) == 2
    st = srv.run()
    assert st["completed"] == 3 and st["recompute_share"] == 0
    assert [rid for rid, _ in built] == [0, 1, 2] and all(h.full for _, h in built[:2])


def test_strict_sizes_an_oversize_ring_to_the_largest_free_block():
    """A gradient whose whole history exceeds the arena, admitted while
    another runs: "strict" sizes its ring to the largest free block, "whole"
    to half of it (more recompute)."""
    small = _reqs(1, grids=((64, 48),), shots=(1, 1), steps=(100, 100))[0]
    big = _reqs(1, seed=5, grids=((96, 64),), shots=(4, 4), steps=(2000, 2000))[0]
    big.rid = 1
    arena = int(0.8 * _bytes(big))
    used = {}
    for rule in ("strict", "whole"):
        srv, built = _server(arena, ticks={0: 50, 1: 3}, histories=rule)
        srv.submit(small)
        srv.step()
        free = srv.arena.largest_free()
        srv.submit(big)
        srv.step()
        (rid, hist), = [b for b in built if b[0] == 1]
        assert not hist.full
        used[rule] = _bytes(big, hist.R)
        assert used[rule] <= free
        small.result = small.done_at = None
    assert used["strict"] > free // 2 >= used["whole"] * 0.99


def test_held_request_blocks_younger_ones_after_its_age_bound():
    """A request whose smallest footprint does not fit waits (a hold); while
    it is young smaller ones pass it, once it is older than max_hold_s
    nothing younger is admitted before it."""
    big = _reqs(1, seed=5, grids=((200, 160),), shots=(8, 8), steps=(300, 300))[0]
    small = _reqs(3, seed=6, grids=((40, 40),), shots=(1, 1), steps=(50, 50))
    for k, r in enumerate(small):
        r.rid = 10 + k
    from fastgpr.serving import _smallest_history
    w = big.wl
    spec = (w.NI, w.NJ, w.S, w.nt, int(w.rx.shape[1]), w.P)
    least = batching.job_bytes(*spec, _smallest_history(spec, (64, 1024, 6)), True)
    arena = least + _bytes(small[0]) // 2      # big fits only when no small one runs
    assert _bytes(small[0]) * 3 < least
    for max_hold, expect_small_first in ((60.0, True), (0.0, False)):
        srv, built = _server(arena, ticks={10: 50}, max_hold_s=max_hold)
        srv.submit(small[0])
        srv.step()                         # small[0] runs for 50 ticks and holds memory
        srv.submit(big)
        srv.submit(small[1])
        srv.submit(small[2])
        srv.step()
        order = [rid for rid, _ in built]
        assert srv.stats.holds >= 1 and big.rid not in order
        assert (11 in order) == expect_small_first, (max_hold, order)
        srv.run()
        order = [rid for rid, _ in built]
        assert all(r.result is not None and "error" not in r.result for r in [big] + small)
        if not expect_small_first:
            assert order.index(big.rid) < order.index(11)


def test_cancel_waiting_and_running():
    reqs = _reqs(3)
    srv, _ = _server(1 << 30, ticks={0: 100})
    for r in reqs[:2]:
        srv.submit(r)
    srv.step()
    srv.submit(reqs[2])
    srv.cancel(0)                          # running
    srv.cancel(2)                          # still waiting
    st = srv.run()
    assert reqs[0].result == {"error": "cancelled"} and reqs[2].result == {"error": "cancelled"}
    assert reqs[1].result == {"traces": 1}
    assert st["cancelled"] == 2 and st["completed"] == 1
    assert srv.arena.free_bytes() == srv.arena.size


def test_request_larger_than_the_arena_is_refused():
    r = _reqs(1)[0]
    srv, _ = _server(1 << 16)
    srv.submit(r)
    st = srv.run()
    assert "error" in r.result and st["refused"] == 1


def test_arena_best_fit_and_merging():
    """Random alloc / release against a brute-force model of the byte range:
    blocks never overlap, best fit takes the smallest free block that holds
    the request, and released neighbours merge."""
    rng = np.random.default_rng(0)
    A = batching.Arena("cpu", 1 << 20)
    live = {}
    for _ in range(3000):
        if live and rng.random() < 0.45:
            off = list(live)[rng.integers(len(live))]
            A.release(off)
            del live[off]
        else:
            n = int(rng.integers(1, 40000))
            frees = [s for _, s in A.free_list if s >= batching._aligned(n)]
            off = A.alloc(n)
            if off is None:
                assert not frees
                continue
            assert A.used[off] == batching._aligned(n)
            live[off] = A.used[off]
        spans = sorted([(o, o + s) for o, s in live.items()] + [(o, o + s) for o, s in A.free_list])
        assert spans[0][0] == 0 and spans[-1][1] == A.size
        for a, b in zip(spans, spans[1:]):
            assert a[1] == b[0]                      # a partition of the range
        for a, b in zip(A.free_list, A.free_list[1:]):
            assert a[0] + a[1] < b[0]                # free neighbours are merged
    for off in list(live):
        A.release(off)
    assert A.free_list == [(0, A.size)]


def test_region_views_are_aligned_and_bounded():
    A = batching.Arena("cpu", 1 << 16)
    off = A.alloc(5000)
    reg = A.region(off)
    a = reg((3, 7), torch.float32)
    b = reg((10,), torch.float64)
    base = A.buf.data_ptr()
    assert (a.data_ptr() - base) % 256 == 0 and (b.data_ptr() - base) % 256 == 0
    assert a.shape == (3, 7)
    assert b.data_ptr() - a.data_ptr() == 256
    with pytest.raises(MemoryError):
        reg((5000,), torch.uint8)
