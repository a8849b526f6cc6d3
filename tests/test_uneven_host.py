"""Host-side planning of uneven batches (no GPU): waves, histories, shot
cuts and the shared buffers of split models."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "fast-gpr-fwi_amd"))

from fastgpr.batching import ALIGN, ShotGroup, job_bytes  # noqa: E402
from fastgpr.uneven import plan_waves  # noqa: E402

TB = (64, 1024, 6)


def _spec(NI, NJ, S, nt):
    return (NI, NJ, S, nt, 10, 20, TB)


def _groups(NI, NJ, S, nt, n=3):
    cuts = [round(i * S / n) for i in range(n + 1)]
    return [_spec(NI, NJ, cuts[i + 1] - cuts[i], nt) for i in range(n)]


def test_owners_keep_each_model_in_waves_of_its_own():
    specs = _groups(300, 150, 16, 400) + _groups(200, 100, 7, 300)
    owners = [0, 0, 0, 1, 1, 1]
    big = 10 * sum(job_bytes(*sp[:6], sp[3], True, sp[6]) for sp in specs)
    packed, _ = plan_waves(specs, big)
    assert len(packed) == 1 and sorted(packed[0]) == list(range(6))
    waves, hist = plan_waves(specs, big, owners)
    assert len(waves) == 2
    for w in waves:
        assert len({owners[k] for k in w}) == 1
    assert sorted(k for w in waves for k in w) == list(range(6))
    assert all(h.full for h in hist)


def test_jobs_fill_waves_first_fit_within_budget():
    specs = _groups(300, 150, 16, 400)
    one = max(job_bytes(*sp[:6], sp[3], True, sp[6]) for sp in specs)
    waves, hist = plan_waves(specs, int(2.2 * one), [0, 0, 0])
    assert [len(w) for w in waves] == [2, 1]
    for w in waves:
        assert sum(job_bytes(*specs[k][:6], specs[k][3], True, TB) for k in w) <= int(2.2 * one)
    assert all(h.full for h in hist)


def test_job_too_large_alone_gets_a_ring_in_its_own_wave():
    specs = [_spec(400, 200, 8, 1000), _spec(100, 50, 4, 200)]
    full = job_bytes(*specs[0][:6], specs[0][3], True, TB)
    budget = int(0.6 * full)
    waves, hist = plan_waves(specs, budget)
    assert waves[0] == [0] and not hist[0].full and hist[0].R < 1000
    assert job_bytes(*specs[0][:6], hist[0], True, TB) <= budget
    assert hist[1].full and [1] in waves


def test_shot_cuts_cover_every_shot_once():
    for S in (1, 2, 3, 4, 5, 11, 15, 16, 22, 64):
        n = max(1, min(3, S))
        cuts = [round(i * S / n) for i in range(n + 1)]
        assert cuts[0] == 0 and cuts[-1] == S
        assert all(cuts[i + 1] > cuts[i] for i in range(n)), (S, cuts)


def test_shot_group_bytes_hold_its_buffers():
    class G:
        S, iterations, NI, NJ = 15, 6000, 1200, 600

        class receiver:
            shape = (2, 119)
    need = (15 * 6000 * 119 * 4 + 15 * 2 * 1200 * 600 * 8 + 2 * 1200 * 600 * 8
            + 2 * 1200 * 600 * 4)
    got = ShotGroup.bytes(G)
    assert need <= got <= need + 6 * ALIGN
    assert ShotGroup.bytes(G, want=("er",)) == got - ((1200 * 600 * 4 + ALIGN - 1) // ALIGN) * ALIGN
