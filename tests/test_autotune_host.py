"""Host logic of the time-blocked autotuner (fastgpr.autotune.choose): the
fastest supported variant wins, variants the library does not compile are
skipped, and winners are cached per (device, grid, P, S, receivers, shot
groups, candidate set) in the process and in the FASTGPR_TUNE_CACHE JSON
file.  The
timing itself runs on the GPU; here it is replaced by a table."""
import json
from types import SimpleNamespace

import pytest
import torch

from fastgpr import autotune


FAKE_US = {(64, 1024, 6): 11.3, (64, 1024, 8): 12.3, (32, 256, 4): 20.0}
CANDIDATES = [(32, 256, 4), (48, 256, 6), (64, 1024, 8), (64, 1024, 6)]   # (48, 256, 6): not compiled
CAND = "|32,256,4;48,256,6;64,1024,6;64,1024,8"                           # the set, sorted, in the key


@pytest.fixture
def tuner(monkeypatch, tmp_path):
    calls = []

    def fake_time(G, er, se, mr, tb, reps=96):
        calls.append((G.options["shot_streams"], tuple(tb)))
        return FAKE_US[tuple(tb)] * 1e-6

    path = tmp_path / "tune.json"
    monkeypatch.setattr(autotune, "time_variant", fake_time)
    monkeypatch.setattr(torch.cuda, "get_device_name", lambda dev=None: "gfx950-test")
    monkeypatch.setenv("FASTGPR_TUNE_CACHE", str(path))
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "4")
    monkeypatch.setattr(autotune, "_cache", {})
    return calls, path


def grid(shot_streams=3, S=13):
    return SimpleNamespace(device="cuda:0", NI=400, NJ=200, pmlthick=20, S=S, nrx=80,
                           options={"shot_streams": shot_streams})


def test_fastest_supported_variant_wins_and_is_cached(tuner):
    calls, path = tuner
    G = grid()
    assert autotune.choose(G, None, None, None, candidates=CANDIDATES) == (64, 1024, 6)
    # the variant the library does not compile is never timed
    assert sorted(tb for _, tb in calls) == sorted(FAKE_US)
    key = "gfx950-test|400x200|P20|S13|rx80|g3" + CAND
    assert json.loads(path.read_text()) == {key: [64, 1024, 6]}

    n = len(calls)
    assert autotune.choose(G, None, None, None, candidates=CANDIDATES) == (64, 1024, 6)
    assert len(calls) == n                       # in-process cache: no timing

    autotune._cache.clear()                      # a new process: the file answers
    assert autotune.choose(G, None, None, None, candidates=CANDIDATES) == (64, 1024, 6)
    assert len(calls) == n


def test_shot_groups_are_part_of_the_key(tuner):
    calls, path = tuner
    autotune.choose(grid(shot_streams=3), None, None, None, candidates=CANDIDATES)
    autotune.choose(grid(shot_streams=1), None, None, None, candidates=CANDIDATES)
    # two shots allow two groups at most, whatever was asked for
    autotune.choose(grid(shot_streams=3, S=2), None, None, None, candidates=CANDIDATES)
    assert [g for g, _ in calls] == [3] * 3 + [1] * 3 + [3] * 3
    keys = sorted(json.loads(path.read_text()))
    assert keys == ["gfx950-test|400x200|P20|S13|rx80|g1" + CAND,
                    "gfx950-test|400x200|P20|S13|rx80|g3" + CAND,
                    "gfx950-test|400x200|P20|S2|rx80|g2" + CAND]


def test_candidate_set_is_part_of_the_key(tuner):
    """A winner among a narrower set (the multi-problem variants) is not
    taken for the answer over all variants, nor the other way round."""
    calls, path = tuner
    G = grid()
    assert autotune.choose(G, None, None, None, candidates=[(64, 1024, 8), (32, 256, 4)]) == (64, 1024, 8)
    assert autotune.choose(G, None, None, None, candidates=CANDIDATES) == (64, 1024, 6)
    assert len(calls) == 2 + 3
    assert set(json.loads(path.read_text())) == {
        "gfx950-test|400x200|P20|S13|rx80|g3|32,256,4;64,1024,8",
        "gfx950-test|400x200|P20|S13|rx80|g3" + CAND}


def test_hardware_queue_limit_caps_the_groups(tuner, monkeypatch):
    calls, path = tuner
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "2")     # one side stream with its own queue
    autotune.choose(grid(shot_streams=3), None, None, None, candidates=CANDIDATES)
    assert list(json.loads(path.read_text())) == ["gfx950-test|400x200|P20|S13|rx80|g1" + CAND]


def test_timing_reps_shrink_with_the_block_and_stay_bounded():
    sizes = [(40, 30, 2), (400, 200, 13), (500, 250, 22), (800, 400, 11), (1600, 800, 32), (4000, 4000, 64)]
    reps = [autotune.timing_reps(SimpleNamespace(NI=a, NJ=b, S=s)) for a, b, s in sizes]
    assert reps[0] == reps[1] == 96                  # the small grids keep the full count
    assert all(x >= y for x, y in zip(reps, reps[1:]))
    assert reps[-1] == 12 and min(reps) == 12
