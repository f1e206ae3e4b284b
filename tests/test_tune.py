"""Autotuner cache logic (CPU): deterministic for a fixed cache."""
import json

import pytest

from psktm.tune import CANDIDATES_2D, CANDIDATES_3D, Autotuner, signature


def test_picks_fastest_and_skips_unsupported(tmp_path):
    times = {(4, 4, 4): 2.0, (4, 4, 2): 1.5, (4, 4, 8): None, (8, 2, 4): 1.7}
    t = Autotuner(str(tmp_path / "c.json"))
    key = signature("batch", ns=1000, nt=1000, dtype="f16")
    assert t.choose(key, times, lambda c: times[c]) == (4, 4, 2)
    saved = json.load(open(tmp_path / "c.json"))
    assert saved[key]["choice"] == [4, 4, 2] and saved[key]["timings"]["4,4,8"] is None


def test_cache_hit_is_deterministic_and_does_not_time(tmp_path):
    path = str(tmp_path / "c.json")
    key = signature("single", ns=1500, nt=1500)
    Autotuner(path).choose(key, CANDIDATES_3D, lambda c: float(sum(c)))
    calls = []
    again = Autotuner(path).choose(key, CANDIDATES_3D, lambda c: calls.append(c) or 0.0)
    assert again == min(CANDIDATES_3D, key=sum) and calls == []
    # retune ignores the cache
    assert Autotuner(path).choose(key, CANDIDATES_3D, lambda c: -c[2], retune=True) == (4, 4, 8)


def test_ties_go_to_first_candidate(tmp_path):
    t = Autotuner(str(tmp_path / "c.json"))
    assert t.choose("k", CANDIDATES_2D, lambda c: 1.0) == CANDIDATES_2D[0]


def test_no_runnable_candidate_raises(tmp_path):
    with pytest.raises(RuntimeError):
        Autotuner(str(tmp_path / "c.json")).choose("k", CANDIDATES_2D, lambda c: None)


def test_candidates_are_valid_tile_sizes():
    for c in list(CANDIDATES_3D) + list(CANDIDATES_2D):
        assert c[0] * c[1] in (8, 16) and 1 <= c[2] <= 8


def test_refine_times_the_best_trials_on_the_whole_workload(tmp_path):
    # trials rank (4,4,2) < (8,2,4) < (2,8,4) < (4,4,4); on the whole workload
    # the third of them is the quickest, the fourth is never timed again
    trial = {(4, 4, 4): 2.0, (4, 4, 2): 1.0, (4, 4, 8): None, (8, 2, 4): 1.1, (2, 8, 4): 1.2}
    whole = {(4, 4, 2): 5.0, (8, 2, 4): None, (2, 8, 4): 4.0}
    refined = []
    t = Autotuner(str(tmp_path / "c.json"))
    got = t.choose("k", trial, lambda c: trial[c],
                   refine_fn=lambda c: refined.append(c) or whole[c], refine_top=3)
    assert got == (2, 8, 4)
    assert refined == [(4, 4, 2), (8, 2, 4), (2, 8, 4)]
    saved = json.load(open(tmp_path / "c.json"))["k"]
    assert saved["choice"] == [2, 8, 4] and saved["seconds"] == 4.0
    assert saved["refined"] == {"4,4,2": 5.0, "8,2,4": None, "2,8,4": 4.0}
    # the cached choice comes back without timing anything
    assert Autotuner(str(tmp_path / "c.json")).choose(
        "k", trial, lambda c: 1 / 0, refine_fn=lambda c: 1 / 0) == (2, 8, 4)
