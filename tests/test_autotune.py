"""Autotuner logic (CPU): candidate order, the results gate, the cache."""
import numpy as np

from fuller_md_en_amd import autotune


def _fake_run(costs, wrong=()):
    def run(p):
        key = tuple(sorted(p.items()))
        pos = np.arange(6.0)
        if key in wrong:
            pos = pos + 1e-3
        return {"seconds": costs[key], "pos": pos, "ep": np.ones(4), "overflow": 0}
    return run


def test_picks_fastest_valid_candidate_and_caches(tmp_path):
    space = {"a": [1, 2], "b": [0, 1]}
    combos = [tuple(sorted(c.items())) for c in autotune.candidates(space)]
    assert combos[0] == (("a", 1), ("b", 0))  # the default comes first
    costs = dict(zip(combos, [1.0, 0.8, 0.5, 0.9]))
    wrong = {combos[2]}  # fastest, but changes the results: rejected
    path = str(tmp_path / "tune.json")
    e = autotune.tune("k", _fake_run(costs, wrong), space, path=path)
    assert e["params"] == {"a": 1, "b": 1} and e["seconds"] == 0.8
    assert e["default_seconds"] == 1.0
    assert [t["valid"] for t in e["tried"]] == [True, True, False, True]
    # cached: the runner is not called again
    e2 = autotune.tune("k", lambda p: 1 / 0, space, path=path)
    assert e2["params"] == e["params"]
    assert "k" in autotune.load_cache(path)


def test_overflow_rejects_candidate(tmp_path):
    space = {"a": [1, 2]}

    def run(p):
        return {"seconds": 1.0 / p["a"], "pos": np.zeros(3), "ep": np.zeros(2),
                "overflow": 1 if p["a"] == 2 else 0}
    e = autotune.tune("k2", run, space, path=str(tmp_path / "t.json"))
    assert e["params"] == {"a": 1}


def test_same_results_tolerance():
    a = {"pos": np.array([1.0, 2.0]), "ep": np.array([-3.0])}
    b = {"pos": np.array([1.0, 2.0 + 1e-9]), "ep": np.array([-3.0])}
    assert autotune.same_results(a, b, 1e-7)
    b["ep"] = np.array([-3.1])
    assert not autotune.same_results(a, b, 1e-7)


def test_stages_fix_earlier_choices(tmp_path):
    """A list of spaces is tuned stage by stage: the second stage's axes are
    swept with the first stage's choice fixed, and every candidate is judged
    against the all-default run."""
    space = [{"a": [1, 2]}, {"b": [0, 1]}]
    assert autotune.defaults(space) == {"a": 1, "b": 0}
    costs = {(("a", 1), ("b", 0)): 1.0, (("a", 2), ("b", 0)): 0.7,
             (("a", 2), ("b", 1)): 0.6, (("a", 1), ("b", 1)): 0.1}
    calls = []

    def run(p):
        calls.append(dict(p))
        return _fake_run(costs)(p)
    e = autotune.tune("k3", run, space, path=str(tmp_path / "t.json"))
    # (a=1, b=1) is never tried: stage 2 runs with a = 2 fixed
    assert calls == [{"a": 1, "b": 0}, {"a": 2, "b": 0}, {"a": 2, "b": 0}, {"a": 2, "b": 1}]
    assert e["params"] == {"a": 2, "b": 1} and e["default_seconds"] == 1.0
