"""The autotuner's search (host logic, no GPU): coordinate descent finds the minimum of a
separable cost, re-measures nothing, and rejects a configuration that changes the result."""
import pytest

from gtap_amd import autotune


def test_finds_separable_minimum():
    calls = []

    def measure(p):
        calls.append(dict(p))
        return (p["a"] - 3) ** 2 + (p["b"] - 20) ** 2 + 1.0, "same"

    space = {"a": (0, 1, 2, 3, 4), "b": (10, 20, 30)}
    best, table = autotune.coordinate_descent(measure, space)
    assert best == {"a": 3, "b": 20}
    keys = [tuple(sorted(c.items())) for c in calls]
    assert len(keys) == len(set(keys))  # every point measured once
    assert len(table) == len(calls)
    assert table[0]["params"] == {"a": 0, "b": 10}


def test_start_point_and_no_improvement():
    space = {"a": (5, 6), "b": (1, 2)}
    best, table = autotune.coordinate_descent(lambda p: (1.0, 0), space, start={"a": 6, "b": 2})
    assert best == {"a": 6, "b": 2}  # ties keep the incumbent
    assert table[0]["params"] == {"a": 6, "b": 2}


def test_result_change_is_an_error():
    def measure(p):
        return 1.0, p["a"]  # the "result" depends on the configuration

    with pytest.raises(RuntimeError, match="result changed"):
        autotune.coordinate_descent(measure, {"a": (0, 1)})


def test_unknown_workload():
    with pytest.raises(ValueError):
        autotune.workload("nosuch")


def test_tuned_table_lookup(tmp_path, monkeypatch):
    """tuned() reads the committed table: the best point of a workload, {} for an unknown
    workload or a missing file."""
    import json
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"nq14": {"best": {"blocks": 512, "probe_lanes": 4}, "best_ms": 1.0}}))
    monkeypatch.setattr(autotune, "TUNED_PATH", str(p))
    assert autotune.tuned("nq14") == {"blocks": 512, "probe_lanes": 4}
    assert autotune.tuned("fib30") == {}
    monkeypatch.setattr(autotune, "TUNED_PATH", str(tmp_path / "absent.json"))
    assert autotune.tuned("nq14") == {}


def test_committed_table_is_well_formed():
    """Every committed entry names only searchable coordinates with integer values."""
    import json
    import os
    if not os.path.exists(autotune.TUNED_PATH):
        pytest.skip("no tuned table committed")
    table = json.load(open(autotune.TUNED_PATH))
    known = set(autotune.CONFIG_KEYS) | {"split", "cutoff", "merge_leaf", "leaf"}
    for w, entry in table.items():
        assert set(entry["best"]) <= known, w
        assert all(isinstance(v, int) for v in entry["best"].values()), w
        assert entry["best_ms"] <= entry["default_ms"] * 1.0001, w
