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
