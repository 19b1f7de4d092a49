"""fib through the C ABI on the GPU, checked against the oracle (bit-exact) and against the
task-count invariant of the fork-join tree (every task runs exactly once)."""
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    gtap_amd.gtap_set_config(timeout_ms=15000)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0


def executions(n):
    # leaves run once, internal nodes run twice (spawn state + resume state)
    return 3 * oracle.fib(n + 1) - 2


@pytest.mark.parametrize("n", list(range(0, 24)) + [27, 30])
def test_fib_parity(n):
    print(f"case fib({n})", flush=True)
    v, st = gtap_amd.fib(n)
    assert v == oracle.fib(n)
    assert st["tasks_executed"] == executions(n)
    assert st["tasks_spawned"] == 2 * oracle.fib(n + 1) - 2
    assert st["jobs_taken"] == 1
    assert st["error_code"] == 0


def test_fib_repeatable_under_nondeterministic_schedules():
    vals = set()
    for _ in range(5):
        v, st = gtap_amd.fib(28)
        vals.add(v)
    assert vals == {oracle.fib(28)}


def test_invalid_argument_is_rejected_without_launch():
    with pytest.raises(gtap_amd.GtapError):
        gtap_amd.fib(-1)
