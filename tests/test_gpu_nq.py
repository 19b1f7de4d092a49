"""n-queens through the C ABI on the GPU against the oracle: bit-exact counts and the
exact number of task executions (search-tree nodes + one resume per internal node)."""
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


@pytest.mark.parametrize("n", list(range(1, 13)) + [14])
def test_nqueens_parity(n):
    print(f"case nqueens({n})", flush=True)
    v, st = gtap_amd.nqueens(n)
    sols, tasks = oracle.nq(n)
    assert v == sols
    assert st["tasks_executed"] == tasks + oracle.nq_internal(n)
    assert st["tasks_spawned"] == tasks - 1
    assert st["error_code"] == 0


@pytest.mark.parametrize("n", [1, 4, 8, 10, 12])
def test_nqueens_slot_join_parity(n):
    """The result-slot lowering: same counts and task executions as the additive join."""
    print(f"case nqueens_slots({n})", flush=True)
    v, st = gtap_amd.nqueens_slots(n)
    sols, tasks = oracle.nq(n)
    assert v == sols
    assert st["tasks_executed"] == tasks + oracle.nq_internal(n)
    assert st["error_code"] == 0


def test_nqueens_rejects_out_of_range():
    with pytest.raises(gtap_amd.GtapError):
        gtap_amd.nqueens(0)
    with pytest.raises(gtap_amd.GtapError):
        gtap_amd.nqueens(17)
    with pytest.raises(gtap_amd.GtapError):
        gtap_amd.nqueens_slots(17)
