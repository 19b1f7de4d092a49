"""n-queens at the reference's benchmark settings (evaluation/2-comparison/nq/gtap_nq.cu:10-99:
tasks to row `cutoff`, sequential search below, no taskwait, one counter) gives the
solution counts of the oracle for every cutoff, including cutoffs at or beyond N."""
import pytest

import gtap_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    G.gtap_set_config(timeout_ms=30000)
    assert G.gtap_initialize() == 0
    yield
    assert G.gtap_finalize() == 0


@pytest.mark.parametrize("n", [1, 4, 6, 8, 10, 12, 13])
def test_cutoff_counts(n):
    for cutoff in (0, 3, 7, n - 1, n, n + 2):
        if cutoff < 0:
            continue
        print(f"case nq cutoff n={n} cutoff={cutoff}", flush=True)
        v, s = G.nqueens_cutoff(n, cutoff)
        assert v == O.nq(n)[0]


def test_cutoff_nq14_reference_setting():
    print("case nq cutoff n=14 cutoff=7", flush=True)
    v, s = G.nqueens_cutoff(14, 7)
    assert v == 365596
