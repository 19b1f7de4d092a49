"""Block-cooperative cilksort through the C ABI on the GPU: the sorted keys are bit-exact
against an independent sort (numpy) and the oracle's cilksort restatement."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    gtap_amd.gtap_set_config(timeout_ms=20000)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0


@pytest.mark.parametrize("n", [0, 1, 2, 3, 100, 4095, 4096, 4097, 16384, 16385, 65537, 262147, 1000003])
def test_cilksort_sizes(n):
    print(f"case cilksort n={n}", flush=True)
    keys = oracle.gen_keys_u32(n, seed=1000 + n)
    out, st = gtap_amd.cilksort(keys)
    assert np.array_equal(out, np.sort(keys))
    assert st["error_code"] == 0


@pytest.mark.parametrize("leaf,merge_leaf", [(64, 256), (4096, 4096), (1, 3), (4096, 1 << 20)])
def test_cilksort_cutoffs(leaf, merge_leaf):
    print(f"case cilksort leaf={leaf} merge_leaf={merge_leaf}", flush=True)
    rng = np.random.default_rng(leaf * 7 + merge_leaf)
    for keys in (rng.integers(0, 2**32, 200003, dtype=np.uint64).astype(np.uint32),
                 rng.integers(0, 5, 100000).astype(np.uint32),
                 np.arange(70000, dtype=np.uint32)[::-1].copy()):
        out, _ = gtap_amd.cilksort(keys, leaf=leaf, merge_leaf=merge_leaf)
        assert np.array_equal(out, np.sort(keys))
        assert np.array_equal(out, oracle.cilksort_u32(keys, leaf, merge_leaf))


def test_cilksort_benchmark_keys():
    n = 1 << 26
    print(f"case cilksort benchmark n={n}", flush=True)
    keys = gtap_amd.gen_keys(n)  # gen_vector keys, seed 123456789
    out, st = gtap_amd.cilksort(keys)
    assert np.array_equal(out, np.sort(keys))
