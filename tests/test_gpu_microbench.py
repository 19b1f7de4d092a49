"""The measurement entry points return finite, positive figures (they feed the benchmark's
roofline fractions and the publication-order choice of the scheduler)."""
import math

import pytest

import gtap_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0


def test_hbm_copy_and_atomics():
    print("case bench_hbm_copy/bench_atomics", flush=True)
    g = gtap_amd.bench_hbm_copy(256 << 20, 2)
    a = gtap_amd.bench_atomics(2)
    assert 100.0 < g < 20000.0
    assert a > 1.0 and math.isfinite(a)


@pytest.mark.parametrize("waves", [64, 8192])
def test_fences(waves):
    print(f"case bench_fences waves={waves}", flush=True)
    f = gtap_amd.bench_fences(waves, 16)
    for k in ("release_fence_ns", "acquire_fence_ns", "store_wait_ns"):
        assert f[k] > 0.0 and math.isfinite(f[k])


def test_fences_rejects_bad_sizes():
    with pytest.raises(Exception):
        gtap_amd.bench_fences(6, 16)


@pytest.mark.parametrize("words", [1, 64, 4096])
def test_contended_atomics(words):
    print(f"case bench_atomics_contended words={words}", flush=True)
    g = gtap_amd.bench_atomics_contended(words, 2)
    assert g > 0.0 and math.isfinite(g)


@pytest.mark.parametrize("table", [1 << 20, 16 << 20])
def test_random_access(table):
    print(f"case bench_random_access table={table}", flush=True)
    r = gtap_amd.bench_random_access(table, 2)
    assert r["load_gps"] > 0.1 and math.isfinite(r["load_gps"])
    assert r["fetch_min_gps"] > 0.1 and math.isfinite(r["fetch_min_gps"])
