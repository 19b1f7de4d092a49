"""The mixed job stream on the GPU: batched (one resident launch), one launch per job, and
served (Poisson arrivals streamed into the resident kernel). Every job's result is checked
against the oracle; each job's result alone equals its result in the mix."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    gtap_amd.gtap_set_config(timeout_ms=60000)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0


_EXPECTED = {}


def expected_results(stream):
    """The oracle's result of every job: fib value, n-queens count, and for a sort job the
    checksum of its sorted keys (every key at its position; oracle.job_result, with numpy's
    sort for speed: the oracle's merge sort equals it, tests/test_oracle.py)."""
    key = (stream.seed, stream.njobs)
    if key not in _EXPECTED:
        out = np.empty(stream.njobs, dtype=np.int64)
        nq_cache = {}
        for i in range(stream.njobs):
            k, p = int(stream.kind[i]), int(stream.param[i])
            if k == 0:
                out[i] = oracle.fib(p)
            elif k == 1:
                if p not in nq_cache:
                    nq_cache[p] = oracle.nq(p)[0]
                out[i] = nq_cache[p]
            else:
                out[i] = oracle.sort_checksum(np.sort(oracle.gen_keys_u32(p, int(stream.key_seed[i]))))
        _EXPECTED[key] = out
    return _EXPECTED[key]


def check_stream(stream, results, sort_jobs=0):
    """Every job's result against the oracle (for a sort job that checks every sorted key);
    the first `sort_jobs` sort jobs' keys are also read back and compared directly."""
    exp = expected_results(stream)
    bad = np.nonzero(np.asarray(results, dtype=np.int64) != exp)[0]
    assert bad.size == 0, [(int(i), int(stream.kind[i]), int(stream.param[i])) for i in bad[:8]]
    for i in np.where(stream.kind == 2)[0][:sort_jobs]:
        keys = oracle.gen_keys_u32(int(stream.param[i]), int(stream.key_seed[i]))
        assert np.array_equal(stream.sorted_keys(int(i)), np.sort(keys)), i


def test_job_generator_matches_oracle():
    k, p, s = gtap_amd.jobs_generate(2024, 4096)
    k2, p2, s2 = oracle.jobs(2024, 4096)
    assert np.array_equal(k, k2) and np.array_equal(p, p2) and np.array_equal(s, s2)


def test_stream_small_all_forms():
    st = gtap_amd.Stream(seed=7, njobs=48)
    try:
        print("case stream small batched", flush=True)
        res_b, lat_b, stats = st.run_batched()
        check_stream(st, res_b, sort_jobs=48)
        assert stats["jobs_taken"] == 48
        ex_b = st.exec_ms()
        # device time from admission to completion: positive, and within the launch-relative
        # completion time (every job is admitted at or after the launch start)
        assert np.all(ex_b > 0) and np.all(ex_b <= lat_b + 1e-6)
        print("case stream small per-job", flush=True)
        res_p, lat_p, tot = st.run_per_job()
        check_stream(st, res_p)
        assert np.array_equal(res_b, res_p)  # alone == in the mix
        print("case stream small served", flush=True)
        res_s, lat_s, tot_s, stats_s = st.serve(rate=2000.0, arrival_seed=3)
        check_stream(st, res_s)
        assert np.all(lat_s > 0)
        ex_s = st.exec_ms()
        assert np.all(ex_s > 0) and np.all(ex_s < tot_s)
        print("case stream small served per-job", flush=True)
        res_q, lat_q, tot_q, _ = st.serve(rate=2000.0, arrival_seed=3, per_job_launch=True)
        check_stream(st, res_q)
    finally:
        st.close()


def test_stream_full_batched():
    st = gtap_amd.Stream(seed=2024, njobs=4096)
    try:
        print("case stream 4096 batched", flush=True)
        res, lat, stats = st.run_batched()
        check_stream(st, res, sort_jobs=16)  # all 4096 results; every sort job by its checksum
        assert stats["jobs_taken"] == 4096
        print("case stream 4096 served", flush=True)
        res_s, lat_s, tot_s, _ = st.serve(rate=40000.0, arrival_seed=11)
        check_stream(st, res_s, sort_jobs=16)
        assert np.array_equal(res, res_s)
    finally:
        st.close()


def test_stream_graph_replays():
    """The batched form is a captured HIP graph: replays with eager launches and host
    synchronisation in between each give the oracle's results, sorted keys included."""
    st = gtap_amd.Stream(seed=99, njobs=96)
    try:
        for rep in range(4):
            print(f"case stream graph replay {rep}", flush=True)
            res, lat, stats = st.run_batched()
            check_stream(st, res, sort_jobs=96)
            assert stats["jobs_taken"] == 96
            v, _ = gtap_amd.fib(15 + rep)  # eager work on the same stream between replays
            assert v == oracle.fib(15 + rep)
    finally:
        st.close()


def test_stream_graph_recaptured_after_config_change():
    st = gtap_amd.Stream(seed=5, njobs=32)
    try:
        print("case stream graph before profiling switch", flush=True)
        res0, _, _ = st.run_batched()
        assert gtap_amd.gtap_set_profiling(True, 256) == 0
        try:
            print("case stream graph with profiling on", flush=True)
            res1, _, _ = st.run_batched()
        finally:
            assert gtap_amd.gtap_set_profiling(False, 0) == 0
        print("case stream graph after profiling switch", flush=True)
        res2, _, _ = st.run_batched()
        check_stream(st, res2)
        assert np.array_equal(res0, res1) and np.array_equal(res0, res2)
    finally:
        st.close()


def test_stream_merge_leaf_setting():
    """The sort jobs' one-wave merge size is a run-time setting: every allowed value gives the
    oracle's results in the batched graph (re-captured per value) and served; values outside
    [512, 16384] are refused and leave the setting unchanged."""
    st = gtap_amd.Stream(seed=31, njobs=120)
    try:
        ref = None
        for leaf in (512, 1024, 2048, 16384, 4096):
            print(f"case stream merge_leaf={leaf} batched", flush=True)
            st.set_merge_leaf(leaf)
            res, _, stats = st.run_batched()
            check_stream(st, res, sort_jobs=120)
            assert stats["jobs_taken"] == 120
            ref = res if ref is None else ref
            assert np.array_equal(res, ref)
        print("case stream merge_leaf=16384 served", flush=True)
        st.set_merge_leaf(16384)
        res_s, _, _, _ = st.serve(rate=5000.0, arrival_seed=2)
        check_stream(st, res_s, sort_jobs=8)
        for bad in (0, 511, 16385, -5):
            with pytest.raises(gtap_amd.GtapError):
                st.set_merge_leaf(bad)
        res2, _, _ = st.run_batched()  # still 16384
        check_stream(st, res2)
    finally:
        st.close()
