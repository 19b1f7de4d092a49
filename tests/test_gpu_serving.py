"""Served admission by free task ids (GPU): with a small task-id pool, the reservation ledger
must hold jobs back; every job still gives the oracle's result, the same as in the batched
run (alone == in the mix), and the job table is reused as a ring when it is smaller than the
stream."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu

SMALL_POOL = 768 << 20  # 4.2 M task records of 192 bytes


@pytest.fixture(scope="module", autouse=True)
def runtime():
    gtap_amd.gtap_set_config(timeout_ms=60000, pool_bytes=SMALL_POOL)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


def expected(st):
    out = []
    for k, p, s in zip(st.kind, st.param, st.key_seed):
        if k == 2:
            out.append(oracle.sort_checksum(np.sort(oracle.gen_keys_u32(int(p), int(s)))))
        else:
            out.append(oracle.job_result(int(k), int(p), int(s)))
    return np.array(out, dtype=np.int64)


def test_admission_holds_jobs_back_and_results_stay_exact():
    st = gtap_amd.Stream(seed=31, njobs=600)
    try:
        exp = expected(st)
        print("case served 600 jobs, small pool, all arriving within ~3 ms", flush=True)
        res, lat, tot, stats = st.serve(rate=200_000.0, arrival_seed=5, max_live=1024)
        assert np.array_equal(res, exp)
        assert stats["admission_holds"] > 0, stats
        assert stats["error_code"] == 0
        print("case batched, same jobs", flush=True)
        res_b, _, _ = st.run_batched()
        assert np.array_equal(res_b, exp)
        print("case served with a 256-slot job ring (max_live 64)", flush=True)
        res_r, lat_r, _, stats_r = st.serve(rate=200_000.0, arrival_seed=6, max_live=64)
        assert np.array_equal(res_r, exp)
        assert stats_r["jobs_taken"] == 600
        assert np.all(lat_r > 0)
    finally:
        st.close()
