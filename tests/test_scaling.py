"""The worker-scaling harness's statistics and CSV layout (CPU only)."""
import csv

from gtap_amd import scaling


def test_median_iqr_nearest_rank():
    # worker_scaling_fib.sh:66-104: q1 index = ceil(m/4) - 1, q3 index = ceil(3m/4) - 1
    assert scaling.median_iqr([5.0]) == (5.0, 0.0, 0.0)
    med, lo, hi = scaling.median_iqr([4, 1, 3, 2])  # sorted 1 2 3 4
    assert med == 2.5 and lo == 2.5 - 1 and hi == 3 - 2.5
    t = list(range(1, 21))  # 20 runs: q1 = s[4] = 5, q3 = s[14] = 15, median 10.5
    assert scaling.median_iqr(t) == (10.5, 5.5, 4.5)


def test_csv_columns_match_reference_layout(tmp_path):
    rows = [{"block_size": 256, "total_threads": 1024, "grid_size": 4, "max_tasks_per_warp": 16384,
             "ws_med": 1.0, "ws_err_low": 0.1, "ws_err_high": 0.2}]
    p = tmp_path / "x.csv"
    scaling.write_csv(rows, str(p))
    with open(p) as f:
        header = f.readline().strip().split(",")
    assert header == ["block_size", "total_threads", "grid_size", "max_tasks_per_warp", "ws_med", "ws_err_low",
                      "ws_err_high", "gq_med", "gq_err_low", "gq_err_high", "chaselev_med", "chaselev_err_low",
                      "chaselev_err_high"]
    r = list(csv.DictReader(open(p)))[0]
    assert r["ws_med"] == "1.000000" and r["gq_med"] == ""
