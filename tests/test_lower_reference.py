"""Every `#pragma gtap` source of the reference, unmodified, lowered by gtap_amd.lower and
compiled by hipcc for gfx950 (CPU only; skipped when the reference tree is absent). The
-D switches are those of the reference's own build files (tests/lowered_ref/build_lowered_ref.py)."""
import pytest

from tests.lowered_ref import build_lowered_ref as B


@pytest.mark.skipif(not B.available(), reason="the reference tree is not present")
def test_every_reference_source_lowers_and_compiles(tmp_path):
    srcs = B.all_sources()
    assert len(srcs) >= 30, srcs  # examples/ and evaluation/ programs
    for must in ("examples/fib/fib.cu", "examples/nq/nq.cu", "examples/mergesort/mergesort.cu",
                 "examples/cilksort/cilksort.cu", "examples/asynchronous_bfs/asynchronous_bfs_thread.cu",
                 "examples/asynchronous_bfs/asynchronous_bfs_block.cu", "evaluation/3-epaq/fib/fib_queue_3.cu"):
        assert must in srcs
    failed = B.compile_all(str(tmp_path))
    assert not failed, "\n".join(f"{k}: {v}" for k, v in failed.items())
