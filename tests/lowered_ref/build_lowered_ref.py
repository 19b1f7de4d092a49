"""Builds the reference's own `#pragma gtap` programs, unmodified, with this repository's
lowering tool (gtap_amd.lower) and hipcc for gfx950.

* build(): the programs the GPU tests run (tests/test_gpu_lowered_reference.py), each into
  tests/lowered_ref/_lib/lib<name>.so, linked against gtap_amd/_lib/libgtap_amd.so. The
  libraries are git-ignored and travel with the tree; nothing at run time reads the
  reference tree.
* compile_all(): every `#pragma gtap` source of the reference lowered and compiled to an
  object, for the CPU test of the lowering (tests/test_lower_reference.py).

The lowered sources are written to a temporary directory and removed after compiling. The
-D switches are the ones the reference's build files pass (evaluation/3-epaq/*/Makefile
GTAP_NUM_QUEUES, evaluation/3-epaq/cilksort/epaq_cilksort.sh:153-154 cutoffs, and
GTAP_TERMINATE_ON_FIRST_TASK_FINISH where those Makefiles set it). Run as
`python -m tests.lowered_ref.build_lowered_ref` from the repository root.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFERENCE = "/root/reference"
LIBDIR = os.path.join(HERE, "_lib")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# name -> (source under the reference tree, -D switches)
PROGRAMS = {
    "ref_fib": ("examples/fib/fib.cu", {}),
    "ref_nq": ("examples/nq/nq.cu", {}),
    "ref_mergesort": ("examples/mergesort/mergesort.cu", {}),
    "ref_cilksort": ("examples/cilksort/cilksort.cu", {}),
    "ref_bfs_thread": ("examples/asynchronous_bfs/asynchronous_bfs_thread.cu", {}),
    "ref_bfs_block": ("examples/asynchronous_bfs/asynchronous_bfs_block.cu", {}),
    "ref_spmv_thread": ("examples/spmv/spmv_thread.cu", {}),
    "ref_tree_thread": ("examples/binary_tree/binary_tree_thread.cu", {}),
    "ref_fib_queue_1": ("evaluation/3-epaq/fib/fib_queue_1.cu",
                        {"GTAP_NUM_QUEUES": "1", "GTAP_TERMINATE_ON_FIRST_TASK_FINISH": ""}),
    "ref_fib_queue_3": ("evaluation/3-epaq/fib/fib_queue_3.cu",
                        {"GTAP_NUM_QUEUES": "3", "GTAP_TERMINATE_ON_FIRST_TASK_FINISH": ""}),
}


def available() -> bool:
    return os.path.isdir(os.path.join(REFERENCE, "examples"))


def defines_for(rel: str) -> dict:
    """The reference build's -D switches for a source (for sources outside PROGRAMS)."""
    d = {}
    m = re.search(r"_queue_(\d)\.cu$", rel)
    if m:
        d["GTAP_NUM_QUEUES"] = m.group(1)
    if "3-epaq/cilksort/" in rel:
        d.update({"TASK_SPAWN_CUTOFF_SORT": "64", "TASK_SPAWN_CUTOFF_MERGE": "256"})
    return d


def all_sources() -> list:
    out = []
    for path in sorted(glob.glob(os.path.join(REFERENCE, "**", "*.cu"), recursive=True)):
        with open(path, encoding="utf-8", errors="replace") as f:
            if re.search(r"^\s*#\s*pragma\s+gtap\b", f.read(), re.M):
                out.append(os.path.relpath(path, REFERENCE))
    return out


def _hipcc_flags() -> list:
    from gtap_amd import build as gbuild
    return [*gbuild.CFLAGS, "-I", gbuild.CSRC]


def _one(rel: str, name: str, defines: dict, out: str, shared: bool) -> tuple:
    from gtap_amd import build as gbuild
    from gtap_amd import lower
    with tempfile.TemporaryDirectory() as d:
        gen = os.path.join(d, name + ".hip")
        try:
            lower.lower_file(os.path.join(REFERENCE, rel), gen, name, defines)
        except lower.LoweringError as e:
            return rel, f"lowering: {e}"
        if shared:
            cmd = [gbuild.HIPCC, *_hipcc_flags(), "-shared", gen, "-L", gbuild.LIBDIR, "-lgtap_amd",
                   "-Wl,-rpath,$ORIGIN/../../../gtap_amd/_lib", "-o", out + ".tmp"]
        else:
            cmd = [gbuild.HIPCC, *_hipcc_flags(), "-c", gen, "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            errs = [l for l in (r.stdout + r.stderr).splitlines() if "error" in l]
            return rel, "hipcc: " + "\n".join(errs[:10])
        if shared:
            os.replace(out + ".tmp", out)
    return rel, ""


def build(jobs: int = 8) -> dict:
    """Builds PROGRAMS into _lib; returns {name: error message} for the ones that failed."""
    if not available():
        return {}
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_one, rel, name, dfn, os.path.join(LIBDIR, f"lib{name}.so"), True): name
                for name, (rel, dfn) in PROGRAMS.items()}
        res = {futs[f]: f.result()[1] for f in futs}
    return {k: v for k, v in res.items() if v}


def compile_all(out_dir: str, jobs: int = 8) -> dict:
    """Lowers and compiles every `#pragma gtap` source of the reference to an object in
    out_dir; returns {source: error message} for the ones that failed."""
    srcs = all_sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_one, rel, re.sub(r"\W", "_", rel), defines_for(rel),
                          os.path.join(out_dir, re.sub(r"\W", "_", rel) + ".o"), False) for rel in srcs]
        res = [f.result() for f in futs]
    return {rel: msg for rel, msg in res if msg}


if __name__ == "__main__":
    bad = build()
    for k, v in bad.items():
        print(f"{k}: {v}", file=sys.stderr)
    sys.exit(1 if bad else 0)
