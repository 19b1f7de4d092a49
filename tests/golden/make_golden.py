"""Write the golden fixtures of tests/golden/ from the reference's own CPU programs.

The programs are compiled from the reference sources into oracle/_ref/ by
``make -C oracle ref`` (the recipe in oracle/Makefile); this script runs them and stores
their outputs as data:

* reference_outputs.json -- stdout values of seq_fib n (n = 0..32), seq_nq N (N = 1..13),
  omp_nq N cutoff (N = 4..11), seq_cilksort / seq_mergesort verdicts on gen_vector files;
* gen_vector_1000.bin / gen_vector_4097.bin -- files written by the reference gen_vector
  (format: size_t n, then n int32 values; evaluation/util/gen_vector.c:21-36).

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")


def run(args: list[str]) -> str:
    return subprocess.run(args, check=True, capture_output=True, text=True, timeout=600).stdout


def main() -> int:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out: dict = {"seq_fib": {}, "seq_nq": {}, "omp_nq": {}, "seq_cilksort": {}, "seq_mergesort": {}}
    for n in range(0, 33):
        m = re.search(r"Fibonacci of (\d+) is (-?\d+)", run([os.path.join(REF, "seq_fib"), str(n)]))
        out["seq_fib"][str(n)] = int(m.group(2))
    for n in range(1, 14):
        m = re.search(r"N-Queens\((\d+)\) = (\d+)", run([os.path.join(REF, "seq_nq"), str(n)]))
        out["seq_nq"][str(n)] = int(m.group(2))
    for n in range(4, 12):
        m = re.search(r"N-Queens\((\d+)\) = (\d+)", run([os.path.join(REF, "omp_nq"), str(n), "3"]))
        out["omp_nq"][str(n)] = int(m.group(2))
    for n in (1000, 4097):
        path = os.path.join(HERE, f"gen_vector_{n}.bin")
        run([os.path.join(REF, "gen_vector"), str(n), path])
    with tempfile.TemporaryDirectory() as td:
        for n in (1, 2, 100, 4097, 100000):
            path = os.path.join(td, f"v{n}.bin")
            run([os.path.join(REF, "gen_vector"), str(n), path])
            for prog in ("seq_cilksort", "seq_mergesort"):
                s = run([os.path.join(REF, prog), path])
                out[prog][str(n)] = "Correct" in s
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "reference_outputs.json"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
