"""Command line: python -m gtap_amd.lower INPUT.cu [-o OUTPUT.hip] [--name NAME]
[-D NAME[=VALUE] ...] [--mode thread|block] [--plain-accesses]

-D takes the compile-time switches the reference's build passes to its compiler
(GTAP_NUM_QUEUES, GTAP_TERMINATE_ON_FIRST_TASK_FINISH, program macros)."""
import argparse
import os
import sys

from . import LoweringError, lower_file


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Lower a #pragma gtap source to HIP for the MI355X runtime")
    ap.add_argument("input")
    ap.add_argument("-o", "--output", default=None)
    ap.add_argument("--name", default=None, help="program name (namespace and entry-point suffixes)")
    ap.add_argument("-D", dest="defines", action="append", default=[], metavar="NAME[=VALUE]")
    ap.add_argument("--mode", choices=("thread", "block"), default=None,
                    help="default: block when the source includes gtap_block.cuh")
    ap.add_argument("--plain-accesses", action="store_true",
                    help="leave pointer accesses in task code as written (no coherent rewrite)")
    a = ap.parse_args(argv)
    defines = {}
    for d in a.defines:
        k, _, v = d.partition("=")
        defines[k] = v
    out = a.output or os.path.splitext(a.input)[0] + ".lowered.hip"
    try:
        lower_file(a.input, out, a.name, defines, a.mode, not a.plain_accesses)
    except LoweringError as e:
        print(f"{a.input}: error: {e}", file=sys.stderr)
        return 1
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
