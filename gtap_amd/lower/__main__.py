"""Command line: python -m gtap_amd.lower INPUT.cu [-o OUTPUT.hip] [--name NAME]"""
import argparse
import os
import sys

from . import LoweringError, lower_file


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Lower a #pragma gtap source to HIP for the MI355X runtime")
    ap.add_argument("input")
    ap.add_argument("-o", "--output", default=None)
    ap.add_argument("--name", default=None, help="program name (namespace and entry-point suffixes)")
    a = ap.parse_args(argv)
    out = a.output or os.path.splitext(a.input)[0] + ".lowered.hip"
    try:
        lower_file(a.input, out, a.name)
    except LoweringError as e:
        print(f"{a.input}: error: {e}", file=sys.stderr)
        return 1
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
