"""Build recipe for libgtap_amd.so (HIP for gfx950 only).

Every ``.hip`` file under ``csrc/host`` and ``csrc/programs`` is compiled with hipcc into
``_build/`` and linked into ``_lib/libgtap_amd.so`` in-tree, so the library travels with
the source tree. ``python -m gtap_amd.build`` runs it.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libgtap_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    # packed fp32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list[str]:
    out = []
    for sub in ("host", "programs"):
        out += sorted(glob.glob(os.path.join(CSRC, sub, "*.hip")))
    return out


def _compile(src: str) -> tuple[str, str]:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and "unknown target feature" not in l.lower() and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msgs}")
    return obj, msgs


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, m in results:
            if m:
                print(m)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_lowered(jobs)
    return LIB


EXAMPLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lower", "examples")
GEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_gen")


def lowered_lib(name: str) -> str:
    return os.path.join(LIBDIR, f"libgtap_lowered_{name}.so")


def _build_one_lowered(src: str) -> str:
    from gtap_amd import lower
    name = os.path.basename(src).split(".")[0]
    gen = os.path.join(GEN, f"{name}.lowered.hip")
    lower.lower_file(src, gen, name)
    out = lowered_lib(name)
    tmp = out + ".tmp"
    cmd = [HIPCC, *CFLAGS, "-I", CSRC, "-shared", gen, "-L", LIBDIR, "-lgtap_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for the lowered {src}:\n{msgs}")
    os.replace(tmp, out)
    return out


def build_lowered(jobs: int | None = None) -> list[str]:
    """Lowers every `#pragma gtap` example (gtap_amd/lower/examples/*.gtap.cu) with the
    lowering tool and builds each into gtap_amd/_lib/libgtap_lowered_<name>.so."""
    os.makedirs(GEN, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(EXAMPLES, "*.gtap.cu")))
    with cf.ThreadPoolExecutor(max_workers=jobs or min(4, max(1, len(srcs)))) as ex:
        return list(ex.map(_build_one_lowered, srcs))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
