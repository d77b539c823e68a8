"""Build libhelm.so (all HIP sources under csrc/) for gfx950 with hipcc.

Every source is compiled to its own object in parallel (nsrno_amd/_build/),
then the objects are linked into nsrno_amd/libhelm.so.  The translation units
share no device symbols, so no relocatable device code is needed.

Usage: python -m nsrno_amd.build      (writes nsrno_amd/libhelm.so)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
OUT = HERE / "libhelm.so"

FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         # keep the compiler from emitting packed FP32 instructions
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False, jobs=None):
    sources = sorted(CSRC.glob("*.hip"))
    OBJ.mkdir(exist_ok=True)
    inc = ["-I", str(HERE.parent / "include")]
    cmds = [[hipcc(), *FLAGS, *inc, "-c", str(src), "-o", str(OBJ / (src.stem + ".o"))] for src in sources]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(lambda c: _run(c, verbose), cmds))
    failed = [r for r in results if r.returncode != 0]
    for r in failed:
        sys.stderr.write(r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed compiling libhelm.so sources")
    tmp = OUT.with_suffix(".so.tmp")
    link = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp),
            *[str(OBJ / (src.stem + ".o")) for src in sources]]
    r = _run(link, verbose)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libhelm.so failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
