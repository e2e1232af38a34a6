"""Build libpiv_hip.so in-tree with hipcc for gfx950.

    python -m fluidimage_amd.build

Every source is compiled to its own object in parallel (``_build/`` next to
the package), then the objects are linked into the shared library, which
replaces the previous one atomically. Packed FP32 instructions are disabled
for all device code (``-target-feature -packed-fp32-ops``).
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
SOURCES = [CSRC / "piv_hip.hip", CSRC / "piv_tps.hip", CSRC / "piv_deform.hip",
           CSRC / "piv_refine.hip", CSRC / "piv_direct.hip", CSRC / "piv_graph.hip",
           CSRC / "piv_preproc.hip"]
OUTPUT = HERE / "libpiv_hip.so"
OBJDIR = HERE.parent / "_build"
ARCH = os.environ.get("PIV_HIP_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found")
    return exe


def sources():
    return [s for s in SOURCES if s.exists()]


def command(output=OUTPUT, extra=()):
    """One-shot compile-and-link command (kept for reference builds)."""
    return [hipcc(), *FLAGS, "-shared", *extra, "-o", str(output), *map(str, sources())]


def _filtered(stderr):
    return "\n".join(line for line in stderr.splitlines() if "packed-fp32-ops" not in line)


def _compile(src, extra):
    obj = OBJDIR / (src.stem + ".o")
    cmd = [hipcc(), *FLAGS, *extra, "-c", "-o", str(obj), str(src)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, proc.returncode, _filtered(proc.stderr), cmd


def build(verbose=True, extra=(), jobs=None):
    """Compile every source (always) and atomically replace the library."""
    OBJDIR.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(lambda s: _compile(s, extra), srcs))
    failed = [r for r in results if r[2] != 0]
    for src, _obj, _rc, errs, cmd in results:
        if verbose:
            print(" ".join(cmd), flush=True)
        if errs.strip() and verbose:
            print(errs, flush=True)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(f"{r[0].name}:\n{r[3]}" for r in failed))
    tmp = OUTPUT.with_suffix(".so.tmp")
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
            *(str(r[1]) for r in results)]
    if verbose:
        print(" ".join(link), flush=True)
    proc = subprocess.run(link, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{_filtered(proc.stderr)}")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    build()
    sys.exit(0)
