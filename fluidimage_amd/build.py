"""Build libpiv_hip.so in-tree with hipcc for gfx950.

    python -m fluidimage_amd.build

Packed FP32 instructions are disabled for all device code
(``-target-feature -packed-fp32-ops``).
"""

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = [HERE / "csrc" / "piv_hip.hip", HERE / "csrc" / "piv_tps.hip",
           HERE / "csrc" / "piv_deform.hip", HERE / "csrc" / "piv_refine.hip",
           HERE / "csrc" / "piv_direct.hip", HERE / "csrc" / "piv_graph.hip"]
OUTPUT = HERE / "libpiv_hip.so"
ARCH = os.environ.get("PIV_HIP_ARCH", "gfx950")


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found")
    return exe


def command(output=OUTPUT, extra=()):
    return [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        *extra,
        "-o", str(output),
        *map(str, SOURCES),
    ]


def build(verbose=True):
    """Compile the library (always) and atomically replace the output."""
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(
        line for line in proc.stderr.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{errs}")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    build()
    sys.exit(0)
