"""Build the gfx950 HIP library in-tree: pm_amd/lib/libpstm_hip.so.

    python -m pm_amd.build

hipcc cross-compiles for gfx950 without a GPU present.  Packed-FP32 VALU ops
are switched off for every HIP translation unit (-packed-fp32-ops).
"""

from __future__ import annotations

import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
# pstm_hip.hip is the main translation unit; gather_part1..7 compile its gather
# kernel groups separately (the slowest part of the build) so they run in parallel
SOURCES = [PKG / "csrc" / "pstm_hip.hip", PKG / "csrc" / "pstm_mlp.hip", PKG / "csrc" / "pstm_output.hip",
           *(PKG / "csrc" / f"gather_part{i}.hip" for i in range(1, 8))]
OUT = PKG / "lib" / "libpstm_hip.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str], verbose: bool) -> subprocess.Popen:
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def _check(p: subprocess.Popen, verbose: bool) -> None:
    _, stderr = p.communicate()
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in stderr.splitlines() if noise not in l)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed ({p.returncode}):\n{err}")
    if err.strip() and verbose:
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    """Compile every source to an object in parallel, then link the library."""
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="pm_amd_build_") as d:
        objs = [Path(d) / (src.stem + ".o") for src in SOURCES]
        procs = [_run([hipcc(), *flags, "-I", str(REPO / "include"), "-c", str(src), "-o", str(obj)], verbose)
                 for src, obj in zip(SOURCES, objs)]
        for p in procs:
            _check(p, verbose)
        _check(_run([hipcc(), *HIPCC_FLAGS, *map(str, objs), "-o", str(tmp)], verbose), verbose)
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
