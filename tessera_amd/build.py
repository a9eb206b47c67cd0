"""Builds the gfx950 HIP library in-tree (tessera_amd/libtessera_mpm_hip.so).

The library is compiled with hipcc for gfx950 only.  Packed FP32 instructions
are disabled for all device code (``-packed-fp32-ops``).  The build always
recompiles; callers that only need the library loaded use
``tessera_amd.mpm.load_library``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_ROOT = PACKAGE_DIR.parent
CSRC = PACKAGE_DIR / "csrc"
LIBRARY = PACKAGE_DIR / "libtessera_mpm_hip.so"
SOURCES = ["mpm_kernels.hip", "mpm_session.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{REPO_ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line and line.strip()
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out:
        print(out, file=sys.stderr)


def build_library(jobs: int = 2) -> Path:
    """Compiles every HIP source for gfx950 and links the shared library."""
    build_dir = PACKAGE_DIR / "_build"
    build_dir.mkdir(exist_ok=True)
    objects = []
    procs = []
    for src in SOURCES:
        obj = build_dir / (Path(src).stem + ".o")
        objects.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for cmd, proc in procs:
        out, _ = proc.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if proc.returncode != 0:
            raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
        if out:
            print(out, file=sys.stderr)
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)])
    os.replace(tmp, LIBRARY)
    return LIBRARY


def build_oracle() -> Path:
    """Compiles the CPU oracle used by the tests (oracle/liboracle.so)."""
    script = REPO_ROOT / "oracle" / "build_oracle.sh"
    _run(["sh", str(script)])
    return REPO_ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_library()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print(LIBRARY)
