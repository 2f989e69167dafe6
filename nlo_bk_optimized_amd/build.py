"""Build ``libnlo_gfx950.so`` in-tree with hipcc for gfx950.

    python -m nlo_bk_optimized_amd.build

Compiles every ``csrc/*.hip`` into one shared library next to this file.
Packed FP32 instructions are disabled for all device code
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_DIR = PACKAGE_DIR.parent
CSRC = PACKAGE_DIR / "csrc"
INCLUDE = REPO_DIR / "include"
LIBRARY = PACKAGE_DIR / "libnlo_gfx950.so"
ARCH = "gfx950"


def hipcc() -> str:
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(found).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libnlo_gfx950.so")
    return found


def compile_command(output: Path) -> list[str]:
    sources = sorted(str(p) for p in CSRC.glob("*.hip"))
    return [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Wall",
        "-Wno-unused-function",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
        "-o",
        str(output),
        *sources,
    ]


def build(verbose: bool = False) -> Path:
    """Compile the library (always recompiles) and return its path."""

    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = compile_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    jobs = os.environ.get("MAX_JOBS", "8")
    env = dict(os.environ)
    env.setdefault("HIPCC_COMPILE_FLAGS_APPEND", "")
    result = subprocess.run(
        cmd + [f"-parallel-jobs={jobs}"] if False else cmd,
        cwd=PACKAGE_DIR,
        capture_output=True,
        text=True,
        env=env,
    )
    if result.returncode != 0:
        sys.stderr.write(result.stdout)
        sys.stderr.write(result.stderr)
        raise RuntimeError(f"hipcc failed with status {result.returncode}")
    if verbose and result.stderr.strip():
        # the host pass does not know the packed-fp32 target feature: harmless
        lines = [
            line
            for line in result.stderr.splitlines()
            if "packed-fp32-ops" not in line and "feature" not in line
        ]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    tmp.replace(LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    path = build(verbose=True)
    print(f"built {path}")
