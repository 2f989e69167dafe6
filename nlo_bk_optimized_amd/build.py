"""Build ``libnlo_gfx950.so`` in-tree with hipcc for gfx950.

    python -m nlo_bk_optimized_amd.build

Compiles every ``csrc/*.hip`` to an object file (in parallel, one hipcc per
source) and links them into one shared library next to this file.  Packed
FP32 instructions are disabled for all device code
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_DIR = PACKAGE_DIR.parent
CSRC = PACKAGE_DIR / "csrc"
INCLUDE = REPO_DIR / "include"
LIBRARY = PACKAGE_DIR / "libnlo_gfx950.so"
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(found).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libnlo_gfx950.so")
    return found


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def compile_command(source: Path, obj: Path) -> list[str]:
    return [hipcc(), *FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", "-o", str(obj), str(source)]


def link_command(objects: list[Path], output: Path) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", str(output), *map(str, objects)]


def _run(cmd: list[str]) -> subprocess.CompletedProcess:
    return subprocess.run(cmd, cwd=PACKAGE_DIR, capture_output=True, text=True)


def _report(result: subprocess.CompletedProcess, verbose: bool) -> None:
    if result.returncode != 0:
        sys.stderr.write(result.stdout)
        sys.stderr.write(result.stderr)
        raise RuntimeError(f"hipcc failed with status {result.returncode}: {' '.join(result.args)}")
    if verbose and result.stderr.strip():
        # the host pass does not know the packed-fp32 target feature: harmless
        lines = [line for line in result.stderr.splitlines()
                 if "packed-fp32-ops" not in line and "feature" not in line]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


def build(verbose: bool = False) -> Path:
    """Compile the library (always recompiles) and return its path."""

    jobs = max(1, min(int(os.environ.get("MAX_JOBS", "8")), 16))
    srcs = sources()
    with tempfile.TemporaryDirectory(prefix="nlo_build_") as tmpdir:
        objs = [Path(tmpdir) / (s.stem + ".o") for s in srcs]
        cmds = [compile_command(s, o) for s, o in zip(srcs, objs)]
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            results = list(pool.map(_run, cmds))
        for r in results:
            _report(r, verbose)
        tmp = LIBRARY.with_suffix(".so.tmp")
        link = link_command(objs, tmp)
        if verbose:
            print(" ".join(link), flush=True)
        _report(_run(link), verbose)
    tmp.replace(LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    path = build(verbose=True)
    print(f"built {path}")
