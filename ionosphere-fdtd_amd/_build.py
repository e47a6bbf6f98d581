"""Build libgeofdtd.so in-tree with hipcc for gfx950."""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_ROOT = PACKAGE_DIR.parent
SOURCES = [PACKAGE_DIR / "csrc" / "geofdtd.hip", PACKAGE_DIR / "csrc" / "surrogate.hip"]
HEADERS = [REPO_ROOT / "include" / "geofdtd.h", REPO_ROOT / "include" / "geosurrogate.h",
           PACKAGE_DIR / "csrc" / "geofdtd_kernels.h", PACKAGE_DIR / "csrc" / "geofdtd_internal.h"]
OUTPUT = PACKAGE_DIR / "libgeofdtd.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep packed FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> Path:
    """Compile every HIP source of the package into libgeofdtd.so (always rebuilds)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPCC_FLAGS, f"-I{REPO_ROOT / 'include'}", *map(str, SOURCES), "-o", str(tmp)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed ({result.returncode}):\n{result.stderr}")
    if verbose and result.stderr:
        print(result.stderr)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
