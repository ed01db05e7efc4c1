"""Build ``libbaking_hip.so`` for gfx950 with hipcc (no CMake, no JIT cache).

Every ``csrc/*.hip`` is compiled to an object in parallel and linked into one
shared library next to this file, so it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libbaking_hip.so"
ARCH = os.environ.get("BAKING_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops off: packed FP32 forms are kept out of every kernel.
# -amdgpu-mfma-vgpr-form: matrix-core results land in ordinary VGPRs, not in
# accumulation registers that the network kernel would copy out one by one.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-mfma-vgpr-form", "-I", str(HERE.parent / "include")]


def _stale(src: Path, obj: Path) -> bool:
    if not obj.exists():
        return True
    deps = [src, *CSRC.glob("*.h"), *(HERE.parent / "include").glob("*.h")]
    return obj.stat().st_mtime < max(d.stat().st_mtime for d in deps)


def _compile(src: Path, force: bool = True) -> Path:
    obj = BUILD / (src.stem + ".o")
    if not force and not _stale(src, obj):
        return obj
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        noise = "is not a recognized feature for this target"
        err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
        raise RuntimeError(f"hipcc failed for {src.name}:\n{err}")
    return obj


def build(verbose: bool = True, force: bool = True) -> Path:
    """Compile every source (or, with force=False, those older than a source or header) and link."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objects = list(pool.map(lambda src: _compile(src, force), sources))
    tmp = LIB.with_name(LIB.name + ".tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objects), "-o", str(tmp)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
