"""The C ABI: the HIP library loads without a GPU and exports every function the header declares."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "baking_abi.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(bk_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("bk_substep", "bk_cell_keys", "bk_sort_pairs", "bk_workspace_create", "bk_ParticleToGrid",
                     "bk_GridToParticle", "bk_UpdateGrid", "bk_UpdateDing", "bk_DiffuseGrid", "bk_ScalarToGrid"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from baking_amd import _native
    lib_path = _native.library_path()
    if not lib_path.exists():
        pytest.skip("libbaking_hip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib_path)], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bk_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared in baking_abi.h but not exported: {missing}"


def test_ctypes_binding_covers_the_header():
    from baking_amd import _native
    if not _native.library_path().exists():
        pytest.skip("libbaking_hip.so not built")
    lib = _native.load()  # loads on a machine without a GPU: no device call happens
    assert set(_native.SIGNATURES) == set(declared_functions())
    assert lib.bk_abi_version() == 1
