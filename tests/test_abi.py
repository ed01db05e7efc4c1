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
    assert lib.bk_abi_version() == 3


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof and field offsets of every ABI struct agree between the header (gcc) and the ctypes mirror."""
    import ctypes
    from baking_amd import _native
    structs = {"bk_params": _native.BkParams, "bk_dough": _native.BkDough, "bk_tuning": _native.BkTuning,
               "bk_state": _native.BkState, "bk_scene": _native.BkScene, "bk_grid": _native.BkGrid,
               "bk_adjoint": _native.BkAdjoint, "bk_settling": _native.BkSettling}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, got):
        cname, field, value = line.split()
        cls = structs[cname]
        want = ctypes.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert int(value) == want, (cname, field, int(value), want)
