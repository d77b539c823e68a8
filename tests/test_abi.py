"""The C-ABI library loads and exports every entry point include/helm.h declares.

CPU only: no compute call is made.  Struct layouts seen by ctypes are checked
against the C compiler's view of include/helm.h.
"""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from nsrno_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "helm.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(helm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    assert "helm_solve" in names and "helm_apply" in names and len(names) >= 12


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_errors_without_gpu():
    lib = _lib.lib()
    assert lib.helm_version().decode().startswith("helm")
    assert lib.helm_error_string(1).decode() == "invalid argument"
    # argument validation happens before any device call
    h = ctypes.c_void_p()
    bad = _lib.HelmDesc(2, 2, 1, 0, 0.0, 0.0, 0.1, 0.0, 1.0, 1500.0, 12.0)
    assert lib.helm_plan_create(ctypes.byref(bad), ctypes.byref(h)) == 1
    assert "3 nodes" in lib.helm_last_error().decode()
    assert lib.helm_set_stencil_shape(None, 4, 4, 1) == 1
    assert "null plan" in lib.helm_last_error().decode()


def test_struct_layout_matches_c_compiler(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "helm.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(helm_desc), sizeof(helm_solve_opts),'
                   ' sizeof(helm_solve_info), offsetof(helm_solve_opts, callback), offsetof(helm_desc, x0_mm));'
                   'printf("%zu %zu\\n", sizeof(helm_tuning), offsetof(helm_tuning, stencil_block));'
                   'return 0;}\n')
    exe = tmp_path / "sizes"
    try:
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    except (OSError, subprocess.CalledProcessError):
        pytest.skip("gcc unavailable")
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(v) for v in out]
    want = [ctypes.sizeof(_lib.HelmDesc), ctypes.sizeof(_lib.HelmSolveOpts), ctypes.sizeof(_lib.HelmSolveInfo),
            _lib.HelmSolveOpts.callback.offset, _lib.HelmDesc.x0_mm.offset, ctypes.sizeof(_lib.HelmTuning),
            _lib.HelmTuning.stencil_block.offset]
    assert got == want
