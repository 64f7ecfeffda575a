"""The C ABI (CPU): libepsx.so loads without a GPU, exports every entry point include/epsx.h
declares, and the Python bindings declare the same set with matching struct sizes."""
import ctypes
import os
import re
import subprocess

import epsx
from epsx import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "epsx.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(epsx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert "epsx_scgd0_run" in names and "epsx_open" in names
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (epsx_[a-z0-9_]+)", out))
    assert set(names) <= exported
    # every exported epsx_ symbol is declared (nothing undocumented crosses the boundary)
    assert exported <= set(names)


def test_python_bindings_cover_the_header():
    assert sorted(n for n, _, _ in _lib.EXPORTS) == declared()
    assert _lib.lib().epsx_abi_version() == _lib.EPSX_ABI_VERSION


def test_struct_layouts():
    # int64 + 6 int64 + 10 pointers + 2 doubles + (int64, 2 doubles, ptr) x2 + int64 + 2 ptrs
    assert ctypes.sizeof(_lib.Scgd0Problem) == 8 * (1 + 6 + 10 + 2 + 4 + 4 + 3)
    assert ctypes.sizeof(_lib.Scgd0Params) == 8 * (1 + 1 + 4 + 2) + 4 * 2 + 8 * 3 + 8 * 2 + 4 * 4 + 8 + 4 * 2
    assert ctypes.sizeof(_lib.Scgd0Outputs) == 8 * (1 + 8 + 1) + 4 * 2 + 8


def test_host_entry_points_without_device():
    # the host-side grid maps need no GPU
    g = epsx.FrequencyGrid(0.0, 0.1, [0, 1, 3, 4])
    j, t = epsx.linear_interpolation_table(g, [-1.0, 0.05, 0.35, 9.0])
    assert list(j) == [1, 1, 3, 3] and list(t[[0, 3]]) == [0.0, 1.0] and abs(t[1] - 0.5) < 1e-15 and abs(t[2] - 0.5) < 1e-15
    klo, wlo, whi = epsx.vertex_shift_tables(g, 3)
    assert klo.shape == (4, 3, 2) and klo.min() >= 1 and klo.max() <= 3
