"""The C ABI library: it loads without a GPU, exports every entry point include/d2fem.h
declares, and its structs agree with the Python binding."""
import ctypes
import re
from pathlib import Path

import pytest

from diff_2d_fem_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "d2fem.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(d2f_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "d2f_step_plain" in names and "d2f_elastic_grad_diag_f64" in names
    assert len(names) >= 9


def test_library_loads_and_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} is not built (python -m diff_2d_fem_amd.build)")
    L = _lib.lib()                      # checks struct sizes and ABI version itself
    raw = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(raw, n)]
    assert not missing, f"declared in d2fem.h but not exported: {missing}"
    assert L.d2f_abi_version() == 1
    assert L.d2f_last_error() is not None
