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
    assert L.d2f_abi_version() == 2
    assert L.d2f_last_error() is not None


def test_scene_device_mirror_is_current():
    """csrc/scene_dev.h (the scene kernel's view of d2f_scene) matches what its generator writes
    from include/d2fem.h today, and names every field of the header's struct."""
    import importlib.util
    csrc = Path(__file__).resolve().parents[1] / "diff_2d_fem_amd" / "csrc"
    spec = importlib.util.spec_from_file_location("gen_scene_dev", csrc / "gen_scene_dev.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    flds = gen.fields(HEADER.read_text())
    assert (csrc / "scene_dev.h").read_text() == gen.render(flds)
    names = [n for _, _, n in flds]
    assert len(names) == len(set(names))
    assert {"x", "v", "tri", "el", "excl_bits", "max_pairs", "xc_lds"} <= set(names)
    assert sum(1 for f in _lib.Scene._fields_) == len(names)


def test_step_scene_rejects_affine_groups_without_lds_capacity():
    """d2f_step_scene checks its arguments before any GPU call: a scene with affine groups but
    no max_ag (the LDS the scene kernel keeps their state in) is refused with a message."""
    L = _lib.lib()
    s, p = _lib.Scene(), _lib.Params()
    dummy = ctypes.create_string_buffer(64)             # never dereferenced by the host checks
    addr = ctypes.cast(dummy, ctypes.c_void_p).value
    s.B, s.N, s.max_cells = 1, 3, 1
    s.sc_tri = s.sc_cs = s.sc_tri_start = s.ag_start = addr
    s.n_ag, s.max_ag = 1, 0
    rc = L.d2f_step_scene(ctypes.byref(s), ctypes.byref(p), None)
    assert rc != 0
    assert "max_ag" in L.d2f_last_error().decode()
