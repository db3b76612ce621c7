"""CPU checks of the C ABI: the built library loads and exports every
function include/bamrex_hip.h declares, and the ctypes binding covers them
(no compute calls; no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bamrex_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"\b(bx_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for n in ("bx_create", "bx_set_level", "bx_sweep", "bx_flux_box", "bx_update_box",
              "bx_fill_boundary", "bx_est_dt", "bx_advance", "bx_run", "bx_autotune",
              "bx_helm_solve", "bx_cnn_train", "bx_cnn_infer"):
        assert n in names


def test_library_exports_every_declared_function():
    from bamrex_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    l = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(l, n)]
    assert not missing, missing


def test_binding_covers_every_declared_function():
    from bamrex_amd import _lib
    assert sorted(_lib.exported_symbols()) == declared_functions()


def test_binding_loads_with_signatures():
    from bamrex_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    l = _lib.lib()
    for n in _lib.exported_symbols():
        assert getattr(l, n).argtypes is not None
    # a call that needs no device: packing offsets is host-only
    boxes = (_lib.bx_box * 2)()
    for i, (lo, hi) in enumerate((((0, 0, 0), (7, 7, 0)), ((8, 0, 0), (15, 7, 0)))):
        boxes[i].lo[:] = lo
        boxes[i].hi[:] = hi
    offs = (ctypes.c_int64 * 2)()
    tot = ctypes.c_int64()
    assert l.bx_pack_offsets(2, 2, boxes, 2, offs, ctypes.byref(tot)) == 0
    assert offs[0] == 0 and offs[1] % 32 == 0 and tot.value >= 2 * 12 * 12 * 4
