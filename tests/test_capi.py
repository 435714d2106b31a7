"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/*.h declares, and reports errors (no device here) through gc_last_error."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(gc_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "gc_laplacian" in names and "gc_ch_integrate" in names and len(names) >= 30


def test_library_exports_every_declared_symbol(gc):
    L = C.CDLL(gc.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header(gc):
    missing = set(_declared()) - set(gc.bound_symbols())
    assert not missing, sorted(missing)


def test_version(gc):
    assert gc.lib().gc_version() >= 10000


def test_errors_without_device(gc):
    if gc.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(gc.GridcalcError) as e:
        gc.Context()
    assert e.value.code == gc.ERR_DEVICE
    assert "no HIP device" in str(e.value)


def test_null_context_rejected(gc):
    g = gc.Grid.make(8, 8, 8)
    rc = gc.lib().gc_laplacian(None, C.byref(g), gc.F64, 2, None, None)
    assert rc == gc.ERR_INVALID
    assert b"null gc_context" in gc.lib().gc_last_error()
