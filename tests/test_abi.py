"""The C ABI: libsbz_hip.so loads without a GPU and exports every function include/sbz.h declares;
calls that need a device fail with a status code and a message instead of crashing."""
import ctypes
import re
from pathlib import Path

from symboltz import _lib

HDR = Path(__file__).resolve().parent.parent / "include" / "sbz.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HDR.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sbz_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_device_calls_fail_cleanly_without_gpu():
    if _lib.device_count() > 0:
        return
    L = _lib.lib()
    h = ctypes.c_void_p()
    st = L.sbz_model_create(None, None, ctypes.byref(h))
    assert st != 0
    assert L.sbz_last_error()
