"""The C-ABI library loads and exports every entry point include/chimes_gpu.h declares."""
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "chimes_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z0-9_]+)\s*\(", text)))


def test_all_declared_symbols_exported(ch):
    L = ch.lib()
    names = declared_symbols()
    assert "cg_assemble" in names and "cg_fit" in names and "cg_force" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_no_device_call_without_gpu_returns_status(ch):
    """cg_create reports a status (never exits) when it cannot use a device."""
    import ctypes
    s = ch.Setup(os.path.join(ROOT, "tests", "golden", "lsq", "nonorth2", "fm_setup.in"))
    L = ch.lib()
    h = ctypes.c_void_p()
    st = L.cg_create(s.model_ptr, 999, ctypes.byref(h))
    assert st != 0
    assert L.cg_last_error(None)
