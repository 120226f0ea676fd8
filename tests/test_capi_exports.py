"""The C-ABI library loads on a machine without a GPU and exports every entry
point declared in include/jfem_hip.h (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "jfem_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jfem_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from jf_2026_05_25_amd import _lib
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from jf_2026_05_25_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    _lib.load()   # binds argtypes for every symbol


def test_context_creation_fails_cleanly_without_device():
    import torch
    if torch.cuda.is_available():
        return
    from jf_2026_05_25_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.jfem_ctx_create(0, ctypes.byref(h)) != 0
