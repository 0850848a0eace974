"""The C-ABI library loads and exports every entry point include/*.h declares
(no device work: these run without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(gpq_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gpq_queue_create" in names and "gpq_des_run" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from gpupq_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from gpupq_amd import _lib
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared_functions() == bound


def test_status_strings_without_device():
    from gpupq_amd import _lib
    lib = _lib.load()
    assert lib.gpq_status_string(0) == b"ok"
    assert lib.gpq_status_string(2) == b"capacity exceeded"
    assert b"gfx950" in lib.gpq_version()


def test_struct_layouts_match_header():
    from gpupq_amd import _lib
    assert ctypes.sizeof(_lib.Event) == 16
    assert ctypes.sizeof(_lib.QueueConfig) == 48
    assert ctypes.sizeof(_lib.PholdConfig) == 32
