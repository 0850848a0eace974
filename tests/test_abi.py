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
    assert ctypes.sizeof(_lib.QueueConfig) == 80
    assert ctypes.sizeof(_lib.PholdConfig) == 32


STRUCTS = {
    "gpq_event": "Event", "gpq_queue_config": "QueueConfig", "gpq_phold_config": "PholdConfig",
    "gpq_des_report": "DesReport", "gpq_queue_stats": "QueueStats", "gpq_kernel_profile": "KernelProfile",
    "gpq_astar_config": "AstarConfig", "gpq_astar_report": "AstarReport", "gpq_query_spec": "QuerySpec",
    "gpq_batch_tuning": "BatchTuning", "gpq_serve_config": "ServeConfig", "gpq_request": "Request",
    "gpq_serve_report": "ServeReport", "gpq_phold_instance": "PholdInstance", "gpq_phold_tuning": "PholdTuning",
}


def test_ctypes_layouts_equal_c_compiler(tmp_path):
    """sizeof and every field offset of each header struct, from gcc, against the ctypes mirror."""
    import subprocess
    from gpupq_amd import _lib
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gpupq.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        lines.append('  printf("%s __sizeof__ %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in getattr(_lib, pyname)._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split("\n")
    c = {}
    for line in out:
        if line.strip():
            a, b, v = line.split()
            c[(a, b)] = int(v)
    for cname, pyname in STRUCTS.items():
        st = getattr(_lib, pyname)
        assert c[(cname, "__sizeof__")] == ctypes.sizeof(st), cname
        for fname, _ in st._fields_:
            assert c[(cname, fname)] == getattr(st, fname).offset, (cname, fname)


def test_synthetic_queries_on_host_match_oracle(oracle):
    """gpq_query_synthetic is host code: shape and endpoints equal the oracle's definition."""
    import gpupq_amd as gpq
    for q in range(6):
        s = gpq.query_synthetic(21, q)
        side, t = oracle.query_shape(21, q)
        assert (s.side, s.threshold) == (side, t)
        ref = oracle.query_spec(21, q, oracle.grid(side, side, 21, 1, q, t))
        assert (s.sx, s.sy, s.tx, s.ty) == (ref.sx, ref.sy, ref.tx, ref.ty)


def test_serve_trace_on_host():
    import gpupq_amd as gpq
    a = gpq.serve_trace(200, seed=5, rate=100.0)
    assert a == gpq.serve_trace(200, seed=5, rate=100.0)
    arr = [r["arrival_ms"] for r in a]
    assert arr == sorted(arr)
    mean_gap = arr[-1] / (len(arr) - 1)
    assert 7.0 < mean_gap < 13.5                      # Exp gaps with mean 10 ms
    kinds = [r["kind"] for r in a]
    assert 70 < sum(kinds) < 130                      # about half A*
    ph = [r["size"] for r in a if r["kind"] == 0]
    assert min(ph) >= 1 << 12 and max(ph) <= 1 << 18
    zero = gpq.serve_trace(10, seed=5, rate=0.0)
    assert all(r["arrival_ms"] == 0.0 for r in zero)


def test_external_model_built_against_public_device_header():
    # tests/ext/phold_ext.hip is compiled against include/gpupq_device.h alone; it loads without a
    # GPU and knows the view struct of every queue kind
    path = os.path.join(ROOT, "tests", "ext", "libphold_ext.so")
    lib = ctypes.CDLL(path)
    lib.ext_view_bytes.restype = ctypes.c_uint64
    lib.ext_view_bytes.argtypes = [ctypes.c_int]
    sizes = [lib.ext_view_bytes(k) for k in range(5)]
    assert all(s >= 32 for s in sizes) and lib.ext_view_bytes(9) == 0
    for name in ("ext_phold_create", "ext_phold_populate", "ext_phold_handle_next", "ext_phold_state", "ext_lp_verbs"):
        assert hasattr(lib, name)
    from gpupq_amd import _lib
    libq = ctypes.CDLL(_lib.LIB_PATH)
    assert hasattr(libq, "gpq_queue_device_view") and hasattr(libq, "gpq_queue_stream")
