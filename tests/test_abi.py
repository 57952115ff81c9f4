"""CPU-only checks of the C ABI: the library loads and exports every symbol
include/endgraph.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(eg_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "eg_graph_build" in names and "eg_last_error" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    import endgraph_amd
    if not os.path.exists(endgraph_amd.LIB_PATH):
        pytest.skip("libendgraph.so not built")
    lib = ctypes.CDLL(endgraph_amd.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_fails_loudly_without_library(monkeypatch, tmp_path):
    import endgraph_amd
    monkeypatch.setattr(endgraph_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(endgraph_amd, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        endgraph_amd.lib()


def test_error_path_without_gpu():
    """Argument errors are reported through status codes, never aborts."""
    import endgraph_amd as eg
    if not os.path.exists(eg.LIB_PATH):
        pytest.skip("libendgraph.so not built")
    h = ctypes.c_void_p()
    rc = eg.lib().eg_graph_create(0, 1, 1, 0, 1, 8, ctypes.byref(h))
    assert rc == -1 and b"vertices_num" in eg.lib().eg_last_error()
    rc = eg.lib().eg_graph_create(100, 4, 9, 0, 1, 8, ctypes.byref(h))
    assert rc == -1
    rc = eg.lib().eg_graph_create(100, 1, 1, 0, 1, 10, ctypes.byref(h))
    assert rc == -1
    rc = eg.lib().eg_graph_build(None, None, 0, 0, None)
    assert rc == -1


def test_edge_file_records_host_only(tmp_path):
    """eg_edge_file_records only stats the file (no device needed)."""
    import ctypes as C
    import endgraph_amd as eg
    p = tmp_path / "e.bin"
    p.write_bytes(bytes(8 * 5))
    n = C.c_uint64()
    assert eg.lib().eg_edge_file_records(str(p).encode(), 8, C.byref(n)) == 0 and n.value == 5
    assert eg.lib().eg_edge_file_records(str(p).encode(), 12, C.byref(n)) != 0  # 40 B is not whole 12 B records
    assert eg.lib().eg_edge_file_records(str(tmp_path / "none").encode(), 8, C.byref(n)) != 0
