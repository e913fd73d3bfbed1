"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re

import pytest

import g500bfs


def _declared():
    with open(g500bfs.HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(g500_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("g500_graph_generate", "g500_graph_construct", "g500_bfs_create",
                 "g500_bfs_run", "g500_bfs_get_pred", "g500_bfs_validate",
                 "g500_bfs_destroy", "g500_find_roots", "g500_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(g500bfs.LIB_PATH):
        pytest.fail("libg500bfs_hip.so not built")
    lib = C.CDLL(g500bfs.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_handles_are_rejected_without_a_device():
    L = g500bfs.load()
    assert L.g500_graph_construct(None, 0) == -1
    assert b"null graph" in L.g500_last_error()
    assert L.g500_bfs_run(None, 0, 16, 10.0, 14.0, None) == -1
    assert L.g500_bfs_get_pred(None, None) == -1
    assert L.g500_graph_free(None) == 0 and L.g500_bfs_destroy(None) == 0


def test_python_binding_covers_the_header():
    L = g500bfs.load()
    for n in _declared():
        assert getattr(L, n).restype is not None
