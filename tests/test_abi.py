"""The C-ABI library loads and exports every entry point include/sse_hip.h
declares (no compute calls: runs without a GPU)."""

import ctypes
import os
import re

from sse import abi
from sse.hip import LIB_PATH, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "sse_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sse_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = load_library()
    declared = _declared()
    assert declared == sorted(abi.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.sse_abi_version() == abi.SSE_ABI_VERSION


def test_descriptor_struct_matches_header_field_order():
    text = open(os.path.join(ROOT, "include", "sse_hip.h")).read()
    body = text[text.index("typedef struct sse_desc {"):text.index("} sse_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z_0-9]*)\s*(?:\[\d+\])?\s*;", body)
    assert names == [f[0] for f in abi.SseDesc._fields_]


def test_create_rejects_bad_descriptor_without_gpu():
    lib = load_library()
    d = abi.SseDesc()
    d.abi_version = 99
    h = ctypes.c_void_p()
    code = lib.sse_create(ctypes.byref(d), ctypes.byref(h))
    assert code == abi.SSE_ERR_INVALID_ARGUMENT
    assert b"ABI" in lib.sse_last_error(None)
    assert not h.value
