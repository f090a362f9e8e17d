"""The C-ABI library loads and exports every entry point include/hdgx.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "hdgx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(hdgx_\w+)\s*\(", txt, re.M)))


def test_header_declarations_are_the_runtime_exports():
    from hdg_amd.runtime import EXPORTS
    assert declared() == sorted(EXPORTS)


def test_library_exports_every_declared_symbol():
    from hdg_amd.runtime import LIB
    if not os.path.exists(LIB):
        pytest.skip("libhdgx.so not built")
    lib = ctypes.CDLL(LIB)
    for name in declared():
        assert hasattr(lib, name), name


def test_runtime_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from hdg_amd import runtime
    with pytest.raises(runtime.HdgxError):
        runtime.Context(0)


def _param_counts():
    txt = open(os.path.join(ROOT, "include", "hdgx.h")).read()
    out = {}
    for m in re.finditer(r"^\s*(?:int|const char\*)\s+(hdgx_\w+)\s*\(([^;]*?)\);", txt, re.M | re.S):
        params = [p for p in m.group(2).split(",") if p.strip()]
        out[m.group(1)] = len(params)
    return out


def test_ctypes_signatures_match_header():
    from hdg_amd import runtime
    if not os.path.exists(runtime.LIB):
        pytest.skip("libhdgx.so not built")
    L = runtime.lib()
    for name, n in _param_counts().items():
        assert len(getattr(L, name).argtypes) == n, name
