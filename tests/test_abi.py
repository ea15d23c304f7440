"""The C-ABI library loads and exports every symbol include/vtem_hip.h declares, and the
ctypes signature table matches the header's arity (no GPU calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vtem_hip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    decls = {}
    for m in re.finditer(r"\bint\s+(vt_\w+)\s*\(([^;]*?)\)\s*;", src, flags=re.S):
        args = [a for a in m.group(2).split(",") if a.strip()]
        decls[m.group(1)] = len(args)
    return decls


def test_header_parses():
    d = _declared()
    assert len(d) >= 30


def test_library_exports_every_declared_symbol():
    from vtem_amd._lib import LIB_PATH
    assert os.path.exists(LIB_PATH), "build the library first (python -m vtem_amd.build)"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from vtem_amd._lib import SIGNATURES
    d = _declared()
    for name, nargs in d.items():
        assert name in SIGNATURES, f"{name} missing from the ctypes table"
        assert len(SIGNATURES[name]) == nargs, (name, len(SIGNATURES[name]), nargs)
    extra = set(SIGNATURES) - set(d)
    assert not extra, extra


def test_error_mapping():
    from vtem_amd._lib import check, VtemError
    check(0)
    with pytest.raises(ValueError):
        check(-1, "x")
    with pytest.raises(VtemError):
        check(-4, "x")
    with pytest.raises(VtemError):
        check(700, "x")


def test_split_workspace_sizes_follow_the_library():
    """The host sizes many skinny problems' split-K workspaces at once from the rule the
    library reports; for every K up to 20000 the sizes equal the library's own
    per-problem answer (host-only calls, no GPU)."""
    import ctypes
    import numpy as np
    from vtem_amd.batch import skinny_workspace_array
    from vtem_amd._lib import lib
    K = np.arange(1, 20001, 7)
    got = skinny_workspace_array(np.full(K.size, 37), K, np.full(K.size, 17))
    out = ctypes.c_longlong(0)
    for k, g in zip(K, got):
        lib.vt_gemm_skinny_workspace_size(37, int(k), 17, ctypes.addressof(out))
        assert int(out.value) == int(g), int(k)
