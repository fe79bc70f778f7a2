"""The C-ABI library loads on the CPU host and exports every symbol the header declares."""
import ctypes
import os
import re

import natrium_amd as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "natrium_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nat_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    decl = _declared()
    assert set(nat.EXPORTS) <= set(decl)
    assert set(decl) <= set(nat.EXPORTS)


def test_library_exports_all_declared_symbols():
    L = ctypes.CDLL(nat.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name


def test_binding_loads_without_gpu():
    L = nat.lib()
    assert L.nat_last_error(None) == b"null context"
