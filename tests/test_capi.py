"""The C-ABI library loads on the CPU host and exports every symbol the header declares."""
import ctypes
import os
import re

import pytest

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


@pytest.mark.parametrize("cname,pyname", [("nat_desc", "_Desc"), ("nat_stats_t", "_Stats"),
                                          ("nat_request", "_Request"), ("nat_result", "_Result"),
                                          ("nat_server_opts", "_ServerOpts"), ("nat_server_stats_t", "_ServerStats")])
def test_struct_layouts_match_header(tmp_path, cname, pyname):
    """Each ctypes mirror of a C struct of the header has the C struct's size and field offsets."""
    import shutil
    import subprocess

    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    cls = getattr(nat, pyname)
    fields = [f[0] for f in cls._fields_]
    prog = tmp_path / "layout.c"
    body = "".join(f'  printf("%zu\\n", offsetof({cname}, {f}));\n' for f in fields)
    prog.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "natrium_amd.h"\nint main(void) {\n'
                    f'  printf("%zu\\n", sizeof({cname}));\n' + body + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(cls)
    for f, off in zip(fields, out[1:]):
        assert getattr(cls, f).offset == off, f


def test_unsupported_descriptors_rejected_before_device_use():
    """Descriptor checks run before any device call: SL walls need the general backend, and
    compressible populations use SL streaming."""
    import numpy as np
    import pytest

    from natrium_amd.problems import box_vertices

    v = box_vertices(2, [3, 3], [0, 0], [1, 1])
    with pytest.raises(nat.NatError, match="periodic boundaries on every axis"):
        nat.Solver("D2Q9", 2, [3, 3], np.asarray(v), periodic=[1, 0], backend="structured")
    with pytest.raises(nat.NatError, match="SL streaming"):
        nat.Solver("D2Q25H", 2, [3, 3], v, compressible=True, scheme="SEDG")
