"""The C-ABI library loads on the CPU and exports every function that
include/*.h declares (no device calls are made here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b(bamhip_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "bamhip_z4c_rhs" in names and "bamhip_init" in names
    assert len(names) > 20


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "baum_amd", "libbamhip.so")
    if not os.path.exists(path):
        pytest.fail("libbamhip.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
