"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU needed)."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(cd_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from inmyroom_vulkan_amd import collide
    lib = ctypes.CDLL(collide.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    assert declared == set(collide.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_python_binding_loads():
    from inmyroom_vulkan_amd import load_library
    assert load_library() is not None
