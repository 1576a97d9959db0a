"""CPU checks of the C-ABI library: it loads and exports every entry point
declared in include/mom_gpu.h (no compute calls without a GPU)."""
import os
import re

from conftest import ROOT
from mom2amd import capi


def header_functions(name="mom_gpu.h"):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mom_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    L = capi.lib()
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(capi.EXPORTS) == names


def test_library_exports_test_hooks():
    L = capi.lib()
    names = header_functions("mom_gpu_testing.h")
    assert sorted(names) == sorted(capi.TEST_EXPORTS)
    assert all(hasattr(L, n) for n in names)
