"""CPU-side checks of the C ABI library: it loads and exports every entry point
declared in include/rbc3d.h (no device calls)."""
import os
import re

from tests.conftest import ROOT


def test_header_symbols_exported():
    import rbc3d_amd
    hdr = open(os.path.join(ROOT, "include", "rbc3d.h")).read()
    declared = set(re.findall(r"\b(rbc3d_[a-z0-9_]+)\s*\(", hdr))
    lib = rbc3d_amd.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(rbc3d_amd.exported_symbols())
