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


def test_fortran_binding_matches_header():
    """Every bind(C) name in the Fortran interface module is a declared entry point,
    and the module compiles (amdflang, host code only)."""
    import subprocess
    import tempfile
    hdr = open(os.path.join(ROOT, "include", "rbc3d.h")).read()
    declared = set(re.findall(r"\b(rbc3d_[a-z0-9_]+)\s*\(", hdr))
    src = os.path.join(ROOT, "rbc3d_amd", "fortran", "rbc3d_bind.f90")
    bound = set(re.findall(r'bind\(C, name="(rbc3d_[a-z0-9_]+)"\)', open(src).read()))
    assert bound and bound <= declared
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([fc, "-c", src, "-J", d, "-o", os.path.join(d, "m.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
