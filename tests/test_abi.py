"""The C-ABI library loads and exports every symbol include/covof_hip.h declares
(no GPU needed: nothing is called)."""
import os
import re

from conftest import ROOT


def test_header_symbols_exported():
    from covof_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "covof_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(covof_\w+)\s*\(", hdr, re.M))
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    L = _lib.load()
    for name in declared:
        assert hasattr(L, name), name
