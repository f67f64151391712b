"""The C-ABI library loads and exports every entry point of
include/costuum_amd.h (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "costuum_amd.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(ctm_\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_matches_binding_list():
    from costuum_amd import _native
    assert sorted(_native.EXPORTS) == declared()


def test_library_exports_all_symbols():
    from costuum_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    syms = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in declared() if s not in syms]
    assert not missing, missing
    lib = _native.lib()
    for s in declared():
        assert hasattr(lib, s)


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "costuum_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text
