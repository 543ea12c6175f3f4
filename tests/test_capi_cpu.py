"""CPU checks of the C-ABI library: it builds, loads without a GPU and
exports every entry point include/ccpprad.h declares."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(REPO, "include", "ccpprad.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ccpprad_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from ccpprad import lib, build
    lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", build.HIP_LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ccpprad_\w+)", out))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    assert set(lib.EXPORTS) <= exported


def test_version_string_without_gpu():
    from ccpprad import lib
    assert b"gfx950" in lib.load().ccpprad_version()
