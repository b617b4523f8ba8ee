"""The C-ABI library loads and exports every function include/kira_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from tests.kira_pkg import ROOT, load


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(kira_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    pkg = load()
    lib = ctypes.CDLL(pkg.LIB_PATH)
    declared = _declared()
    assert "kira_sieve_run_batch" in declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    pkg = load()
    assert _declared() == set(pkg.SIGNATURES)


def test_no_device_is_reported_cleanly():
    pkg = load()
    # On a host without a GPU this is False; with one it is True.  Either way
    # it must not raise.
    assert pkg.is_available() in (True, False)
