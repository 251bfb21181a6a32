"""The C-ABI library loads on a CPU host and exports every entry point include/*.h declares."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "mocat_amd" / "libmocat_amd.so"


def declared_entry_points():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        names += re.findall(r"^\s*int\s+(mocat_\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_entry_points():
    names = declared_entry_points()
    assert "mocat_random_split" in names and len(names) == len(set(names))


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_python_signatures_cover_header():
    from mocat_amd import _lib
    import mocat_amd  # noqa: F401  (registers every module's signatures)
    missing = [n for n in declared_entry_points() if n not in _lib.SIGNATURES]
    assert not missing, f"declared but not bound in Python: {missing}"


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_abi_version_without_gpu():
    lib = ctypes.CDLL(str(LIB))
    assert lib.mocat_abi_version() == 5
