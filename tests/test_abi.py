"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DECL = re.compile(r"^[A-Za-z_][\w\s\*]*?\b(ymx_\w+)\s*\(", re.M)


def declared_symbols():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(m.group(1) for m in DECL.finditer(text))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "ymx_accel_create" in names and "ymx_accel_intersect" in names
    assert len(names) >= 8


def test_library_exports_all_declared():
    from libyafaray_amd import native
    lib = native.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", str(native.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_symbols() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for n in declared_symbols():
        assert isinstance(getattr(lib, n), ctypes._CFuncPtr)


def test_python_signatures_cover_header():
    from libyafaray_amd import native
    assert sorted(native.SIGNATURES) == [n for n in declared_symbols() if n in native.SIGNATURES]
    assert set(declared_symbols()) <= set(native.SIGNATURES)


def test_argument_errors_need_no_device():
    """Argument validation happens before any device call."""
    from libyafaray_amd import native
    lib = native.lib()
    assert lib.ymx_accel_create(None, 0, None, None, 0, None, None) == native.YMX_OK - 1  # out == NULL
    assert b"out is NULL" in lib.ymx_last_error()
    assert lib.ymx_accel_intersect(None, None, None, 1, None) == -1
    assert lib.ymx_accel_destroy(None) == native.YMX_OK
