"""CPU: the C-ABI library loads and exports every symbol include/*.h declares.

No compute calls are made (there is no GPU here).
"""
import ctypes
import re
from pathlib import Path

import gjxl_amd

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        names.update(re.findall(r"GJXL_AMD_API\s+[\w\s\*]+?\b(gjxl_amd_\w+)\s*\(",
                                text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(gjxl_amd.LIBRARY_PATH))
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(gjxl_amd.exported_symbols())


def test_library_loads_through_package():
    assert gjxl_amd.library() is not None
