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


def test_cpp_caller_links_only_declared_symbols():
    """tests/plain_path_caller.cpp builds with g++ against include/gjxl_amd.h
    alone, and every library symbol it imports is declared there."""
    import subprocess

    from gjxl_amd import build

    caller = build.build_caller()
    out = subprocess.run(["nm", "-D", "--undefined-only", str(caller)], check=True,
                         capture_output=True, text=True).stdout
    imported = {line.split()[-1] for line in out.splitlines() if "gjxl_amd_" in line}
    assert {"gjxl_amd_plain_path_run", "gjxl_amd_host_alloc"} <= imported
    assert imported <= set(declared_symbols())
