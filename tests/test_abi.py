"""The C-ABI libraries load on a machine without a GPU, export every entry
point their headers declare, and agree with the ctypes struct layouts.
No compute calls are made here."""
import ctypes as C
import re
from pathlib import Path

import pytest

import cpvulkan_amd
from cpvulkan_amd import abi

ROOT = Path(__file__).resolve().parent.parent


def declared(header: Path, prefix: str) -> list[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(rf"\b({prefix}\w+)\s*\(", text)) - {"cpv_context", "cpv_program"})


def test_libcpv_exports_every_declared_symbol():
    lib = C.CDLL(str(cpvulkan_amd.LIBCPV))
    names = declared(ROOT / "include" / "cpv.h", "cpv_")
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_library():
    lib = cpvulkan_amd.load_library()
    abi.check_layout(lib)


def test_last_error_is_empty_string_initially():
    lib = cpvulkan_amd.load_library()
    assert isinstance(lib.cpv_last_error(), bytes)


def test_context_create_rejects_bad_ordinal():
    lib = cpvulkan_amd.load_library()
    h = C.c_void_p()
    rc = lib.cpv_context_create(9999, C.byref(h))
    assert rc != 0 and not h.value
