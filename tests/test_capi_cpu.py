"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "su2_mi355x.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(su2gpu_\w+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    import su2_amd
    if not os.path.exists(su2_amd.LIB_PATH):
        from su2_amd import build
        build.build()
    return su2_amd.load_library()


def test_header_declares_entry_points():
    names = _declared()
    assert "su2gpu_create" in names and "su2gpu_md_update" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import su2_amd
    assert set(_declared()) <= set(su2_amd.exported_symbols())


@pytest.mark.parametrize("args,msg", [
    ((7, 4, 4, 4, 4, 4, 2.3, 1.0, 0), "group"),
    ((1, 4, 4, 4, 4, 5, 2.3, 1.0, 0), "ndims"),
    ((1, 4, 4, 4, 4, 4, 2.3, 1.0, 3), "precision"),
    ((1, 4, 4, 2, 4, 3, 2.3, 1.0, 0), "Lz"),
    ((1, 4, 0, 4, 4, 4, 2.3, 1.0, 0), "extents"),
    ((1, 4, 4, 4, 4, 4, 2.3, -1.0, 0), "xi"),
    ((1, 2048, 2048, 2048, 2048, 4, 2.3, 1.0, 0), "too large"),
])
def test_create_rejects_bad_arguments(lib, args, msg):
    h = C.c_void_p()
    rc = lib.su2gpu_create(C.byref(h), *args)
    assert rc == 1
    assert msg in lib.su2gpu_last_error().decode()
    assert not h.value


def test_null_handle_is_an_error(lib):
    r = C.c_double()
    assert lib.su2gpu_gauge_action(None, C.byref(r)) == 1
    assert lib.su2gpu_destroy(None) == 0


def test_host_program_builds_and_links():
    """The C++ host program (hmc driver over the adapter headers) is built by build()."""
    from su2_amd import host
    path = host.build()
    assert os.path.exists(path)
