"""CPU checks of the boundary: the C ABI library loads and exports every
entry point that include/tmpg.h declares; the Python binding names them all;
the input reader parses the reference's input format.  No GPU needed."""
import os
import re
import ctypes

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tmpg.h")
LIB = os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd", "lib", "libtmpg.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tmpg_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "tmpg_create" in names and "tmpg_step" in names and len(names) >= 25


@pytest.mark.skipif(not os.path.exists(LIB), reason="libtmpg.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import tmpg
    assert sorted(tmpg.EXPORTS) == declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libtmpg.so not built")
def test_phase_cycles_needs_the_stamped_build():
    """the product library has no phase stamps: tmpg_phase_cycles reports
    TMPG_ESTATE (-5) without touching the device, and rejects null arrays"""
    lib = ctypes.CDLL(LIB)
    buf = (ctypes.c_ulonglong * 8)()
    cnt = (ctypes.c_ulonglong * 8)()
    assert lib.tmpg_phase_cycles(buf, cnt, 1) == -5
    assert lib.tmpg_phase_cycles(None, cnt, 0) == -1


@pytest.mark.skipif(not os.path.exists(LIB), reason="libtmpg.so not built")
def test_input_reader(tmp_path):
    lib = ctypes.CDLL(LIB)
    p = tmp_path / "input"
    p.write_text("<time>\nc\t= .45   # light speed\n<grid>\nmx0 = 128\n<particles>\nmaxptl0 = 1e8\n")
    h = ctypes.c_void_p()
    assert lib.tmpg_input_open(str(p).encode(), ctypes.byref(h)) == 0
    d = ctypes.c_double()
    i = ctypes.c_int()
    assert lib.tmpg_input_get_double(h, b"time", b"c", ctypes.c_double(0.0), ctypes.byref(d)) == 0
    assert abs(d.value - 0.45) < 1e-15
    assert lib.tmpg_input_get_int(h, b"grid", b"mx0", 0, ctypes.byref(i)) == 0 and i.value == 128
    assert lib.tmpg_input_get_double(h, b"particles", b"maxptl0", ctypes.c_double(0), ctypes.byref(d)) == 0
    assert d.value == 1e8
    assert lib.tmpg_input_get_int(h, b"grid", b"my0", 77, ctypes.byref(i)) == 0 and i.value == 77
    lib.tmpg_input_close(h)
