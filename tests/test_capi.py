"""CPU checks of the C ABI: the library loads and exports every declared
symbol; host-only entry points agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(mapp_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from mapp4py_amd import _lib

    L = C.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers every declared symbol
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_version_and_error_text():
    from mapp4py_amd import _lib

    L = _lib.lib()
    assert b"gfx950" in L.mapp_version()
    h = C.c_void_p()
    st = L.mapp_eam_create_setfl(b"/nonexistent/file.eam.alloy", None, 0, 0, C.byref(h))
    assert st == 2
    assert b"cannot open" in L.mapp_last_error()


def test_create_T_matches_reference_stream():
    import mapp4py_amd as M

    elem = (np.arange(500) % 2).astype(np.uint8)
    masses = [63.546, 1.00794]
    v = M.create_temp(elem, masses, 300.0, 17)
    vo = O.o_create_T(elem, masses, 8.617330350e-5, 300.0, 17)
    np.testing.assert_array_equal(v, vo)
