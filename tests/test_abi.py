"""The C-ABI library loads on a machine without a GPU and exports every declared symbol.
No compute call is made here."""
import ctypes
import re

from pointvortex_amd import backend


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(backend.LIB_PATH))
    names = backend.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    names = set(backend.declared_symbols())
    assert names == set(backend._SIGNATURES), names ^ set(backend._SIGNATURES)


def test_header_is_plain_c():
    text = backend.HEADER_PATH.read_text()
    assert 'extern "C"' in text
    code = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    assert not re.search(r"\b(std::|torch|hipStream_t|template)\b", code)


def test_last_error_is_a_string():
    lib = backend.load_library()
    assert isinstance(lib.pv_last_error(), bytes)


def test_vortex_system_owns_its_arrays():
    """A system copies what it is built from (the reference's struct holds std::vectors), so
    writing a downloaded state into it leaves the caller's arrays unchanged."""
    import numpy as np

    x, y, g = np.arange(4.0), np.ones(4), np.full(4, 0.5)
    s = backend.VortexSystem(x, y, g)
    s.x[:] = -1.0
    assert np.array_equal(x, np.arange(4.0))
    assert s.x.flags["C_CONTIGUOUS"] and s.x.dtype == np.float64
