"""The C-ABI library loads and exports every function include/fibgpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from fib_md_amd import _lib

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "fibgpu.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fib_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    assert os.path.exists(_lib.LIBPATH), "build libfibgpu.so first"
    L = ctypes.CDLL(_lib.LIBPATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes binding covers the header
    assert set(names) == set(_lib.SIGNATURES)


def test_default_params_are_the_reference_constants():
    p = _lib.fib_params()
    _lib.lib().fib_default_params(ctypes.byref(p))
    assert p.sw_A == 7.049556277 and p.sw_B == 0.6022245584 and p.sw_a == 1.8
    assert p.eps_si == 3.473928e-19 and p.sigma_si == 2.0951e-10
    assert p.listfac == 1.1 and p.ntable == 12000
    assert list(p.mol_c) == [0.35, 0.55, 0.10] and list(p.mol_b) == [0.3, 1.2, 6.0]
