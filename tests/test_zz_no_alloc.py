"""fib_force_noalloc (the allocation-free force evaluation) against fib_force
of a one-instance handle, through the C-ABI: the same kernels on the same
inputs give the same forces and energy bit for bit.  Two calls in a row,
the second with a smaller configuration (the shared device array keeps the
first call's contents, which the second must not see), in both precisions;
sizes beyond the array are rejected with a message."""
import ctypes as C

import numpy as np
import pytest

from tests import gpu_common as gc

pytestmark = pytest.mark.gpu


def _handle_force(X, at, L, nsg, nlj, impact, prec):
    from fib_md_amd import FibGPU
    g = FibGPU([(len(X), nsg, nlj, L)], precision=prec)
    g.set_state(X, atype=at)
    g.build_nlist()
    r = g.force(impact=impact, energy=True)
    g.close()
    return r["F"], r["utot"]


def _hip():
    """The HIP runtime the library itself links (device memory for the
    caller's inputs and outputs)."""
    import os
    from fib_md_amd import _lib
    _lib.lib()
    H = C.CDLL(os.path.join("/opt/rocm/lib", "libamdhip64.so.7"))
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    return H


def _noalloc_force(X, at, L, nsg, nlj, impact, prec):
    from fib_md_amd import _lib
    from fib_md_amd.gpu import default_params
    L_ = _lib.lib()
    H = _hip()
    n = len(X)
    Xh = np.asfortranarray(np.asarray(X, np.float64)).ravel(order="F").copy()
    ah = np.ascontiguousarray(at, np.int32)
    Fh = np.full(3 * n, np.nan)
    uh = np.zeros(1)
    ptr = {}
    for name, nbytes in (("X", Xh.nbytes), ("a", ah.nbytes), ("F", Fh.nbytes), ("u", 8)):
        q = C.c_void_p()
        assert H.hipMalloc(C.byref(q), nbytes) == 0
        ptr[name] = q
    try:
        assert H.hipMemcpy(ptr["X"], Xh.ctypes.data, Xh.nbytes, 1) == 0
        assert H.hipMemcpy(ptr["a"], ah.ctypes.data, ah.nbytes, 1) == 0
        assert H.hipMemcpy(ptr["F"], Fh.ctypes.data, Fh.nbytes, 1) == 0
        lb = np.asarray(L, np.float64)
        p = default_params()
        prec_id = {"fp64": _lib.FIB_PREC_FP64, "mixed": _lib.FIB_PREC_MIXED}[prec]
        rc = L_.fib_force_noalloc(C.byref(p), prec_id, n, nsg, nlj,
                                  lb.ctypes.data_as(C.c_void_p), impact, ptr["X"], ptr["a"],
                                  ptr["F"], ptr["u"], None)
        assert rc == 0, (rc, L_.fib_last_error(None))
        assert H.hipDeviceSynchronize() == 0
        assert H.hipMemcpy(Fh.ctypes.data, ptr["F"], Fh.nbytes, 2) == 0
        assert H.hipMemcpy(uh.ctypes.data, ptr["u"], 8, 2) == 0
    finally:
        for q in ptr.values():
            H.hipFree(q)
    return np.ascontiguousarray(Fh.reshape(3, n).T), float(uh[0])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_equals_handle_force(prec):
    big = gc.crystal_with_ions(4, 0.12e-10, 7, 3)     # 512 Si + 3 Ga
    small = gc.crystal_with_ions(3, 0.10e-10, 11, 2)  # 216 Si + 2 Ga
    for X, at, L, nsg, nlj in (big, small):
        print(f"{prec}: {len(X)} atoms", flush=True)
        Fh, uh = _handle_force(X, at, L, nsg, nlj, nlj, prec)
        Fn, un = _noalloc_force(X, at, L, nsg, nlj, nlj, prec)
        assert np.all(np.isfinite(Fn))
        np.testing.assert_array_equal(Fn, Fh)
        assert un == uh


def test_noalloc_rejects_too_many_atoms():
    from fib_md_amd import _lib
    from fib_md_amd.gpu import default_params
    L_ = _lib.lib()
    p = default_params()
    lb = np.full(3, 1e-8)
    n = _lib.FIB_NOALLOC_MAX_ATOMS + 1
    dummy = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = L_.fib_force_noalloc(C.byref(p), 0, n, n, 0, lb.ctypes.data_as(C.c_void_p), 0,
                              dummy, dummy, dummy, None, None)
    assert rc == _lib.FIB_ERR_ARG
    assert b"at most" in L_.fib_last_error(None)
