"""The C ABI without a GPU: the library loads, exports every entry point include/ffmm.h
declares (and the Python binding knows each one), and reports argument errors through
status codes and messages before touching a device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "ffmm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ffmm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from ffmm import _lib
    names = declared()
    assert len(names) >= 18
    L = ctypes.CDLL(_lib.LIBPATH)
    for n in names:
        assert hasattr(L, n), f"{n} is declared in include/ffmm.h but not exported"
    assert set(names) == set(_lib.SIGNATURES), "the ctypes binding and the header disagree"
    assert _lib.lib().ffmm_abi_version() == 3


def test_status_codes_and_messages_without_a_device():
    from ffmm import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    # a null output handle is an argument error (FFMM_ERR_ARGUMENT = 1); the device count
    # is queried first, so on a machine without a GPU the call fails with a device error
    rc = L.ffmm_context_create(0, None, None)
    assert rc in (_lib.FFMM_ERR_ARGUMENT, _lib.FFMM_ERR_DEVICE)
    assert L.ffmm_last_error().decode()
    assert L.ffmm_opcache_create(None, 8, 16, 0.1, ctypes.byref(h)) == _lib.FFMM_ERR_ARGUMENT
    assert "null" in L.ffmm_last_error().decode()
    with pytest.raises(_lib.ArgumentError):
        _lib.check(_lib.FFMM_ERR_ARGUMENT)
    with pytest.raises(_lib.DimensionMismatch):
        _lib.check(_lib.FFMM_ERR_DIMENSION)
    # destroying null handles is a no-op
    assert L.ffmm_plan_destroy(None) == 0 and L.ffmm_opcache_destroy(None) == 0 and L.ffmm_context_destroy(None) == 0


def test_host_api_validation():
    import ffmm
    with pytest.raises(ffmm.ArgumentError, match="p must be at least 2"):
        ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=1)
    with pytest.raises(ffmm.ArgumentError, match="pc must be at least p"):
        ffmm.CubeKernelIndependent(ffmm.Laplace3D(), p=8, pc=6)
    with pytest.raises(ffmm.ArgumentError, match="d must lie"):
        ffmm.CubeKernelIndependent(ffmm.Laplace3D(), d=0.9)
    c = ffmm.OperatorCache(ffmm.CubeKernelIndependent(ffmm.Laplace3D()))
    import numpy as np
    with pytest.raises(ffmm.ArgumentError, match="finite"):
        c.fit_box([0, 0, np.nan], [1, 1, 1], (np.zeros(3), 1.0))
    with pytest.raises(ffmm.ArgumentError, match="outside the requested box"):
        c.fit_box([0, 0, 0], [1, 1, 1.5], (np.zeros(3), 1.0))
    o, w, e = c.fit_box([0.1, 0.1, 0.1], [0.9, 0.9, 0.9], (np.zeros(3), 1.0))
    assert w == 1.0 and e == 0 and c.W0 == 1.0
    o, w, e = c.fit_box([0.1, 0.1, 0.1], [0.3, 0.3, 0.3], (np.full(3, 0.1), 0.2))
    assert w == 0.25 and e == -2                       # snapped to ldexp(W0, e) >= W (src/operators/cache.jl:57-83)
    o, w, e = c.fit_box([0, 0, 0], [3, 3, 3], (np.zeros(3), 3.0))
    assert w == 4.0 and e == 2
    origin, W = ffmm.square_bbox([0, 0, 0], [1, 2, 0.5])
    assert W >= 2 and np.all(origin <= 0) and np.all(origin + W >= [1, 2, 0.5])
