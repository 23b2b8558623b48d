"""The C ABI without a GPU: the library loads, exports every function that
include/hpxmi.h declares, and rejects bad arguments with the documented codes
before touching the device; the Python layer raises the reference's errors."""

import ctypes
import os
import re

import pytest
import torch

from cuhpx_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hpxmi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hpx_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version():
    assert _lib.lib().hpx_version() >= 100


@pytest.mark.parametrize("nside", [0, 3, 12, -4])
def test_remap_rejects_bad_nside(nside):
    rc = _lib.lib().hpx_remap(0, nside, 0, 0, 0, 0, None, None, 1, 4, None)
    assert rc == _lib.HPX_ERR_ARG
    assert "power of two" in _lib.last_error()


def test_remap_rejects_bad_origin_and_elem_size():
    lib = _lib.lib()
    dummy = ctypes.c_void_p(16)
    other = ctypes.c_void_p(32)
    assert lib.hpx_remap(2, 4, 7, 0, 0, 0, dummy, other, 1, 4, None) == _lib.HPX_ERR_ARG
    assert "S, E, N, W" in _lib.last_error()
    assert lib.hpx_remap(2, 4, 0, 0, 0, 0, dummy, other, 1, 3, None) == _lib.HPX_ERR_ARG
    assert lib.hpx_remap(9, 4, 0, 0, 0, 0, dummy, other, 1, 4, None) == _lib.HPX_ERR_ARG
    assert lib.hpx_remap(0, 4, 0, 0, 0, 0, dummy, dummy, 1, 4, None) == _lib.HPX_ERR_ARG
    # zero rows is a no-op
    assert lib.hpx_remap(0, 4, 0, 0, 0, 0, None, None, 0, 4, None) == _lib.HPX_OK


def test_plan_rejects_bad_arguments():
    lib = _lib.lib()
    h = ctypes.c_void_p()
    w = (ctypes.c_double * 7)(*([1.0] * 7))
    assert lib.hpx_sht_plan_create_healpix(0, 5, 5, w, None, ctypes.byref(h)) == _lib.HPX_ERR_ARG
    assert lib.hpx_sht_plan_create_healpix(2, -1, 5, w, None, ctypes.byref(h)) == _lib.HPX_ERR_ARG
    assert lib.hpx_sht_plan_create_healpix(2, 5, 0, w, None, ctypes.byref(h)) == _lib.HPX_ERR_ARG
    # asymmetric ring grid
    nphi = (ctypes.c_int32 * 3)(4, 8, 6)
    th = (ctypes.c_double * 3)(0.5, 1.5, 2.6)
    ph = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    assert lib.hpx_sht_plan_create_rings(3, nphi, th, ph, w, 4, 4, None, ctypes.byref(h)) == _lib.HPX_ERR_ARG
    assert "symmetric" in _lib.last_error()
    # compute calls validate the plan and precision before any launch
    assert lib.hpx_sht_forward(None, 0, None, None, 1, None, 0, None) == _lib.HPX_ERR_ARG
    assert lib.hpx_sht_workspace_bytes(None, 1) == -1
    assert lib.hpx_sht_max_batch(None) == -1


def test_check_maps_codes_to_exceptions():
    with pytest.raises(ValueError):
        _lib.check(_lib.HPX_ERR_ARG, "x")
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.HPX_ERR_UNSUPPORTED, "x")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.HPX_ERR_HIP, "x")


def test_python_remap_errors_match_reference():
    import cuhpx_amd
    x = torch.zeros(12 * 9)
    with pytest.raises(ValueError, match="power of two"):
        cuhpx_amd.nest2ring(x, 3)
    with pytest.raises(RuntimeError, match="number of elements"):
        cuhpx_amd.nest2ring(torch.zeros(13), 1)
    with pytest.raises(RuntimeError, match="S, E, N, W"):
        cuhpx_amd.nest2flat(torch.zeros(12), "Q", False, 1)
    with pytest.raises(RuntimeError, match="S, E, N, W"):
        cuhpx_amd.flat2flat(torch.zeros(12), "S", False, "Z", False, 1)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error")
def test_sht_modules_need_a_gpu():
    import cuhpx_amd
    with pytest.raises(RuntimeError, match="GPU"):
        cuhpx_amd.SHTCUDA(4)
    with pytest.raises(RuntimeError, match="GPU"):
        cuhpx_amd.iSHTCUDA(4)


def test_mixed_and_autotune_reject_bad_arguments():
    lib = _lib.lib()
    h = ctypes.c_void_p()
    plans = (ctypes.c_void_p * 1)(None)
    counts = (ctypes.c_int64 * 1)(4)
    assert lib.hpx_sht_mixed_workspace_bytes(1, plans, counts) == -1
    assert lib.hpx_sht_mixed_workspace_bytes(0, None, None) == 0
    assert lib.hpx_sht_mixed_create(0, plans, counts, 1, 3, None, None, None, None, 0,
                                    ctypes.byref(h)) == _lib.HPX_ERR_ARG
    assert lib.hpx_sht_mixed_create(_lib.MIXED_MAX_PLANS + 1, plans, counts, 1, 3, None, None, None, None, 0,
                                    ctypes.byref(h)) == _lib.HPX_ERR_ARG
    assert lib.hpx_sht_mixed_create(1, plans, counts, 1, 0, None, None, None, None, 0,
                                    ctypes.byref(h)) == _lib.HPX_ERR_ARG
    assert lib.hpx_sht_mixed_launch(None, None) == _lib.HPX_ERR_ARG
    lib.hpx_sht_mixed_destroy(None)
    assert lib.hpx_sht_autotune(None, 1, 4, None, 0, None) == _lib.HPX_ERR_ARG
    shapes = (ctypes.c_int32 * 15)()
    assert lib.hpx_sht_tuned_shapes(None, 1, 4, shapes) == _lib.HPX_ERR_ARG
