"""ctypes binding of ``libhpxmi.so`` (the C ABI declared in include/hpxmi.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already loaded (both carry the soname libamdhip64.so.7); device
pointers and streams therefore come straight from torch tensors and
``torch.cuda.current_stream()``.  A missing library raises at first use: there
is no fallback path.
"""

import atexit
import ctypes
import os
import threading

import torch  # noqa: F401  (must be loaded before the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhpxmi.so")

HPX_OK = 0
HPX_ERR_ARG = 1
HPX_ERR_HIP = 2
HPX_ERR_UNSUPPORTED = 3

REMAP_KINDS = {
    "ring2nest": 0,
    "nest2ring": 1,
    "nest2flat": 2,
    "flat2nest": 3,
    "ring2flat": 4,
    "flat2ring": 5,
    "flat2flat": 6,
}

PRECISIONS = {"fp32": 0, "bf16x3": 1, "bf16": 2}
NORM_ORTHO, NORM_FOUR_PI, NORM_SCHMIDT = 0, 1, 2


def norm_code(norm):
    """The reference's norm strings (cuhpx/sht_tools.py:196-221): "ortho",
    "schmidt", and every other value scaled as "four-pi"."""
    if norm == "ortho":
        return NORM_ORTHO
    if norm == "schmidt":
        return NORM_SCHMIDT
    return NORM_FOUR_PI
ADJOINT_MODES = {"exact": 0, "compat": 1}

_lock = threading.Lock()
_lib = None

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); every function of include/hpxmi.h
SIGNATURES = {
    "hpx_last_error": (ctypes.c_char_p, []),
    "hpx_version": (_c_int, []),
    "hpx_remap": (_c_int, [_c_int, _c_i64, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _c_i64, _c_int, _vp]),
    "hpx_copy": (_c_int, [_vp, _vp, _c_i64, _vp]),
    "hpx_sht_plan_create_healpix": (_c_int, [_c_int, _c_int, _c_int, _dp, _vp, ctypes.POINTER(_vp)]),
    "hpx_sht_plan_create_rings": (_c_int, [_c_int, _i32p, _dp, _dp, _dp, _c_int, _c_int, _vp,
                                           ctypes.POINTER(_vp)]),
    "hpx_sht_plan_destroy": (None, [_vp]),
    "hpx_sht_plan_npix": (_c_i64, [_vp]),
    "hpx_sht_plan_table_bytes": (_c_i64, [_vp]),
    "hpx_sht_workspace_bytes": (_c_i64, [_vp, _c_i64]),
    "hpx_sht_max_batch": (_c_i64, [_vp]),
    "hpx_sht_forward": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_inverse": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_forward_noalloc": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp]),
    "hpx_sht_forward_adjoint": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_inverse_adjoint": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_plan_enable_f64": (_c_int, [_vp, _vp]),
    "hpx_sht_workspace_bytes_f64": (_c_i64, [_vp, _c_i64]),
    "hpx_sht_forward_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_inverse_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_forward_adjoint_f64": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_inverse_adjoint_f64": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_vsht_workspace_bytes_f64": (_c_i64, [_vp, _c_i64]),
    "hpx_vsht_forward_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_vsht_inverse_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_ring_rfft_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _vp]),
    "hpx_ring_irfft_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _vp]),
    "hpx_vsht_plan_create_healpix": (_c_int, [_c_int, _c_int, _c_int, _dp, _c_int, _c_int, _c_int, _vp,
                                              ctypes.POINTER(_vp)]),
    "hpx_vsht_workspace_bytes": (_c_i64, [_vp, _c_i64]),
    "hpx_vsht_forward": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_vsht_inverse": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_ring_rfft": (_c_int, [_vp, _vp, _vp, _c_i64, _vp]),
    "hpx_ring_irfft": (_c_int, [_vp, _vp, _vp, _c_i64, _vp]),
    "hpx_sht_stage_ring_analysis": (_c_int, [_vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_stage_legendre_analysis": (_c_int, [_vp, _c_int, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_stage_legendre_synthesis": (_c_int, [_vp, _c_int, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_stage_ring_synthesis": (_c_int, [_vp, _vp, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_autotune": (_c_int, [_vp, _c_int, _c_i64, _vp, _c_i64, _vp]),
    "hpx_sht_tuned_shapes": (_c_int, [_vp, _c_int, _c_i64, _i32p]),
    "hpx_sht_set_shapes": (_c_int, [_vp, _c_int, _c_i64, _i32p]),
    "hpx_sht_mixed_workspace_bytes": (_c_i64, [_c_int, ctypes.POINTER(_vp), ctypes.POINTER(_c_i64)]),
    "hpx_sht_mixed_create": (_c_int, [_c_int, ctypes.POINTER(_vp), ctypes.POINTER(_c_i64), _c_int, _c_int,
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp,
                                      _c_i64, ctypes.POINTER(_vp)]),
    "hpx_sht_mixed_launch": (_c_int, [_vp, _vp]),
    "hpx_sht_mixed_destroy": (None, [_vp]),
}

MIXED_FORWARD = 1
MIXED_INVERSE = 2
MIXED_MAX_PLANS = 16


def lib():
    """The loaded library (loads it on first use)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build it with `python cuhpx_amd/csrc/build.py`")
                handle = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(handle, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = handle
    return _lib


_exiting = False


@atexit.register
def _mark_exit():
    # Registered after torch's own handlers, so it runs before them.  From
    # here on the finalisers of plans and mixed requests make no native
    # calls: the process is ending and the driver reclaims device memory,
    # while the HIP runtime may already be tearing down.
    global _exiting
    _exiting = True


def exiting():
    """True once the interpreter has started to exit."""
    return _exiting


def last_error():
    msg = lib().hpx_last_error()
    return msg.decode() if msg else ""


def check(rc, what):
    """Raise the Python exception matching a C-ABI status code."""
    if rc == HPX_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == HPX_ERR_ARG:
        raise ValueError(msg)
    if rc == HPX_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())
