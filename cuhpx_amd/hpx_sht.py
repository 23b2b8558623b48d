"""Spherical harmonic transforms on HEALPix (and other iso-latitude) grids.

Drop-in for cuhpx/hpx_sht.py:515-735: ``SHTCUDA`` / ``iSHTCUDA`` take the
same constructor arguments (nside, lmax, mmax, quad_weights, norm, csphase;
defaults lmax = 4*nside-1, mmax = 2*nside+1, hpx_sht.py:654-655), map
real [..., 12*nside^2] to complex [..., lmax, mmax] (a[l, m], zero for m > l)
and back, raise ValueError on complex SHT input (687-688) and RuntimeError
without a GPU (658-659), and are differentiable through autograd Functions
with hand-written adjoints (515-641).  As in the reference, ``norm`` and
``csphase`` are accepted and ignored (the functions are orthonormal, without
the Condon-Shortley phase).

Differences, all deliberate:
  * the work runs on the caller's current stream (the reference pins the
    stream current at construction, hpx_sht.py:656, 690);
  * the input of the inverse transform is never modified;
  * the backward passes are exact adjoints by default; ``adjoint="compat"``
    reproduces the reference's formulas (0.5x / 2x on m = 0 and Nyquist);
  * extra keyword ``precision`` selects the Legendre arithmetic: "fp32"
    (default, FP32 MFMA), "bf16x3" (3-term split BF16 MFMA, FP32 accumulate)
    or "bf16" (single-term BF16);
  * float64 input computes in FP64 end to end (FP64 ring FFTs, FP64 table and
    GEMMs; ``precision`` applies to float32 input only), returning complex128,
    as the reference does (hpx_fft.cpp:45; hpx_sht.py:557, 599).

All device work is in libhpxmi.so (ring FFTs, Legendre GEMMs, table
generation); this module only allocates outputs and the workspace.
"""

import ctypes
import threading
import weakref

import numpy as np
import torch
import torch.nn as nn

from . import _lib
from .sht_tools import W_helper, healpix_weights

# Device plans shared by every module of the same grid, band limit and device.
# The cache holds them weakly: a plan (and its device tables) is freed when
# the last module, autograd context or caller holding it goes away.
_plan_lock = threading.Lock()
_plans = weakref.WeakValueDictionary()


def _create(fn, *args):
    """Run a plan-creation call; when the device lacks room for the tables,
    return torch's cached free blocks to the device and try once more."""
    rc = fn(*args)
    if rc == _lib.HPX_ERR_UNSUPPORTED and "free" in _lib.last_error():
        torch.cuda.empty_cache()
        rc = fn(*args)
    return rc


class ShtPlan:
    """A device plan of libhpxmi.so for one grid, lmax, mmax and device
    (scalar, or with ``vector`` the tables of the vector transform)."""

    def __init__(self, handle, device, npix, nring, lmax, mmax, vector=False):
        self._handle = handle
        self.device = device
        self.npix = npix
        self.nring = nring
        self.lmax = lmax
        self.mmax = mmax
        self.vector = vector
        self.max_batch = int(_lib.lib().hpx_sht_max_batch(handle))
        self._f64 = False

    @property
    def handle(self):
        return self._handle

    @classmethod
    def healpix(cls, nside, lmax, mmax, quad_weights="ring", device=None):
        device = torch.device("cuda", torch.cuda.current_device() if device is None else device.index)
        key = ("healpix", nside, lmax, mmax, quad_weights, device.index)
        with _plan_lock:
            plan = _plans.get(key)
            if plan is None:
                _, w = healpix_weights(nside, quad_weights)
                w = np.ascontiguousarray(w, dtype=np.float64)
                h = ctypes.c_void_p()
                with torch.cuda.device(device):
                    rc = _create(_lib.lib().hpx_sht_plan_create_healpix,
                                 nside, lmax, mmax, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 _lib.stream_ptr(device), ctypes.byref(h))
                _lib.check(rc, "SHT plan")
                plan = cls(h, device, 12 * nside * nside, 4 * nside - 1, lmax, mmax)
                _plans[key] = plan
        return plan

    @classmethod
    def healpix_vector(cls, nside, lmax, mmax, quad_weights="ring", csphase=True, device=None, norm="ortho",
                       inverse=False):
        """Plan of the vector transform (VectorSHT / VectoriSHT tables).
        ``norm`` scales the tables as the reference's legpoly does; with a
        norm other than "ortho" the forward (VectorSHT) and inverse
        (VectoriSHT, ``inverse=True``) tables differ and are separate plans."""
        device = torch.device("cuda", torch.cuda.current_device() if device is None else device.index)
        code = _lib.norm_code(norm)
        inv = bool(inverse) and code != _lib.NORM_ORTHO
        key = ("healpix-vector", nside, lmax, mmax, quad_weights, bool(csphase), code, inv, device.index)
        with _plan_lock:
            plan = _plans.get(key)
            if plan is None:
                _, w = healpix_weights(nside, quad_weights)
                w = np.ascontiguousarray(w, dtype=np.float64)
                h = ctypes.c_void_p()
                with torch.cuda.device(device):
                    rc = _create(_lib.lib().hpx_vsht_plan_create_healpix,
                                 nside, lmax, mmax, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 int(bool(csphase)), code, int(inv), _lib.stream_ptr(device), ctypes.byref(h))
                _lib.check(rc, "vector SHT plan")
                plan = cls(h, device, 12 * nside * nside, 4 * nside - 1, lmax, mmax, vector=True)
                _plans[key] = plan
        return plan

    @classmethod
    def rings(cls, nphi, theta, phi0, weights, lmax, mmax, device=None, key=None):
        device = torch.device("cuda", torch.cuda.current_device() if device is None else device.index)
        cache_key = None if key is None else ("rings", key, lmax, mmax, device.index)
        with _plan_lock:
            plan = _plans.get(cache_key) if cache_key is not None else None
            if plan is None:
                nphi = np.ascontiguousarray(nphi, dtype=np.int32)
                theta = np.ascontiguousarray(theta, dtype=np.float64)
                phi0 = np.ascontiguousarray(phi0, dtype=np.float64)
                weights = np.ascontiguousarray(weights, dtype=np.float64)
                dptr = ctypes.POINTER(ctypes.c_double)
                h = ctypes.c_void_p()
                with torch.cuda.device(device):
                    rc = _create(
                        _lib.lib().hpx_sht_plan_create_rings, len(nphi),
                        nphi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), theta.ctypes.data_as(dptr), phi0.ctypes.data_as(dptr),
                        weights.ctypes.data_as(dptr), lmax, mmax, _lib.stream_ptr(device),
                        ctypes.byref(h))
                _lib.check(rc, "SHT plan")
                plan = cls(h, device, int(nphi.sum()), len(nphi), lmax, mmax)
                if cache_key is not None:
                    _plans[cache_key] = plan
        return plan

    def __del__(self):
        h = getattr(self, "_handle", None)
        # once the interpreter exits (or this module is torn down) no native call
        if h is not None and h.value and getattr(_lib, "_exiting", True) is False:
            try:
                _lib.lib().hpx_sht_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._handle = None

    @property
    def table_bytes(self):
        return int(_lib.lib().hpx_sht_plan_table_bytes(self._handle))

    def workspace(self, nbatch, f64=False):
        """Workspace for nbatch maps (scalar plan) or nbatch vector fields;
        ``f64``: for the FP64 calls."""
        if f64:
            fn = _lib.lib().hpx_vsht_workspace_bytes_f64 if self.vector else _lib.lib().hpx_sht_workspace_bytes_f64
        else:
            fn = _lib.lib().hpx_vsht_workspace_bytes if self.vector else _lib.lib().hpx_sht_workspace_bytes
        nbytes = int(fn(self._handle, nbatch))
        return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device), nbytes

    def enable_f64(self):
        """Build the plan's FP64 tables (once; allocates and synchronises)."""
        with _plan_lock:
            if not self._f64:
                with torch.cuda.device(self.device):
                    rc = _lib.lib().hpx_sht_plan_enable_f64(self._handle, _lib.stream_ptr(self.device))
                _lib.check(rc, "FP64 plan tables")
                self._f64 = True

    def autotune(self, precision, nbatch):
        """Measure and keep the fastest FFT / GEMM launch shapes for this
        precision and batch (synchronises the current stream)."""
        ws, nbytes = self.workspace(nbatch)
        with torch.cuda.device(self.device):
            rc = _lib.lib().hpx_sht_autotune(self._handle, _lib.PRECISIONS[precision], nbatch, _lib.ptr(ws),
                                             nbytes, _lib.stream_ptr(self.device))
        _lib.check(rc, "SHT autotune")
        return self.tuned_shapes(precision, nbatch)

    def tuned_shapes(self, precision, nbatch):
        """Launch shapes used for (precision, nbatch)."""
        out = (ctypes.c_int32 * 15)()
        rc = _lib.lib().hpx_sht_tuned_shapes(self._handle, _lib.PRECISIONS[precision], nbatch, out)
        _lib.check(rc, "SHT tuned shapes")
        v = list(out)
        return {"fft_analysis_waves": v[0:3], "fft_synthesis_waves": v[3:6], "gemm_fwd_col_tiles": v[6],
                "gemm_inv_col_tiles": v[7], "gemm_inv_row_blocks": v[8],
                "fft_analysis_threads_per_field": v[9:12], "fft_synthesis_threads_per_field": v[12:15]}

    def set_shapes(self, precision, nbatch, shapes):
        """Set the launch shapes used for (precision, nbatch) without timing
        (a dict in the form tuned_shapes returns; values are validated)."""
        v = (list(shapes["fft_analysis_waves"]) + list(shapes["fft_synthesis_waves"])
             + [shapes["gemm_fwd_col_tiles"], shapes["gemm_inv_col_tiles"], shapes["gemm_inv_row_blocks"]]
             + list(shapes["fft_analysis_threads_per_field"]) + list(shapes["fft_synthesis_threads_per_field"]))
        arr = (ctypes.c_int32 * 15)(*v)
        rc = _lib.lib().hpx_sht_set_shapes(self._handle, _lib.PRECISIONS[precision], nbatch, arr)
        _lib.check(rc, "SHT set shapes")

    def _run(self, fn, *args, nbatch, f64=False):
        ws, nbytes = self.workspace(nbatch, f64)
        with torch.cuda.device(self.device):
            rc = fn(self._handle, *args, nbatch, _lib.ptr(ws), nbytes, _lib.stream_ptr(self.device))
        return rc

    def _call(self, fn, head, inp, out, what, f64=False):
        """fn(handle, *head, in, out, nbatch, workspace, bytes, stream) over the
        rows of the contiguous inp and out, at most max_batch rows per call
        (half as many in FP64; one call, with nbatch 0, for an empty batch)."""
        n = inp.shape[0]
        step = max(1, self.max_batch // 2 if f64 else self.max_batch)
        ib = inp.stride(0) * inp.element_size()
        ob = out.stride(0) * out.element_size()
        for s in range(0, max(n, 1), step):
            c = min(step, n - s)
            rc = self._run(fn, *head, ctypes.c_void_p(inp.data_ptr() + s * ib),
                           ctypes.c_void_p(out.data_ptr() + s * ob), nbatch=c, f64=f64)
            _lib.check(rc, what)

    def forward(self, f, precision):
        """f [n, npix] float32 (or float64) contiguous -> alm [n, lmax, mmax]
        complex64 (complex128, computed in FP64)."""
        n = f.shape[0]
        if f.dtype == torch.float64:
            self.enable_f64()
            out = torch.empty((n, self.lmax, self.mmax), dtype=torch.complex128, device=f.device)
            self._call(_lib.lib().hpx_sht_forward_f64, (), f, out, "SHT forward (FP64)", f64=True)
            return out
        out = torch.empty((n, self.lmax, self.mmax), dtype=torch.complex64, device=f.device)
        self._call(_lib.lib().hpx_sht_forward, (_lib.PRECISIONS[precision],), f, out, "SHT forward")
        return out

    def inverse(self, alm, precision):
        """alm [n, lmax, mmax] complex64 (or complex128) contiguous -> f [n, npix]
        float32 (float64, computed in FP64)."""
        n = alm.shape[0]
        if alm.dtype == torch.complex128:
            self.enable_f64()
            out = torch.empty((n, self.npix), dtype=torch.float64, device=alm.device)
            self._call(_lib.lib().hpx_sht_inverse_f64, (), alm, out, "SHT inverse (FP64)", f64=True)
            return out
        out = torch.empty((n, self.npix), dtype=torch.float32, device=alm.device)
        self._call(_lib.lib().hpx_sht_inverse, (_lib.PRECISIONS[precision],), alm, out, "SHT inverse")
        return out

    def vector_forward(self, f, precision):
        """f [n, 2, npix] float32 (float64) contiguous -> alm [n, 2, lmax, mmax]
        complex64 (complex128, computed in FP64)."""
        n = f.shape[0]
        if f.dtype == torch.float64:
            self.enable_f64()
            out = torch.empty((n, 2, self.lmax, self.mmax), dtype=torch.complex128, device=f.device)
            self._call(_lib.lib().hpx_vsht_forward_f64, (), f, out, "vector SHT forward (FP64)", f64=True)
            return out
        out = torch.empty((n, 2, self.lmax, self.mmax), dtype=torch.complex64, device=f.device)
        self._call(_lib.lib().hpx_vsht_forward, (_lib.PRECISIONS[precision],), f, out, "vector SHT forward")
        return out

    def vector_inverse(self, alm, precision):
        """alm [n, 2, lmax, mmax] complex64 (complex128) contiguous -> f [n, 2, npix]
        float32 (float64, computed in FP64)."""
        n = alm.shape[0]
        if alm.dtype == torch.complex128:
            self.enable_f64()
            out = torch.empty((n, 2, self.npix), dtype=torch.float64, device=alm.device)
            self._call(_lib.lib().hpx_vsht_inverse_f64, (), alm, out, "vector SHT inverse (FP64)", f64=True)
            return out
        out = torch.empty((n, 2, self.npix), dtype=torch.float32, device=alm.device)
        self._call(_lib.lib().hpx_vsht_inverse, (_lib.PRECISIONS[precision],), alm, out, "vector SHT inverse")
        return out

    def forward_adjoint(self, galm, precision, mode):
        n = galm.shape[0]
        if galm.dtype == torch.complex128:
            self.enable_f64()
            out = torch.empty((n, self.npix), dtype=torch.float64, device=galm.device)
            self._call(_lib.lib().hpx_sht_forward_adjoint_f64, (_lib.ADJOINT_MODES[mode],), galm, out,
                       "SHT forward adjoint (FP64)", f64=True)
            return out
        out = torch.empty((n, self.npix), dtype=torch.float32, device=galm.device)
        self._call(_lib.lib().hpx_sht_forward_adjoint, (_lib.PRECISIONS[precision], _lib.ADJOINT_MODES[mode]),
                   galm, out, "SHT forward adjoint")
        return out

    def inverse_adjoint(self, gf, precision, mode):
        n = gf.shape[0]
        if gf.dtype == torch.float64:
            self.enable_f64()
            out = torch.empty((n, self.lmax, self.mmax), dtype=torch.complex128, device=gf.device)
            self._call(_lib.lib().hpx_sht_inverse_adjoint_f64, (_lib.ADJOINT_MODES[mode],), gf, out,
                       "SHT inverse adjoint (FP64)", f64=True)
            return out
        out = torch.empty((n, self.lmax, self.mmax), dtype=torch.complex64, device=gf.device)
        self._call(_lib.lib().hpx_sht_inverse_adjoint, (_lib.PRECISIONS[precision], _lib.ADJOINT_MODES[mode]),
                   gf, out, "SHT inverse adjoint")
        return out


def _as_rows(x, width):
    lead = x.shape[:-1]
    return x.reshape(-1, width).contiguous(), lead


def _as_alm_rows(a, lmax, mmax):
    lead = a.shape[:-2]
    return a.reshape(-1, lmax, mmax).contiguous(), lead


class SHTFunction(torch.autograd.Function):
    """Forward SHT with a hand-written adjoint (hpx_sht.py:515-583)."""

    @staticmethod
    def forward(ctx, x, plan, precision, adjoint):
        ctx.plan, ctx.precision, ctx.adjoint = plan, precision, adjoint
        ctx.cdtype = torch.complex128 if x.dtype == torch.float64 else torch.complex64
        rows, lead = _as_rows(x, plan.npix)
        out = plan.forward(rows, precision)
        return out.reshape(*lead, plan.lmax, plan.mmax)

    @staticmethod
    def backward(ctx, grad):
        plan = ctx.plan
        rows, lead = _as_alm_rows(grad.to(ctx.cdtype), plan.lmax, plan.mmax)
        gx = plan.forward_adjoint(rows, ctx.precision, ctx.adjoint)
        return gx.reshape(*lead, plan.npix), None, None, None


class iSHTFunction(torch.autograd.Function):
    """Inverse SHT with a hand-written adjoint (hpx_sht.py:586-641)."""

    @staticmethod
    def forward(ctx, a, plan, precision, adjoint):
        ctx.plan, ctx.precision, ctx.adjoint = plan, precision, adjoint
        ctx.rdtype = torch.float64 if a.dtype == torch.complex128 else torch.float32
        rows, lead = _as_alm_rows(a, plan.lmax, plan.mmax)
        out = plan.inverse(rows, precision)
        return out.reshape(*lead, plan.npix)

    @staticmethod
    def backward(ctx, grad):
        plan = ctx.plan
        rows, lead = _as_rows(grad.to(ctx.rdtype), plan.npix)
        ga = plan.inverse_adjoint(rows, ctx.precision, ctx.adjoint)
        return ga.reshape(*lead, plan.lmax, plan.mmax), None, None, None


def _check_opts(precision, adjoint):
    if precision not in _lib.PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}")
    if adjoint not in _lib.ADJOINT_MODES:
        raise ValueError(f"adjoint must be one of {sorted(_lib.ADJOINT_MODES)}")


class _HealpixSHTBase(nn.Module):
    def __init__(self, nside, lmax=None, mmax=None, quad_weights="ring", norm="ortho", csphase=True,
                 precision="fp32", adjoint="exact"):
        super().__init__()
        self.nside = nside
        self.norm = norm
        self.csphase = csphase
        self.nlat = 4 * nside - 1
        self.nlon = 4 * nside
        self.quad_weights = quad_weights
        self.lmax = lmax or self.nlat
        self.mmax = mmax or (self.nlon // 2 + 1)
        _check_opts(precision, adjoint)
        self.precision = precision
        self.adjoint = adjoint
        if not torch.cuda.is_available():
            raise RuntimeError("CUDA device is not available. This class requires a GPU.")
        self.device = torch.device("cuda")
        _, w = healpix_weights(nside, quad_weights)
        self.register_buffer("weights", torch.from_numpy(w).to(torch.float32).to(self.device),
                             persistent=False)
        self.register_buffer("W", torch.from_numpy(W_helper(w, nside)).to(self.device),
                             persistent=False)
        self._plan = ShtPlan.healpix(nside, self.lmax, self.mmax, quad_weights)

    def plan_for(self, device):
        if device.type != "cuda":
            raise RuntimeError("SHTCUDA / iSHTCUDA need a GPU tensor")
        if self._plan.device.index != device.index:
            self._plan = ShtPlan.healpix(self.nside, self.lmax, self.mmax, self.quad_weights, device)
        return self._plan


class SHTCUDA(_HealpixSHTBase):
    """Forward SHT: real [..., 12 nside^2] (RING) -> complex [..., lmax, mmax]."""

    def forward(self, x):
        if torch.is_complex(x):
            raise ValueError("Input tensor must be real.")
        if x.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"SHTCUDA takes float32 or float64 maps; got {x.dtype}")
        plan = self.plan_for(x.device)
        if x.shape[-1] != plan.npix:
            raise RuntimeError(f"last dimension must be {plan.npix}, got {x.shape[-1]}")
        return SHTFunction.apply(x, plan, self.precision, self.adjoint)


class iSHTCUDA(_HealpixSHTBase):
    """Inverse SHT: complex [..., lmax, mmax] -> real [..., 12 nside^2] (RING)."""

    def forward(self, x):
        if not torch.is_complex(x):
            raise ValueError("Input tensor must be complex.")
        if x.dtype not in (torch.complex64, torch.complex128):
            raise TypeError(f"iSHTCUDA takes complex64 or complex128 coefficients; got {x.dtype}")
        plan = self.plan_for(x.device)
        if tuple(x.shape[-2:]) != (plan.lmax, plan.mmax):
            raise RuntimeError(f"last dimensions must be ({plan.lmax}, {plan.mmax}), got {tuple(x.shape[-2:])}")
        return iSHTFunction.apply(x, plan, self.precision, self.adjoint)


class _HealpixVectorBase(nn.Module):
    _inverse = False

    def __init__(self, nside, lmax, mmax, grid, quad_weights, norm, csphase, use_bluestein, precision):
        super().__init__()
        if grid != "healpix":
            raise ValueError("Unknown quadrature mode")
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}")
        self.nside = nside
        self.grid = grid
        self.norm = norm
        self.csphase = csphase
        self.nlat = 4 * nside - 1
        self.nlon = 4 * nside
        self.quad_weights = quad_weights
        self.use_bluestein = use_bluestein
        self.lmax = lmax or self.nlat
        self.mmax = mmax or (self.nlon // 2 + 1)
        self.precision = precision
        if not torch.cuda.is_available():
            raise RuntimeError("CUDA device is not available. This class requires a GPU.")
        self._plan = ShtPlan.healpix_vector(nside, self.lmax, self.mmax, quad_weights, csphase, norm=norm,
                                            inverse=self._inverse)

    def plan_for(self, device):
        if device.type != "cuda":
            raise RuntimeError("VectorSHT / VectoriSHT need a GPU tensor")
        if self._plan.device.index != device.index:
            self._plan = ShtPlan.healpix_vector(self.nside, self.lmax, self.mmax, self.quad_weights, self.csphase,
                                                device, norm=self.norm, inverse=self._inverse)
        return self._plan


class VectorSHT(_HealpixVectorBase):
    """Forward vector SHT (cuhpx/hpx_sht.py:286-374): real [..., 2, 12 nside^2]
    (theta and phi components, RING) -> complex [..., 2, lmax, mmax]
    (spheroidal, toroidal coefficients).

    Same constructor as the reference (``grid`` must be "healpix";
    ``norm`` scales the tables as the reference's legpoly(norm=...) does,
    sht_tools.py:196-221: "ortho" (default) leaves them, "schmidt" scales
    the rows of degree L by sqrt(4 pi / (2 L + 1)) (L = l for the
    theta-derivative, l + 1 for the m / sin(theta) table, as the reference
    builds them), any other value by sqrt(4 pi); ``use_bluestein`` is
    accepted and, as there, has no effect; ``csphase`` applies the
    Condon-Shortley sign to odd m, default True as in sht_tools.py:302-304),
    plus ``precision`` ("fp32", "bf16x3", "bf16") for the Legendre step.
    Like the reference, whose ring FFTs carry no autograd, the output does not
    track gradients."""

    def __init__(self, nside, lmax=None, mmax=None, grid="healpix", quad_weights="ring", norm="ortho",
                 csphase=True, use_bluestein=False, precision="fp32"):
        super().__init__(nside, lmax, mmax, grid, quad_weights, norm, csphase, use_bluestein, precision)

    def forward(self, x):
        if torch.is_complex(x):
            raise ValueError("Input tensor must be real.")
        if x.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"VectorSHT takes float32 or float64 fields; got {x.dtype}")
        plan = self.plan_for(x.device)
        if x.dim() < 2 or x.shape[-2] != 2 or x.shape[-1] != plan.npix:
            raise RuntimeError(f"input must be [..., 2, {plan.npix}], got {tuple(x.shape)}")
        lead = x.shape[:-2]
        with torch.no_grad():
            out = plan.vector_forward(x.detach().reshape(-1, 2, plan.npix).contiguous(), self.precision)
        return out.reshape(*lead, 2, plan.lmax, plan.mmax)


class VectoriSHT(_HealpixVectorBase):
    """Inverse vector SHT (cuhpx/hpx_sht.py:377-444): complex
    [..., 2, lmax, mmax] (spheroidal, toroidal) -> real [..., 2, 12 nside^2]
    (theta and phi components, RING).  Constructor as the reference's plus
    ``precision``; ``norm`` as in VectorSHT with the reciprocal factors
    (legpoly(inverse=True), hpx_sht.py:396); the output does not track
    gradients."""

    _inverse = True

    def __init__(self, nside, lmax=None, mmax=None, grid="healpix", norm="ortho", csphase=True,
                 use_bluestein=False, precision="fp32"):
        super().__init__(nside, lmax, mmax, grid, "ring", norm, csphase, use_bluestein, precision)

    def forward(self, x):
        if not torch.is_complex(x):
            raise ValueError("Input tensor must be complex.")
        if x.dtype not in (torch.complex64, torch.complex128):
            raise TypeError(f"VectoriSHT takes complex64 or complex128 coefficients; got {x.dtype}")
        plan = self.plan_for(x.device)
        if x.dim() < 3 or tuple(x.shape[-3:]) != (2, plan.lmax, plan.mmax):
            raise RuntimeError(f"input must be [..., 2, {plan.lmax}, {plan.mmax}], got {tuple(x.shape)}")
        lead = x.shape[:-3]
        with torch.no_grad():
            out = plan.vector_inverse(x.detach().reshape(-1, 2, plan.lmax, plan.mmax).contiguous(), self.precision)
        return out.reshape(*lead, 2, plan.npix)
