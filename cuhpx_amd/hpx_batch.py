"""Mixed-size SHT requests: several HEALPix sizes, each with its own batch of
maps, served by one HIP graph launch.

The reference serves mixed sizes with one SHTCUDA / iSHTCUDA call per size
(cuhpx/hpx_sht.py:685-691, 732-735); its FFT plan cache holds one (n, nside)
and is rebuilt whenever either changes (src/harmonic_transform/hpx_fft.cpp:
49-58).  Here every size keeps its own device plan, the launch shapes of each
size are autotuned on the device (``ShtPlan.autotune``), and the whole request
(ring analysis, Legendre analysis, Legendre synthesis, ring synthesis for all
sizes, sizes as concurrent graph branches) is captured once into a HIP graph
by ``hpx_sht_mixed_create``; a call is one ``hipGraphLaunch``.

The graph is bound to static device buffers (``maps_in``, ``alm``,
``maps_out``, one tensor per size) owned by this object: write inputs into
them (or pass tensors to ``__call__``, which copies), launch, read outputs.
Outputs are overwritten by the next launch.
"""

import ctypes

import torch

from . import _lib
from .hpx_sht import ShtPlan

_DIRECTIONS = {"forward": _lib.MIXED_FORWARD, "inverse": _lib.MIXED_INVERSE,
               "both": _lib.MIXED_FORWARD | _lib.MIXED_INVERSE}


class MixedSHT:
    """One graph over several (nside, nbatch[, lmax, mmax]) requests.

    sizes      list of tuples (nside, nbatch) or (nside, nbatch, lmax, mmax);
               lmax / mmax default to the reference's 4*nside-1 / 2*nside+1
    precision  "fp32", "bf16x3" or "bf16" (Legendre arithmetic)
    direction  "forward" (maps -> alm), "inverse" (alm -> maps) or "both"
               (maps -> alm -> maps)
    autotune   measure each size's launch shapes before capturing
    """

    def __init__(self, sizes, precision="bf16x3", direction="both", autotune=True, quad_weights="ring",
                 device=None):
        if not torch.cuda.is_available():
            raise RuntimeError("MixedSHT needs a GPU")
        if not 1 <= len(sizes) <= _lib.MIXED_MAX_PLANS:
            raise ValueError(f"between 1 and {_lib.MIXED_MAX_PLANS} sizes")
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}")
        if direction not in _DIRECTIONS:
            raise ValueError(f"unknown direction {direction!r}")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else
                                   torch.device(device).index)
        self.precision = precision
        self.direction = direction
        self.specs = []
        self.plans = []
        for s in sizes:
            nside, nbatch = int(s[0]), int(s[1])
            lmax = int(s[2]) if len(s) > 2 and s[2] is not None else 4 * nside - 1
            mmax = int(s[3]) if len(s) > 3 and s[3] is not None else 2 * nside + 1
            if nbatch < 1:
                raise ValueError("every size needs nbatch >= 1")
            self.specs.append((nside, nbatch, lmax, mmax))
            self.plans.append(ShtPlan.healpix(nside, lmax, mmax, quad_weights, self.device))
        self.tuned = []
        if autotune:
            for plan, (_, nbatch, _, _) in zip(self.plans, self.specs):
                self.tuned.append(plan.autotune(precision, nbatch))
        else:
            for plan, (_, nbatch, _, _) in zip(self.plans, self.specs):
                self.tuned.append(plan.tuned_shapes(precision, nbatch))
        dev = self.device
        self.maps_in = [torch.zeros((b, 12 * n * n), dtype=torch.float32, device=dev) for n, b, _, _ in self.specs]
        self.alm = [torch.zeros((b, lm, mm), dtype=torch.complex64, device=dev) for _, b, lm, mm in self.specs]
        self.maps_out = [torch.zeros((b, 12 * n * n), dtype=torch.float32, device=dev)
                         for n, b, _, _ in self.specs]
        k = len(self.specs)
        handles = (ctypes.c_void_p * k)(*[p.handle.value for p in self.plans])
        counts = (ctypes.c_int64 * k)(*[b for _, b, _, _ in self.specs])
        nbytes = int(_lib.lib().hpx_sht_mixed_workspace_bytes(k, handles, counts))
        if nbytes < 0:
            raise ValueError("bad mixed request")
        self._workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        ins = (ctypes.c_void_p * k)(*[t.data_ptr() for t in self.maps_in])
        alms = (ctypes.c_void_p * k)(*[t.data_ptr() for t in self.alm])
        outs = (ctypes.c_void_p * k)(*[t.data_ptr() for t in self.maps_out])
        torch.cuda.synchronize(dev)
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            rc = _lib.lib().hpx_sht_mixed_create(k, handles, counts, _lib.PRECISIONS[precision],
                                                 _DIRECTIONS[direction], ins, alms, outs,
                                                 _lib.ptr(self._workspace), nbytes, ctypes.byref(h))
        _lib.check(rc, "mixed SHT request")
        self._handle = h

    def launch(self):
        """Run the captured request once on the current stream."""
        with torch.cuda.device(self.device):
            rc = _lib.lib().hpx_sht_mixed_launch(self._handle, _lib.stream_ptr(self.device))
        _lib.check(rc, "mixed SHT launch")

    def __call__(self, inputs=None):
        """Copy ``inputs`` (maps for "forward" / "both", alm for "inverse",
        one tensor per size) into the static buffers, launch, and return the
        static outputs: (alm, maps_out) lists."""
        if inputs is not None:
            dst = self.alm if self.direction == "inverse" else self.maps_in
            if len(inputs) != len(dst):
                raise ValueError(f"expected {len(dst)} inputs")
            for d, x in zip(dst, inputs):
                d.copy_(x.reshape(d.shape))
        self.launch()
        return self.alm, self.maps_out

    def __del__(self):
        h = getattr(self, "_handle", None)
        # once the interpreter exits (or this module is torn down) no native call
        if h is not None and h.value and getattr(_lib, "_exiting", True) is False:
            try:
                _lib.lib().hpx_sht_mixed_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._handle = None
