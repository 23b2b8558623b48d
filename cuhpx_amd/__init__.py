"""cuhpx_amd: MI355X-native HEALPix spherical harmonic transforms and remaps.

Public surface mirrors the reference package's hot path (cuhpx/__init__.py:16-19):
SHTCUDA, iSHTCUDA, VectorSHT, VectoriSHT, Grid, Regridding, ring2nest,
nest2ring, nest2flat, flat2nest, ring2flat, flat2ring, flat2flat and
sht_tools, plus MixedSHT
(mixed-size requests through one HIP graph launch).  The device work is
hand-written HIP for gfx950 in libhpxmi.so, reached through a C ABI
(include/hpxmi.h).
"""

from . import cuhpx_fft, sht_tools
from .hpx_batch import MixedSHT
from .hpx_regrid import Grid, InverseLatLonSHT, LatLonSHT, Regridding
from .hpx_remap import flat2flat, flat2nest, flat2ring, nest2flat, nest2ring, ring2flat, ring2nest
from .hpx_sht import SHTCUDA, ShtPlan, VectoriSHT, VectorSHT, iSHTCUDA

__all__ = [
    "Grid",
    "Regridding",
    "LatLonSHT",
    "InverseLatLonSHT",
    "cuhpx_fft",
    "SHTCUDA",
    "iSHTCUDA",
    "VectorSHT",
    "VectoriSHT",
    "ShtPlan",
    "MixedSHT",
    "ring2nest",
    "nest2ring",
    "nest2flat",
    "flat2nest",
    "ring2flat",
    "flat2ring",
    "flat2flat",
    "sht_tools",
]
