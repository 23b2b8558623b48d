"""FP64 CPU oracle for the HEALPix SHT and remap hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``cuhpx_amd``)
imports this package; it is used by ``tests/``, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py``'s ``cpu_baseline`` leg.  Every function
restates, in NumPy (FP64) or in PyTorch where the reference itself is the
PyTorch CPU path, an algorithm of NVlabs/cuHPX and cites the reference file and
lines it follows.

Modules
-------
geometry     ring layout, ring quadrature weights, FITS reader
legendre     normalised associated Legendre functions (reference recurrence,
             plus an underflow-scaled FP64 variant)
sht          FP64 forward / inverse SHT and their exact adjoints
remap        RING / NESTED / flat index permutations (integer formulas)
cpu_path     restatement of the reference's CPU torch SHT path (timing baseline)
equiangular  FP64 SHT on an equiangular lat/lon grid (regridding leg)

Pinning
-------
Each function is pinned by tests in ``tests/test_oracle_*.py`` against golden
vectors and mathematical properties.  Functions with no such check are listed
as "parity unpinned" below.

parity unpinned:
  * equiangular.*  -- the lat/lon leg of the reference calls torch_harmonics,
    which is not vendored; only its round trip and the closed-form checks of
    the quadrature are pinned.
"""
