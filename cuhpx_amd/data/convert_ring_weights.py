"""Convert the HEALPix ring-weight FITS tables into ``ring_weights.npz``.

The reference reads ``cuhpx/data/weight_ring_n{nside:05d}.fits`` with astropy
(cuhpx/sht_tools.py:34-44): HDU 1 is a BINTABLE whose column 0 holds the
"TEMPERATURE WEIGHTS" (the D_weight of each north-hemisphere ring), and the
reference flattens that column.  Small tables store one big-endian float64 per
row (TFORM1 = '1D', 2*nside rows); large tables pack 1024 values per row
(TFORM1 = '1024D', 2*nside/1024 rows).  This script parses the FITS headers
itself (astropy is not a dependency) and writes one float64 array of length
2*nside per nside into ``ring_weights.npz`` (keys ``n00002`` ... ``n08192``).

Run once, from the repository root:
    python cuhpx_amd/data/convert_ring_weights.py /path/to/cuhpx/data
"""

import os
import re
import sys

import numpy as np

BLOCK = 2880
CARD = 80


def _read_header(buf, start):
    """Parse one FITS header starting at byte ``start``.

    Returns (dict of keyword -> raw value string, byte offset of the data).
    """
    cards = {}
    pos = start
    while True:
        card = buf[pos:pos + CARD].decode("ascii", "replace")
        pos += CARD
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            value = card[10:].split("/")[0].strip()
            cards[key] = value.strip("'").strip()
    data_start = ((pos - start + BLOCK - 1) // BLOCK) * BLOCK + start
    return cards, data_start


def read_fits_column0(path):
    """Column 0 of the first BINTABLE extension, flattened, as float64."""
    with open(path, "rb") as fh:
        buf = fh.read()
    primary, data0 = _read_header(buf, 0)
    naxis = int(primary.get("NAXIS", "0"))
    size = 0
    if naxis > 0:
        size = abs(int(primary["BITPIX"])) // 8
        for i in range(1, naxis + 1):
            size *= int(primary[f"NAXIS{i}"])
    ext_start = data0 + ((size + BLOCK - 1) // BLOCK) * BLOCK
    ext, data = _read_header(buf, ext_start)
    if ext.get("XTENSION") != "BINTABLE":
        raise ValueError(f"{path}: HDU 1 is not a BINTABLE")
    row_bytes = int(ext["NAXIS1"])
    nrows = int(ext["NAXIS2"])
    match = re.fullmatch(r"(\d*)D", ext["TFORM1"].strip())
    if match is None:
        raise ValueError(f"{path}: column 0 is not float64 ({ext['TFORM1']})")
    count = int(match.group(1) or "1")
    table = np.frombuffer(buf, dtype=np.uint8, count=row_bytes * nrows, offset=data)
    table = table.reshape(nrows, row_bytes)[:, :8 * count]
    return np.ascontiguousarray(table).view(">f8").astype(np.float64).reshape(-1)


def convert(src_dir, out_path):
    arrays = {}
    for name in sorted(os.listdir(src_dir)):
        match = re.fullmatch(r"weight_ring_n(\d{5})\.fits", name)
        if match is None:
            continue
        nside = int(match.group(1))
        w = read_fits_column0(os.path.join(src_dir, name))
        if w.shape != (2 * nside,):
            raise ValueError(f"{name}: expected {2 * nside} weights, got {w.shape}")
        arrays[f"n{nside:05d}"] = w
    np.savez_compressed(out_path, **arrays)
    return arrays


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/cuhpx/data"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ring_weights.npz")
    got = convert(src, out)
    print(f"wrote {out}: nside {sorted(int(k[1:]) for k in got)}")
