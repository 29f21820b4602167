"""Lune/Ketzner-format files read with numpy alone: the independent check of
the library's C++ reader (gahhg_read_lune_points / gahhg_read_lune_polygon).

The reference ships copies of its polygons and query points in the format of
the comparators it was measured against (README.md:5, test/data_for_Lune):
one `colatitude,longitude` pair per line in radians, the longitude counted
from 180 deg West, so lat = 90 - degrees(colat), lon = (degrees(lon) mod 360)
- 180.  A polygon file is a single ring closed by repeating its first vertex,
followed by one more line, a point the comparators take as inside.  Under
this rule the format's Square_72, Radiometer and Tennessee are the TXT files'
polygons (test_formats.py).
"""
import gzip
import os
import shutil

import numpy as np


def unpacked(path, tmpdir):
    """path itself, or for a .gz fixture its contents written under tmpdir."""
    path = str(path)
    if not path.endswith(".gz"):
        return path
    out = os.path.join(str(tmpdir), os.path.basename(path)[:-3])
    if not os.path.exists(out):
        with gzip.open(path, "rb") as f, open(out, "wb") as g:
            shutil.copyfileobj(f, g)
    return out


def _read_pairs(path):
    a = np.loadtxt(path, delimiter=",", dtype=np.float64, ndmin=2)
    return a[:, 0], a[:, 1]


def colat_lon_to_lonlat(colat, lon):
    lat = 90.0 - np.degrees(colat)
    lo = np.mod(np.degrees(lon), 360.0) - 180.0
    return lo, lat


def read_points(path):
    colat, lon = _read_pairs(path)
    return colat_lon_to_lonlat(colat, lon)


def read_polygon(path):
    """(lon, lat, ring_off, ref): one ring with its closing vertex, and the
    point after it (lon, lat) or None."""
    colat, lon = _read_pairs(path)
    lo, la = colat_lon_to_lonlat(colat, lon)
    close = [i for i in range(1, len(lo)) if colat[i] == colat[0] and lon[i] == lon[0]]
    if not close:
        raise ValueError(f"{path}: the ring is not closed")
    end = close[0] + 1
    if end < 4:
        raise ValueError(f"{path}: a ring needs at least 3 vertices")
    ref = (float(lo[end]), float(la[end])) if end < len(lo) else None
    return lo[:end].copy(), la[:end].copy(), np.array([0, end], dtype=np.int64), ref
