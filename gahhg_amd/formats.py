"""Adapters for the point and polygon files the reference ships for the
external comparators (test/data_for_Lune, README.md:5): one vertex or point
per line as `colatitude_rad,longitude_rad`, the longitude counted from
180 deg West in [0, 2 pi) (lon_deg = degrees(longitude) - 180: the files'
Square_72 and Radiometer are the TXT files' polygons under that rule); a polygon
file is a single closed ring (first vertex repeated) followed by one more
line, a point the polygon contains (the reference point of Ketzner's
containment test).

They convert to and from the reference's own convention -- lon/lat degrees,
lon in [-180, 180), rings closed as in the polygon TXT files
(source/GPISPPolygon.cu:58-186) -- so such files can be loaded through
GPISP.set_polygon_lonlat / query_lonlat like the reference's own data.
"""
import math

import numpy as np


def _read_pairs(path):
    a = np.loadtxt(path, delimiter=",", dtype=np.float64, ndmin=2)
    return a[:, 0], a[:, 1]


def colat_lon_to_lonlat(colat, lon):
    lat = 90.0 - np.degrees(colat)
    lo = np.mod(np.degrees(lon), 360.0) - 180.0
    return lo, lat


def lonlat_to_colat_lon(lon, lat):
    colat = np.radians(90.0 - np.asarray(lat, dtype=np.float64))
    lo = np.radians(np.mod(np.asarray(lon, dtype=np.float64) + 180.0, 360.0))
    return colat, lo


def read_lune_polygon(path, with_reference_point=False):
    """A Lune-format polygon -> (lon, lat, ring_off) in the reference's
    convention (one ring, closing vertex kept); with_reference_point also
    returns the contained point (lon, lat) the file ends with."""
    colat, lon = _read_pairs(path)
    lo, la = colat_lon_to_lonlat(colat, lon)
    # the ring closes where the first vertex comes back; what follows is the
    # reference point
    close = [i for i in range(1, len(lo)) if colat[i] == colat[0] and lon[i] == lon[0]]
    if not close:
        raise ValueError(f"{path}: the ring is not closed")
    end = close[0] + 1
    if end < 4:
        raise ValueError(f"{path}: a ring needs at least 3 vertices")
    ring = (lo[:end].copy(), la[:end].copy(), np.array([0, end], dtype=np.int64))
    if with_reference_point:
        ref = (float(lo[end]), float(la[end])) if end < len(lo) else None
        return ring + (ref,)
    return ring


def read_lune_points(path):
    """Lune-format points -> (lon, lat) degrees."""
    colat, lon = _read_pairs(path)
    return colat_lon_to_lonlat(colat, lon)


def write_lune_points(path, lon, lat):
    colat, lo = lonlat_to_colat_lon(lon, lat)
    with open(path, "w") as f:
        for a, b in zip(colat, lo):
            f.write(f"{a:.16f},{b:.16f}\n")


def write_lune_polygon(path, lon, lat, ring_off, reference_point):
    """The outer ring of a polygon in Lune format (the format has no holes),
    followed by a point it contains (lon, lat degrees)."""
    s, e = int(ring_off[0]), int(ring_off[1])
    lo = np.append(np.asarray(lon[s:e], dtype=np.float64), reference_point[0])
    la = np.append(np.asarray(lat[s:e], dtype=np.float64), reference_point[1])
    write_lune_points(path, lo, la)
