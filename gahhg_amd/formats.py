"""Lune/Ketzner-format files (test/data_for_Lune, README.md:5): the format of
the comparators the reference was measured against, one point per line as
`colatitude_rad,longitude_rad` with the longitude counted from 180 deg West
in [0, 2 pi) (lon_deg = degrees(longitude) - 180, taken mod 360).  A polygon
file is a single closed ring (first vertex repeated) followed by one more
line, a point the comparators take as inside.

Reading is the library's (gahhg_read_lune_points, gahhg_read_lune_polygon,
gahhg_load_polygon_lune in include/gahhg.h); the result is in the
reference's own convention -- lon/lat degrees, lon in [-180, 180), rings
closed as in the polygon TXT files (source/GPISPPolygon.cu:58-186) -- so it
goes through GPISP.set_polygon_lonlat / query_lonlat like the reference's own
data.  The writers below produce the format from that convention.
"""
import numpy as np

from . import read_lune_points as _read_points, read_lune_polygon as _read_polygon


def lonlat_to_colat_lon(lon, lat):
    colat = np.radians(90.0 - np.asarray(lat, dtype=np.float64))
    lo = np.radians(np.mod(np.asarray(lon, dtype=np.float64) + 180.0, 360.0))
    return colat, lo


def read_lune_polygon(path, with_reference_point=False):
    """A Lune-format polygon -> (lon, lat, ring_off) in the reference's
    convention (one ring, closing vertex kept); with_reference_point also
    returns the contained point (lon, lat) the file ends with, or None."""
    lon, lat, off, ref = _read_polygon(path)
    return (lon, lat, off, ref) if with_reference_point else (lon, lat, off)


def read_lune_points(path):
    """Lune-format points -> (lon, lat) degrees."""
    return _read_points(path)


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
