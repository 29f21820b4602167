"""Lune/Ketzner-format files (test/data_for_Lune): the library's C++ reader
(gahhg_read_lune_points / gahhg_read_lune_polygon) against the numpy
restatement of the format, bit for bit, on every such file the reference
ships; the format's Square_72 and Radiometer describe the same polygons as
the reference's TXT files, and the oracle labels their points alike."""
import os

import numpy as np
import pytest

from gahhg_amd import formats
import lune_numpy as ln
import oracle_lib as ol

HERE = os.path.dirname(__file__)
LUNE = os.path.join(HERE, "golden", "lune")
REF = os.path.join(HERE, "golden", "reference")
POLYGONS = ["Antarctic.csv.gz", "NA7_nohole.csv.gz", "Radiometer.csv", "Square_72.csv", "Tennessee.csv.gz",
            "Zigzag.csv.gz"]
QUERIES = ["Radiometer_Queries.csv", "Square_72_Queries.csv", "Tennessee_Queries.csv.gz"]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("name", POLYGONS)
def test_polygon_reader_equals_numpy(name, tmp_path):
    path = ln.unpacked(os.path.join(LUNE, name), tmp_path)
    lon, lat, off, ref = formats.read_lune_polygon(path, with_reference_point=True)
    elon, elat, eoff, eref = ln.read_polygon(path)
    assert np.array_equal(bits(lon), bits(elon)) and np.array_equal(bits(lat), bits(elat))
    assert list(off) == list(eoff) and ref == eref
    assert lon[0] == lon[-1] and lat[0] == lat[-1]
    assert np.all(lon >= -180.0) and np.all(lon < 180.0)


@pytest.mark.parametrize("name", QUERIES)
def test_point_reader_equals_numpy(name, tmp_path):
    path = ln.unpacked(os.path.join(LUNE, name), tmp_path)
    lon, lat = formats.read_lune_points(path)
    elon, elat = ln.read_points(path)
    assert len(lon) == 10_000
    assert np.array_equal(bits(lon), bits(elon)) and np.array_equal(bits(lat), bits(elat))


def test_negative_longitudes_wrap_like_numpy(tmp_path):
    # the Tennessee points are stored with negative longitudes; the mod takes
    # the divisor's sign, and exact multiples of 360 deg give +0
    p = tmp_path / "pts.csv"
    vals = [-1.4258267048884783, -np.pi, -2 * np.pi, 0.0, 2 * np.pi, 1e-300, -1e-300, 7.5, -7.5]
    p.write_text("".join(f"{0.5:.17g},{v:.17g}\n" for v in vals))
    lon, _ = formats.read_lune_points(p)
    elon, _ = ln.read_points(p)
    assert np.array_equal(bits(lon), bits(elon))


def test_malformed_polygons_are_refused(tmp_path):
    import gahhg_amd
    p = tmp_path / "open.csv"
    p.write_text("0.1,0.1\n0.2,0.1\n0.2,0.2\n0.1,0.2\n")
    with pytest.raises(gahhg_amd.GahhgError) as e:
        formats.read_lune_polygon(p)
    assert e.value.code == -3
    p.write_text("0.1,0.1\n0.2,0.1\n0.1,0.1\n")
    with pytest.raises(gahhg_amd.GahhgError):
        formats.read_lune_polygon(p)
    p.write_text("0.1,0.1\n0.2;0.1\n")
    with pytest.raises(gahhg_amd.GahhgError):
        formats.read_lune_points(p)


def test_square72_vertices_match_txt():
    lon, lat, off, ref = formats.read_lune_polygon(os.path.join(LUNE, "Square_72.csv"), with_reference_point=True)
    assert list(off) == [0, 5]
    assert ref[1] == 90.0  # the North Pole, inside the square
    # the TXT file lists -135, -45, 45, 135 at latitude 18, in that order
    assert np.allclose(lat, 18.0, atol=1e-12)
    assert np.round(lon[:4], 9).tolist() == [-135.0, -45.0, 45.0, 135.0]


def test_radiometer_vertices_match_txt():
    lon, lat, off = formats.read_lune_polygon(os.path.join(LUNE, "Radiometer.csv"))
    rows = open(os.path.join(REF, "Radiometer.txt")).read().split("\n")[2:2 + 241]
    tl = np.array([[float(v) for v in r.split()[:2]] for r in rows])
    assert len(lon) == len(tl)
    assert np.max(np.abs(lon - tl[:, 0])) < 1e-9 and np.max(np.abs(lat - tl[:, 1])) < 1e-9


def test_tennessee_vertices_match_txt(tmp_path):
    lon, lat, off = formats.read_lune_polygon(ln.unpacked(os.path.join(LUNE, "Tennessee.csv.gz"), tmp_path))
    rows = open(os.path.join(REF, "Tennessee.txt")).read().split("\n")
    n = int(rows[1].split()[1])
    tl = np.array([[float(v) for v in r.split()[:2]] for r in rows[2:2 + n]])
    assert len(lon) == len(tl)
    assert np.max(np.abs(lon - tl[:, 0])) < 1e-9 and np.max(np.abs(lat - tl[:, 1])) < 1e-9


def test_round_trip(tmp_path):
    lon, lat = formats.read_lune_points(os.path.join(LUNE, "Radiometer_Queries.csv"))
    p = tmp_path / "pts.csv"
    formats.write_lune_points(p, lon, lat)
    lo2, la2 = formats.read_lune_points(p)
    assert np.allclose(lo2, lon, atol=1e-12) and np.allclose(la2, lat, atol=1e-12)


def test_labels_match_txt_polygon():
    for poly, pts in (("Square_72", "Square_72_Queries.csv"), ("Radiometer", "Radiometer_Queries.csv")):
        lon, lat, off = formats.read_lune_polygon(os.path.join(LUNE, poly + ".csv"))
        a = ol.Oracle().set_lonlat(lon, lat, off).build()
        b = ol.Oracle().load_txt(os.path.join(REF, poly + ".txt")).build()
        qlon, qlat = formats.read_lune_points(os.path.join(LUNE, pts))
        xyz = ol.query_lonlat_to_xyz(qlon, qlat)
        la, lb = a.query_xyz(xyz), b.query_xyz(xyz)
        # the same region: labels agree except possibly within ~1e-9 deg of
        # an edge (the two files round the vertices differently)
        assert np.mean(la == lb) > 0.999, poly
        assert set(np.unique(la)) <= {1, 2}
        # the file's reference point is inside
        _, _, _, ref = formats.read_lune_polygon(os.path.join(LUNE, poly + ".csv"), with_reference_point=True)
        assert a.query_xyz(ol.query_lonlat_to_xyz(np.array([ref[0]]), np.array([ref[1]])))[0] == 2, poly


def test_polygon_round_trip(tmp_path):
    lon, lat, off, ref = formats.read_lune_polygon(os.path.join(LUNE, "Radiometer.csv"), with_reference_point=True)
    p = tmp_path / "poly.csv"
    formats.write_lune_polygon(p, lon, lat, off, ref)
    lo2, la2, off2, ref2 = formats.read_lune_polygon(p, with_reference_point=True)
    assert list(off2) == list(off)
    assert np.allclose(lo2, lon, atol=1e-12) and np.allclose(la2, lat, atol=1e-12)
    assert np.allclose(ref2, ref, atol=1e-12)
