"""Lune/Ketzner-format adapter (test/data_for_Lune): the reference's own
copies of Square_72 and Radiometer in that format describe the same
polygons as its TXT files, and the oracle labels their points alike."""
import os

import numpy as np

from gahhg_amd import formats
import oracle_lib as ol

HERE = os.path.dirname(__file__)
LUNE = os.path.join(HERE, "golden", "lune")
REF = os.path.join(HERE, "golden", "reference")


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
