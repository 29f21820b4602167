"""Text input parsed on the GPU.

* The reference's query files, loaded on the device, give the host loader's
  points bit for bit, and the oracle's (same text, same conversion), with no
  line left to the host.
* Crafted text (CRLF, no final newline, spaces, signs, exponents, 17-digit
  and 19-digit numbers, zeros past the 19th digit) equals the host loader;
  text the device does not convert (hex, inf, 21 significant digits,
  subnormals) is parsed on the host, and the count of such lines says so.
* Malformed files give the host loader's error, naming the same line.
* 2M generated lines: device equals host; labels of device-loaded points equal
  the oracle's.
* Polygon files parsed on the device build the hierarchy of the host loader
  (every array bitwise), including %.17g vertex text; malformed polygon
  files give the host loader's error.
"""
import os
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import datasets
import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host_points(path):
    lon, lat = gahhg_amd.read_query_csv(path)
    return gahhg_amd.query_lonlat_to_xyz(lon, lat)


def device_points(g, path=None, text=None):
    v, n = g.load_query_csv_device(path) if path is not None else g.parse_query_csv_device(text)
    return v.download(np.empty((n, 3))), g.load_info


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.fixture(scope="module")
def g():
    ctx = gahhg_amd.GPISP(device=0)
    yield ctx
    ctx.close()


@pytest.mark.parametrize("name", ["Mars_query.csv", "Square_72_query.csv", "Radiometer_query.csv",
                                  "Tennessee_query.csv"])
def test_reference_query_files(g, name):
    say("device load", name)
    path = os.path.join(GOLD, name)
    xyz, info = device_points(g, path)
    assert same_bits(xyz, host_points(path))
    lon, lat = ol.load_query_csv(path)
    assert same_bits(xyz, ol.query_lonlat_to_xyz(lon, lat))
    assert info["lines"] == len(lon) and info["host_lines"] == 0 and info["bytes"] == os.path.getsize(path)


def crafted():
    lines = ["-102.781724,17.46", "  +1.5,\t-2.25", "1e2,3E-1", ".5,5.", "-0.0,0", "0,-0",
             "12.3456789012345678,-45.678901234567891", "179.99999999999997,89.999999999999986",
             "9007199254740993e-14,1234567890123456789e-18", "100.000000000000000000000,2.50000000000000000000",
             "1.5, 2.5", "7,8,9", "1.25,2.5 trailing", "-180,-90", "360,180", "1e-5,-1E-05"]
    host = ["0x1p3,5", "inf,0", "1,nan", "1234567890.12345678901,2", "1e-320,3", "5,1e400", "1,0x10"]
    return lines, host


def test_crafted_text(g, tmp_path):
    lines, host = crafted()
    for variant in ("lf", "crlf", "nofinal"):
        sep = "\r\n" if variant == "crlf" else "\n"
        text = sep.join(lines + host) + ("" if variant == "nofinal" else sep)
        f = tmp_path / f"q_{variant}.csv"
        f.write_bytes(text.encode())
        xyz, info = device_points(g, text=text.encode())
        ref = host_points(f)
        assert info["lines"] == len(lines) + len(host)
        assert info["host_lines"] == len(host)
        ok = ~np.isnan(ref).any(axis=1)
        assert same_bits(xyz[ok], ref[ok]), variant
        assert np.isnan(xyz[~ok]).any(axis=1).all()
        xyz2, _ = device_points(g, path=f)
        assert same_bits(np.nan_to_num(xyz2), np.nan_to_num(xyz))


@pytest.mark.parametrize("text,line", [(b"1.0,2.0\n1.0;2.0\n3,4\n", 2), (b"1,2\n\n3,4\n", 2), (b"abc\n", 1),
                                       (b"1,2\n3,\n", 2), (b"1,2\n3 4", 2), (b"1,2\n\n", 2),
                                       (b"1,2\n3,4\n5,x\n0x1,2\n", 3)])
def test_malformed_query_text_same_error(g, tmp_path, text, line):
    """The first line sscanf("%lf,%lf") does not read twice is named, as the
    host loader (and the reference's loop) stop there."""
    f = tmp_path / "bad.csv"
    f.write_bytes(text)
    with pytest.raises(gahhg_amd.GahhgError) as host_err:
        gahhg_amd.read_query_csv(f)
    assert host_err.value.code == -3
    with pytest.raises(gahhg_amd.GahhgError) as dev_err:
        g.load_query_csv_device(f)
    assert dev_err.value.code == -3
    assert f": line {line}: wrong format" in str(dev_err.value)


def test_empty_file(g, tmp_path):
    f = tmp_path / "empty.csv"
    f.write_bytes(b"")
    v, n = g.load_query_csv_device(f)
    assert n == 0


def test_two_million_lines_and_labels(g, tmp_path):
    say("device load: 2M generated lines")
    rng = np.random.default_rng(21)
    w = datasets.star_1k(n_queries=2_000_000)
    text = datasets.query_csv_bytes(w["qlon"], w["qlat"], decimals=6)
    f = tmp_path / "big.csv"
    f.write_bytes(text)
    v, n = g.load_query_csv_device(f)
    assert n == 2_000_000 and g.load_info["host_lines"] == 0
    xyz = v.download(np.empty((n, 3)))
    assert same_bits(xyz, host_points(f))
    # 17-digit text of random points
    lon = rng.uniform(-180, 180, 200_000)
    lat = rng.uniform(-90, 90, 200_000)
    t17 = "".join("%.17g,%.17g\n" % (a, b) for a, b in zip(lon, lat)).encode()
    x17, info = device_points(g, text=t17)
    assert info["host_lines"] == 0
    assert same_bits(x17, gahhg_amd.query_lonlat_to_xyz(lon, lat))
    # labels of device-resident points
    g2 = gahhg_amd.GPISP(device=0)
    g2.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    v, n = g2.load_query_csv_device(f)
    dl = g2.alloc(4 * n)
    g2.query_device(v, n, dl)
    lab = dl.download(np.empty(n, dtype=np.int32))
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    assert np.count_nonzero(lab != o.query_xyz(xyz)) == 0
    dl.free()
    g2.close()


def grids_equal(a, b):
    ga, gb = a.grid(), b.grid()
    for key in ("cells", "borders", "refs"):
        assert np.array_equal(ga[key], gb[key]), key
    assert same_bits(ga["v"], gb["v"])


@pytest.mark.parametrize("name", ["Square_72.txt", "Radiometer.txt", "NA0.txt", "Mars_300.txt",
                                  "Tennessee_Tennessee_02999.txt", "Tennessee.txt"])
def test_polygon_device_load(name):
    say("device polygon load", name)
    path = os.path.join(GOLD, name)
    a = gahhg_amd.GPISP(device=0).load_polygon_txt(path).build()
    b = gahhg_amd.GPISP(device=0).load_polygon_txt_device(path).build()
    assert b.load_info["host_lines"] == 0
    grids_equal(a, b)
    assert a.stats()["depth"] == b.stats()["depth"] and a.stats()["n_edges"] == b.stats()["n_edges"]
    a.close()
    b.close()


def test_polygon_device_load_17_digits(tmp_path):
    w = datasets.fractal_100k(n_queries=10, n_outer=20_000)
    f = tmp_path / "fractal.txt"
    datasets.write_polygon_txt(f, w["lon"], w["lat"], w["ring_off"])
    a = gahhg_amd.GPISP(device=0).load_polygon_txt(f).build()
    b = gahhg_amd.GPISP(device=0).load_polygon_txt_device(f).build()
    assert b.load_info["host_lines"] == 0
    grids_equal(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("text", [
    "POLYGON 1 0\n0 5\n0 0\n1 0\n1 1\n0 1\n",               # ring shorter than its count
    "POLYGON 1 0\n0 5\n0 0\n1 0\nx 1\n0 1\n0 0\n",          # bad vertex
    "POLYGON 2 0\n0 4\n0 0\n1 0\n1 1\n0 0\n",               # bad header
    "POLYGON 1 1\n0 4\n0 0\n1 0\n1 1\n0 0\n1 3\n",          # hole with too few edges
    "POLYGON 1 1\n0 5\n0 0\n1 0\n1 1\n0 1\n0 0\n1 4\n.2 .2\n.4 .2\n",  # truncated hole
])
def test_malformed_polygon_same_error(tmp_path, text):
    f = tmp_path / "bad.txt"
    f.write_text(text)
    a = gahhg_amd.GPISP(device=0)
    with pytest.raises(gahhg_amd.GahhgError) as e1:
        a.load_polygon_txt(f)
    with pytest.raises(gahhg_amd.GahhgError) as e2:
        a.load_polygon_txt_device(f)
    assert str(e1.value) == str(e2.value)
    a.close()
