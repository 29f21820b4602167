"""Correctly rounded sin and cos in the lon/lat -> xyz conversions.

* The oracle's certified routine equals the exact-decimal evaluation
  (tests/exact_trig.py) on hard arguments: the doubles nearest k pi/2 and
  multiples of the reference's 3.14159265/2 in degrees, tiny, large, signed
  zero, and random degrees.
* The C library's sin/cos are not correctly rounded everywhere, which is why
  neither side uses them: on a fixed sample glibc differs from the certified
  values, and each difference checked is glibc's error, not the oracle's.
* The product's host conversion (the same source as the device conversion)
  gives the oracle's bits for every point of a 1M-point sample and of the
  reference's query files.
"""
import math
import os

import numpy as np

import exact_trig
import gahhg_amd
import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference")
PI_REF = 3.14159265  # source/GPISPDS.cuh:7


def hard_arguments():
    xs = [0.0, -0.0, 1e-300, -1e-300, 5e-324, 1e-10, 2.0 ** -27, 0.5, 1.0, 3.0, 7.0, 100.0, -1000.0, 12345.678]
    for k in range(1, 9):
        t = k * math.pi / 2
        xs += [t, math.nextafter(t, 0.0), math.nextafter(t, 10.0), -t]
    for deg in (30.0, 45.0, 60.0, 90.0, 135.0, 180.0, -90.0, -180.0, 270.0, 360.0, 89.99999999, 1e-7):
        xs.append(deg * PI_REF / 180.0)
        xs.append(deg * math.pi / 180.0)
    rng = np.random.default_rng(11)
    d = rng.uniform(-360.0, 360.0, 1200)
    xs += (d * PI_REF / 180.0).tolist()
    xs += (np.round(d, 6) * PI_REF / 180.0).tolist()
    return np.array(xs)


def test_oracle_trig_equals_exact_evaluation():
    x = hard_arguments()
    s, c = ol.sincos_cr(x)
    for i, v in enumerate(x):
        es, ec = exact_trig.cr_sin(v), exact_trig.cr_cos(v)
        assert s[i] == es and math.copysign(1, s[i]) == math.copysign(1, es), (v.hex(), s[i], es)
        assert c[i] == ec, (v.hex(), c[i], ec)


def test_glibc_is_not_correctly_rounded_everywhere():
    rng = np.random.default_rng(12)
    x = rng.uniform(-180.0, 180.0, 100_000) * PI_REF / 180.0
    s, c = ol.sincos_cr(x)
    gs = np.sin(x)
    diff = np.nonzero(gs != s)[0]
    assert len(diff) > 20
    for i in diff[:6]:
        assert s[i] == exact_trig.cr_sin(x[i]) != gs[i]
        assert abs(np.int64(np.float64(gs[i]).view(np.int64)) - np.int64(np.float64(s[i]).view(np.int64))) == 1


def test_product_conversion_equals_oracle():
    rng = np.random.default_rng(13)
    lon = np.concatenate([rng.uniform(-180, 180, 1_000_000), np.round(rng.uniform(-180, 180, 200_000), 6),
                          [0.0, -0.0, 180.0, -180.0, 90.0, 360.0]])
    lat = np.concatenate([np.degrees(np.arcsin(rng.uniform(-1, 1, 1_000_000))),
                          np.round(rng.uniform(-90, 90, 200_000), 6), [0.0, -0.0, 90.0, -90.0, 45.0, 89.999999]])
    for name in ("Mars_query.csv", "Square_72_query.csv", "Radiometer_query.csv", "Tennessee_query.csv"):
        qlon, qlat = ol.load_query_csv(os.path.join(GOLD, name))
        lon = np.concatenate([lon, qlon])
        lat = np.concatenate([lat, qlat])
    got = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    exp = ol.query_lonlat_to_xyz(lon, lat)
    bad = np.nonzero((got.view(np.int64) != exp.view(np.int64)).any(axis=1))[0]
    assert len(bad) == 0, (len(bad), lon[bad[:3]], lat[bad[:3]])
