"""C-ABI library checks that need no GPU: it loads, exports every declared
symbol, and its host-side paths match the oracle bit for bit."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gahhg_amd
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gahhg.h")
REF = os.path.join(ROOT, "tests", "golden", "reference")


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(gahhg_amd.LIB_PATH):
        gahhg_amd.build_library()
    return gahhg_amd.lib()


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gahhg_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(L):
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(gahhg_amd.EXPORTED) == names
    out = subprocess.run(["nm", "-D", "--defined-only", gahhg_amd.LIB_PATH], capture_output=True,
                         text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}$", out, re.M), n


def test_library_is_gfx950_code_object(L):
    # the fat binary built here carries a gfx950 code object and no other ISA
    blob = open(gahhg_amd.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert not re.search(rb"amdgcn-amd-amdhsa--gfx9(0[0-9]|4[0-2])", blob)


def test_query_conversion_matches_oracle(L):
    rng = np.random.default_rng(3)
    lon = rng.uniform(-180, 180, 20000)
    lat = rng.uniform(-90, 90, 20000)
    lat[:4] = [90.0, -90.0, 0.0, 1e-20]
    assert (gahhg_amd.query_lonlat_to_xyz(lon, lat) == ol.query_lonlat_to_xyz(lon, lat)).all()


def test_read_query_csv_matches_oracle(L, tmp_path):
    p = os.path.join(REF, "Radiometer_query.csv")
    a = gahhg_amd.read_query_csv(p)
    b = ol.load_query_csv(p)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    bad = tmp_path / "bad.csv"
    bad.write_text("1,2\nfoo\n")
    with pytest.raises(gahhg_amd.GahhgError):
        gahhg_amd.read_query_csv(bad)
    with pytest.raises(gahhg_amd.GahhgError):
        gahhg_amd.read_query_csv(tmp_path / "missing.csv")


def test_read_query_csv_edge_lines_match_oracle(L, tmp_path):
    """The reader's decimal fast path and its sscanf fallback give what the
    oracle's sscanf("%lf,%lf") gives: signs, exponents, more than 19
    significant digits, hex, inf/nan, blanks and CRLF, trailing text."""
    lines = ["1,2", " -0.5,+3e2", "12.5e-3, 7", "1.2345678901234567890123456789,-0.000000000000000000000001",
             "0x1p3,5", "inf,-inf", "45.000000000000000000001,1", "3,4 trailing", "7,8\r", "1e308,-1e-300",
             "0.1,0.2", "-0,0"]
    p = tmp_path / "edge.csv"
    p.write_bytes(("\n".join(lines) + "\n").encode())
    a = gahhg_amd.read_query_csv(p)
    b = ol.load_query_csv(str(p))
    assert len(a[0]) == len(lines) == len(b[0])
    for x, y in ((a[0], b[0]), (a[1], b[1])):
        assert (np.isnan(x) == np.isnan(y)).all()
        ok = ~np.isnan(x)
        assert (x[ok].view(np.int64) == y[ok].view(np.int64)).all()
    for bad_line in ("1.5 ,2", "1,", ",2", "abc,1", ""):
        q = tmp_path / "bad_edge.csv"
        q.write_bytes(("0,0\n" + bad_line + "\n1,1\n").encode())
        with pytest.raises(gahhg_amd.GahhgError):
            gahhg_amd.read_query_csv(q)


def test_write_result_csv(L, tmp_path):
    p = tmp_path / "r.csv"
    gahhg_amd.write_result_csv(p, np.array([2, 1, 0, 2], dtype=np.int32))
    assert p.read_text() == "1\n0\n0\n1\n"


def test_params_default(L):
    p = gahhg_amd.default_params()
    assert (p.max_depth, p.auto_depth, p.block_query) == (15, 1, 256)


def test_create_rejects_bad_params(L):
    h = ctypes.c_void_p()
    p = gahhg_amd.default_params(max_depth=0)
    assert L.gahhg_create(ctypes.byref(h), ctypes.byref(p)) == -1
    p = gahhg_amd.default_params(block_query=100)
    assert L.gahhg_create(ctypes.byref(h), ctypes.byref(p)) == -1


def test_null_context_calls_fail_cleanly(L):
    assert L.gahhg_build(None) == -1
    assert L.gahhg_last_error(None) == b"null context"
