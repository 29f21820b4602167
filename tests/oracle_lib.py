"""ctypes binding of the CPU oracle (oracle/libgahhg_oracle.so).

Test infrastructure: imported only by tests, the smoke check and the CPU
baseline leg of the benchmark, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "build", "libgahhg_oracle.so")

_lib = None


class OgLevel(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("vstart", "vnum", "bstart", "bnum", "cstart", "cnum", "rnum")]


class OgDiag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "walk_steps_mark", "walk_steps_dispatch", "walk_partner_missing",
        "walk_step_cap_hits", "walk_unlocated", "overlap_bad_index",
        "zero_ref_edges", "undefined_leaves_snapshot", "undefined_leaves_final",
        "partner_conflicts")] + [("mark_passes", ctypes.c_int32), ("pad", ctypes.c_int32)]


class OgQCount(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "plane_tests", "boundary_hits", "ref_tests", "bnum_sum", "unlocated", "inside")]


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build_oracle()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.og_new.restype = P
        L.og_free.argtypes = [P]
        L.og_last_error.restype = ctypes.c_char_p
        L.og_last_error.argtypes = [P]
        L.og_load_polygon_txt.argtypes = [P, ctypes.c_char_p]
        L.og_set_polygon_lonlat.argtypes = [P, dp, dp, lp, ctypes.c_int32]
        L.og_num_edges.argtypes = [P]
        L.og_num_edges.restype = ctypes.c_int32
        L.og_orientation_flips.argtypes = [P]
        L.og_orientation_flips.restype = ctypes.c_int32
        L.og_get_polygon.argtypes = [P, dp, ip]
        L.og_ll2xyz.argtypes = [dp, dp, ctypes.c_int64, dp]
        L.og_query_lonlat_to_xyz.argtypes = [dp, dp, ctypes.c_int64, dp]
        L.og_sincos_cr.argtypes = [dp, ctypes.c_int64, dp, dp]
        L.og_load_query_csv.argtypes = [P, ctypes.c_char_p, P, P, ctypes.c_int64]
        L.og_load_query_csv.restype = ctypes.c_int64
        L.og_grid0.argtypes = [dp, ip, dp, ip]
        L.og_build.argtypes = [P, ctypes.c_int32, ctypes.c_int32]
        L.og_build_create_only.argtypes = [P, ctypes.c_int32, ctypes.c_int32]
        L.og_classify.argtypes = [P]
        L.og_counts.argtypes = [P, lp]
        L.og_gamma.argtypes = [P]
        L.og_gamma.restype = ctypes.c_float
        L.og_levels.argtypes = [P, ctypes.POINTER(OgLevel)]
        L.og_diagnostics.argtypes = [P, ctypes.POINTER(OgDiag)]
        L.og_export.argtypes = [P, dp, ip, ip, dp, ip]
        L.og_query_xyz.argtypes = [P, dp, ctypes.c_int64, ip, ctypes.c_int32,
                                   ctypes.POINTER(OgQCount)]
        L.og_query_leaf.argtypes = [P, dp, ctypes.c_int64, ip]
        L.og_query_leaf.restype = None
        L.og_provenance.argtypes = [P, ip, ip, ip]
        L.og_provenance.restype = None
        L.og_debug_set_type.argtypes = [P, ctypes.c_int32, ctypes.c_int32]
        L.og_debug_set_type.restype = None
        L.og_kat_normalize.argtypes = [dp, dp]
        L.og_kat_cross.argtypes = [dp, dp, dp]
        L.og_kat_is_ee_inter.argtypes = [dp, dp, dp, dp]
        L.og_kat_is_ee_inter.restype = ctypes.c_int32
        L.og_kat_edge_border_inter.argtypes = [dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int32)]
        L.og_kat_edge_border_inter.restype = ctypes.c_int32
        L.og_kat_cp_prop.argtypes = [dp, ip, ip, ctypes.c_int32, dp, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]
        L.og_kat_cp_prop.restype = ctypes.c_int32
        _lib = L
    return _lib


def cp_prop(pv, pe, refs, center):
    """getCPProp over refs at center: (type, slot of the chosen edge in
    refs or -1, its nearest vertex or -1, squared distance)."""
    L = lib()
    sl, nv, md = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
    t = L.og_kat_cp_prop(np.ascontiguousarray(pv, dtype=np.float64).reshape(-1),
                         np.ascontiguousarray(pe, dtype=np.int32).reshape(-1),
                         np.ascontiguousarray(refs, dtype=np.int32), len(refs),
                         np.ascontiguousarray(center, dtype=np.float64), ctypes.byref(sl), ctypes.byref(nv),
                         ctypes.byref(md))
    return int(t), sl.value, nv.value, md.value


def sincos_cr(x):
    """Correctly rounded sin and cos of every element of x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().og_sincos_cr(x, x.size, s, c)
    return s, c


def grid0():
    L = lib()
    v = np.zeros(80 * 3)
    bi = np.zeros(240 * 6, dtype=np.int32)
    bn = np.zeros(240 * 3)
    ci = np.zeros(42 * 7, dtype=np.int32)
    L.og_grid0(v, bi, bn, ci)
    return v.reshape(80, 3), bi.reshape(240, 6), bn.reshape(240, 3), ci.reshape(42, 7)


def ll2xyz(lon, lat):
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    out = np.empty(lon.size * 3)
    lib().og_ll2xyz(lon, lat, lon.size, out)
    return out.reshape(-1, 3)


def query_lonlat_to_xyz(lon, lat):
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    out = np.empty(lon.size * 3)
    lib().og_query_lonlat_to_xyz(lon, lat, lon.size, out)
    return out.reshape(-1, 3)


class Oracle:
    """One polygon, its hierarchy and queries."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.og_new()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.og_free(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.og_last_error(self.h).decode())

    def error(self):
        return self.L.og_last_error(self.h).decode()

    def load_txt(self, path):
        self._check(self.L.og_load_polygon_txt(self.h, str(path).encode()))
        return self

    def set_lonlat(self, lon, lat, ring_off):
        self._check(self.L.og_set_polygon_lonlat(
            self.h, np.ascontiguousarray(lon, dtype=np.float64),
            np.ascontiguousarray(lat, dtype=np.float64),
            np.ascontiguousarray(ring_off, dtype=np.int64), len(ring_off) - 1))
        return self

    @property
    def n_edges(self):
        return self.L.og_num_edges(self.h)

    @property
    def flips(self):
        return self.L.og_orientation_flips(self.h)

    def polygon(self):
        n = self.n_edges
        pv = np.zeros(n * 3)
        pe = np.zeros(n * 2, dtype=np.int32)
        self.L.og_get_polygon(self.h, pv, pe)
        return pv.reshape(n, 3), pe.reshape(n, 2)

    def build(self, max_depth=15, auto_depth=True):
        self._check(self.L.og_build(self.h, max_depth, 1 if auto_depth else 0))
        return self

    def build_create_only(self, max_depth=15, auto_depth=True):
        self._check(self.L.og_build_create_only(self.h, max_depth, 1 if auto_depth else 0))
        return self

    def classify(self):
        self._check(self.L.og_classify(self.h))
        return self

    def counts(self):
        out = np.zeros(5, dtype=np.int64)
        self.L.og_counts(self.h, out)
        return dict(depth=int(out[0]), V=int(out[1]), B=int(out[2]), C=int(out[3]), R=int(out[4]))

    def gamma(self):
        return self.L.og_gamma(self.h)

    def levels(self):
        d = self.counts()["depth"]
        arr = (OgLevel * max(d, 1))()
        self.L.og_levels(self.h, arr)
        return [dict((f, getattr(arr[i], f)) for f, _ in OgLevel._fields_) for i in range(d)]

    def diag(self):
        d = OgDiag()
        self.L.og_diagnostics(self.h, ctypes.byref(d))
        return dict((f, getattr(d, f)) for f, _ in OgDiag._fields_ if f != "pad")

    def grid(self):
        c = self.counts()
        v = np.zeros(c["V"] * 3)
        cells = np.zeros(c["C"] * 7, dtype=np.int32)
        bi = np.zeros(c["B"] * 6, dtype=np.int32)
        bn = np.zeros(c["B"] * 3)
        refs = np.zeros(max(c["R"], 1), dtype=np.int32)
        self.L.og_export(self.h, v, cells, bi, bn, refs)
        return dict(v=v.reshape(-1, 3), cells=cells.reshape(-1, 7), borders=bi.reshape(-1, 6),
                    normals=bn.reshape(-1, 3), refs=refs[:c["R"]])

    def space_bytes(self):
        c = self.counts()
        return 24 * c["V"] + 48 * c["B"] + 28 * c["C"] + 4 * c["R"]

    def query_xyz(self, xyz, nthreads=0, with_counts=False):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        n = xyz.size // 3
        labels = np.zeros(n, dtype=np.int32)
        q = OgQCount()
        self.L.og_query_xyz(self.h, xyz, n, labels, nthreads, ctypes.byref(q))
        if with_counts:
            return labels, dict((f, getattr(q, f)) for f, _ in OgQCount._fields_)
        return labels

    def provenance(self):
        """(rule, source cell, border) per cell: which classification rule
        typed it (0 none, 1 setBottomNECell, 2 setBottomDECell, 3 setIECell,
        4 setDECell copy, 5 setDECell parity, 6 setTopECell)."""
        C = self.counts()["C"]
        r, s, b = (np.zeros(C, dtype=np.int32) for _ in range(3))
        self.L.og_provenance(self.h, r, s, b)
        return r, s, b

    def debug_set_type(self, cell, t):
        self.L.og_debug_set_type(self.h, int(cell), int(t))

    def query_leaf(self, xyz):
        """Leaf cell the first-match descent reaches for each point (-1 = none)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        n = xyz.size // 3
        leaf = np.zeros(n, dtype=np.int32)
        self.L.og_query_leaf(self.h, xyz, n, leaf)
        return leaf


def load_query_csv(path):
    L = lib()
    h = L.og_new()
    try:
        n = L.og_load_query_csv(h, str(path).encode(), None, None, 0)
        if n < 0:
            raise RuntimeError(L.og_last_error(h).decode())
        lon = np.zeros(n)
        lat = np.zeros(n)
        n2 = L.og_load_query_csv(h, str(path).encode(), lon.ctypes.data, lat.ctypes.data, n)
        assert n2 == n
        return lon, lat
    finally:
        L.og_free(h)
