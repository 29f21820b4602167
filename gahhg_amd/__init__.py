"""gahhg_amd -- MI355X (gfx950) GAHHG point-in-spherical-polygon library.

The product is the native library lib/libgahhg.so (C ABI in include/gahhg.h,
host C++ + hand-written HIP kernels).  This module is a thin ctypes binding
whose class mirrors the reference's driver sequence (source/GPISP.cu:47-67):

    GPISP.load_polygon_txt   <- GPISPPolygon::loadPolyFromTXT
    GPISP.build              <- GPISP::createAHHGInit + createAHHG + classifyAHHG
    GPISP.query_lonlat/xyz   <- GPISPTP::loadTPFromCSV + GPISP::test
    GPISP.stats              <- GPISP::getStat
    GPISP.write_report       <- GPISP::exportStat2TXT
    write_result_csv         <- GPISPTP::exportResult2CSV

There is no CPU fallback: if the library is missing or no GPU is visible the
calls raise.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libgahhg.so")
CSRC = os.path.join(PKG_DIR, "csrc")

LABEL_UNDEFINED, LABEL_OUT, LABEL_IN = 0, 1, 2

ERRORS = {-1: "bad argument", -2: "cannot open file", -3: "malformed file",
          -4: "call out of order", -5: "HIP runtime error", -6: "out of memory",
          -7: "size beyond int32 indexing"}


class GahhgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gahhg error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "max_depth", "auto_depth", "block_create", "block_classify", "block_query",
        "device", "verbose", "query_layout", "query_bins", "query_group", "query_sub",
    "query_sub_budget", "query_sort", "query_sort_bits", "query_index", "query_xcd",
    "query_table_budget", "query_graph", "build_device", "build_graph", "query_persistent")]


_STAT_FIELDS = [
    ("pre_ms", ctypes.c_double), ("create_ms", ctypes.c_double), ("classify_ms", ctypes.c_double),
    ("query_ms", ctypes.c_double), ("query_unit_s", ctypes.c_double), ("depth", ctypes.c_int32),
    ("gamma", ctypes.c_float), ("n_edges", ctypes.c_int64), ("n_queries", ctypes.c_int64),
    ("vnum", ctypes.c_int64), ("bnum", ctypes.c_int64), ("cnum", ctypes.c_int64), ("rnum", ctypes.c_int64),
    ("vspace", ctypes.c_int64), ("bspace", ctypes.c_int64), ("cspace", ctypes.c_int64),
    ("rspace", ctypes.c_int64), ("gpu_space_bytes", ctypes.c_int64), ("cpu_space_bytes", ctypes.c_int64),
    ("device_bytes", ctypes.c_int64), ("inside", ctypes.c_int64), ("outside", ctypes.c_int64),
    ("unlocated", ctypes.c_int64), ("orientation_flipped", ctypes.c_int32), ("mark_passes", ctypes.c_int32),
    ("walk_steps_mark", ctypes.c_int64), ("walk_steps_dispatch", ctypes.c_int64),
    ("walk_partner_missing", ctypes.c_int64), ("walk_step_cap_hits", ctypes.c_int64),
    ("walk_unlocated", ctypes.c_int64), ("zero_ref_edges", ctypes.c_int64),
    ("undefined_leaves_snapshot", ctypes.c_int64), ("undefined_leaves_final", ctypes.c_int64),
    ("query_kernel_ms", ctypes.c_double), ("query_sort_ms", ctypes.c_double),
    ("layout_ms", ctypes.c_double), ("build_replayed", ctypes.c_int32),
    ("build_attempts", ctypes.c_int32),
]


class Stats(ctypes.Structure):
    _fields_ = _STAT_FIELDS


# Every symbol include/gahhg.h declares (checked by the CPU test suite).
EXPORTED = [
    "gahhg_params_default", "gahhg_create", "gahhg_destroy", "gahhg_last_error",
    "gahhg_load_polygon_txt", "gahhg_set_polygon_lonlat", "gahhg_build", "gahhg_query_lonlat",
    "gahhg_query_xyz", "gahhg_query_xyz_device", "gahhg_read_query_csv", "gahhg_query_lonlat_to_xyz",
    "gahhg_get_stats", "gahhg_write_report", "gahhg_write_result_csv", "gahhg_device_alloc",
    "gahhg_device_free", "gahhg_memcpy_h2d", "gahhg_memcpy_d2h", "gahhg_grid_counts",
    "gahhg_export_grid", "gahhg_kat_device", "gahhg_set_polygons_lonlat", "gahhg_num_polygons",
    "gahhg_query_batch_xyz", "gahhg_query_batch_xyz_device", "gahhg_get_polygon_stats",
    "gahhg_polygon_grid_counts", "gahhg_export_polygon_grid", "gahhg_set_polygon_depths",
    "gahhg_trace_query_device", "gahhg_build_trace", "gahhg_mlp_forward", "gahhg_mlp_backward",
    "gahhg_mlp_adam", "gahhg_set_depth_model", "gahhg_set_query_hint", "gahhg_depth_features",
    "gahhg_autotune_query", "gahhg_save_grid", "gahhg_load_grid", "gahhg_write_polygon_report",
    "gahhg_load_polygons_txt", "gahhg_load_query_csv_device", "gahhg_parse_query_csv_device",
    "gahhg_load_polygon_txt_device", "gahhg_sort_pairs", "gahhg_query_pairs_xyz", "gahhg_autotune_layout",
    "gahhg_query_batch_xyz_device_noalloc", "gahhg_set_table_model", "gahhg_set_table_choice",
    "gahhg_get_table_layout", "gahhg_table_features", "gahhg_read_lune_points", "gahhg_read_lune_polygon",
    "gahhg_load_polygon_lune", "gahhg_kat_cp_prop",
]

# Start-table candidates of the learned table depth (bins per face side,
# sub-table side); a batch considers the first TABLE_CAND_BATCH.
TABLE_BINS = (0, 16, 32, 64, 128, 256, 256, 256)
TABLE_SUB = (1, 1, 1, 1, 1, 1, 8, 16)
TABLE_CAND_BATCH = 5

# Points per call of the allocation-free query (GAHHG_NOALLOC_MAX_POINTS).
NOALLOC_MAX_POINTS = 4194304

# Query trace record fields (6 int32 per point) and build trace layout.
TRACE_FIELDS = ("leaf", "root", "level", "start_level", "plane_tests", "records")
TRACE_LEVELS = 32
TOP_CELLS = 42
FIXED_POINT = 2.0 ** 40

_lib = None


def build_library(jobs=8):
    import subprocess
    subprocess.check_call(["make", "-s", f"-j{jobs}", "-C", CSRC])


def lib():
    """Loads lib/libgahhg.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GahhgError(-4, f"{LIB_PATH} not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    sig = {
        "gahhg_params_default": ([ctypes.POINTER(Params)], ctypes.c_int),
        "gahhg_create": ([ctypes.POINTER(P), ctypes.POINTER(Params)], ctypes.c_int),
        "gahhg_destroy": ([P], None),
        "gahhg_last_error": ([P], ctypes.c_char_p),
        "gahhg_load_polygon_txt": ([P, ctypes.c_char_p], ctypes.c_int),
        "gahhg_set_polygon_lonlat": ([P, dp, dp, lp, i32], ctypes.c_int),
        "gahhg_build": ([P], ctypes.c_int),
        "gahhg_query_lonlat": ([P, dp, dp, i64, ip], ctypes.c_int),
        "gahhg_query_xyz": ([P, dp, i64, ip], ctypes.c_int),
        "gahhg_query_xyz_device": ([P, P, i64, P], ctypes.c_int),
        "gahhg_read_query_csv": ([ctypes.c_char_p, P, P, i64, ctypes.POINTER(i64)], ctypes.c_int),
        "gahhg_query_lonlat_to_xyz": ([dp, dp, i64, dp], ctypes.c_int),
        "gahhg_get_stats": ([P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gahhg_write_report": ([P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p], ctypes.c_int),
        "gahhg_write_result_csv": ([ctypes.c_char_p, ip, i64], ctypes.c_int),
        "gahhg_device_alloc": ([P, i64, ctypes.POINTER(P)], ctypes.c_int),
        "gahhg_device_free": ([P, P], ctypes.c_int),
        "gahhg_memcpy_h2d": ([P, P, P, i64], ctypes.c_int),
        "gahhg_memcpy_d2h": ([P, P, P, i64], ctypes.c_int),
        "gahhg_grid_counts": ([P, lp], ctypes.c_int),
        "gahhg_export_grid": ([P, P, P, P, P, P, P], ctypes.c_int),
        "gahhg_kat_device": ([P, i32, dp, i64, dp], ctypes.c_int),
        "gahhg_sort_pairs": ([P, P, P, i64, i64, i32], ctypes.c_int),
        "gahhg_set_polygons_lonlat": ([P, dp, dp, lp, ip, i32], ctypes.c_int),
        "gahhg_num_polygons": ([P, ctypes.POINTER(i32)], ctypes.c_int),
        "gahhg_query_batch_xyz": ([P, dp, ip, i64, ip], ctypes.c_int),
        "gahhg_query_batch_xyz_device": ([P, P, P, i64, P], ctypes.c_int),
        "gahhg_query_batch_xyz_device_noalloc": ([P, P, P, i64, P], ctypes.c_int),
        "gahhg_query_pairs_xyz": ([P, dp, i64, ip], ctypes.c_int),
        "gahhg_autotune_layout": ([P, P, P, i64, P, ctypes.c_double, P], ctypes.c_int),
        "gahhg_get_polygon_stats": ([P, i32, ctypes.POINTER(Stats)], ctypes.c_int),
        "gahhg_polygon_grid_counts": ([P, i32, lp], ctypes.c_int),
        "gahhg_export_polygon_grid": ([P, i32, P, P, P, P, P, P], ctypes.c_int),
        "gahhg_set_polygon_depths": ([P, ip, i32], ctypes.c_int),
        "gahhg_trace_query_device": ([P, P, P, i64, P, P], ctypes.c_int),
        "gahhg_build_trace": ([P, P, P], ctypes.c_int),
        "gahhg_mlp_forward": ([P, P, i64, P, P], ctypes.c_int),
        "gahhg_mlp_backward": ([P, P, i64, P, P, P], ctypes.c_int),
        "gahhg_mlp_adam": ([P, P, P, P, P, i32, ctypes.c_float, P], ctypes.c_int),
        "gahhg_set_depth_model": ([P, P, P], ctypes.c_int),
        "gahhg_set_table_model": ([P, P, P], ctypes.c_int),
        "gahhg_set_table_choice": ([P, P, i32], ctypes.c_int),
        "gahhg_get_table_layout": ([P, P], ctypes.c_int),
        "gahhg_table_features": ([P, P], ctypes.c_int),
        "gahhg_set_query_hint": ([P, P, P, i64, ctypes.c_double], ctypes.c_int),
        "gahhg_depth_features": ([P, i32, P], ctypes.c_int),
        "gahhg_autotune_query": ([P, P, P, i64, P, P], ctypes.c_int),
        "gahhg_save_grid": ([P, ctypes.c_char_p], ctypes.c_int),
        "gahhg_load_grid": ([P, ctypes.c_char_p], ctypes.c_int),
        "gahhg_read_lune_points": ([ctypes.c_char_p, P, P, i64, ctypes.POINTER(i64)], ctypes.c_int),
        "gahhg_read_lune_polygon": ([ctypes.c_char_p, P, P, i64, ctypes.POINTER(i64), dp, ip], ctypes.c_int),
        "gahhg_load_polygon_lune": ([P, ctypes.c_char_p, dp, ip], ctypes.c_int),
        "gahhg_kat_cp_prop": ([P, ip, i64, ip, dp, i64, i32, ip, dp], ctypes.c_int),
        "gahhg_load_polygons_txt": ([P, ctypes.POINTER(ctypes.c_char_p), i32], ctypes.c_int),
        "gahhg_load_query_csv_device": ([P, ctypes.c_char_p, P, P, P], ctypes.c_int),
        "gahhg_parse_query_csv_device": ([P, ctypes.c_char_p, i64, P, P, P], ctypes.c_int),
        "gahhg_load_polygon_txt_device": ([P, ctypes.c_char_p, P], ctypes.c_int),
        "gahhg_write_polygon_report": ([P, i32, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i64, i64,
                                        ctypes.c_double, ctypes.c_double], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def default_params(**kw):
    p = Params()
    lib().gahhg_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, int(v))
    return p


def query_lonlat_to_xyz(lon, lat):
    """Reference test-point conversion (FakeLLVertex round trip), host side."""
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    out = np.empty(lon.size * 3)
    rc = lib().gahhg_query_lonlat_to_xyz(lon, lat, lon.size, out)
    if rc:
        raise GahhgError(rc, "lonlat_to_xyz")
    return out.reshape(-1, 3)


def read_query_csv(path):
    L = lib()
    n = ctypes.c_int64(0)
    rc = L.gahhg_read_query_csv(str(path).encode(), None, None, 0, ctypes.byref(n))
    if rc:
        raise GahhgError(rc, f"read {path}")
    lon = np.zeros(n.value)
    lat = np.zeros(n.value)
    rc = L.gahhg_read_query_csv(str(path).encode(), lon.ctypes.data, lat.ctypes.data, n.value,
                                ctypes.byref(n))
    if rc:
        raise GahhgError(rc, f"read {path}")
    return lon, lat


def read_lune_points(path):
    """Lune/Ketzner-format points (test/data_for_Lune) -> lon, lat degrees
    in the reference's convention, read by the library."""
    L = lib()
    n = ctypes.c_int64(0)
    rc = L.gahhg_read_lune_points(str(path).encode(), None, None, 0, ctypes.byref(n))
    if rc:
        raise GahhgError(rc, f"read {path}")
    lon = np.zeros(n.value)
    lat = np.zeros(n.value)
    rc = L.gahhg_read_lune_points(str(path).encode(), lon.ctypes.data, lat.ctypes.data, n.value, ctypes.byref(n))
    if rc:
        raise GahhgError(rc, f"read {path}")
    return lon, lat


def read_lune_polygon(path):
    """A Lune/Ketzner-format polygon -> (lon, lat, ring_off, ref) read by
    the library: one ring with its closing vertex, and the point the file
    gives as inside (lon, lat), or None."""
    L = lib()
    n = ctypes.c_int64(0)
    ref = np.zeros(2)
    has = np.zeros(1, dtype=np.int32)
    rc = L.gahhg_read_lune_polygon(str(path).encode(), None, None, 0, ctypes.byref(n), ref, has)
    if rc:
        raise GahhgError(rc, f"read {path}")
    lon = np.zeros(n.value)
    lat = np.zeros(n.value)
    rc = L.gahhg_read_lune_polygon(str(path).encode(), lon.ctypes.data, lat.ctypes.data, n.value,
                                   ctypes.byref(n), ref, has)
    if rc:
        raise GahhgError(rc, f"read {path}")
    return lon, lat, np.array([0, n.value], dtype=np.int64), ((float(ref[0]), float(ref[1])) if has[0] else None)


def write_result_csv(path, labels):
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    rc = lib().gahhg_write_result_csv(str(path).encode(), labels, labels.size)
    if rc:
        raise GahhgError(rc, f"write {path}")


class LoadInfo(ctypes.Structure):
    _fields_ = [("lines", ctypes.c_int64), ("bytes", ctypes.c_int64), ("host_lines", ctypes.c_int64),
                ("read_ms", ctypes.c_double), ("parse_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]

    def as_dict(self):
        return dict((f, getattr(self, f)) for f, _ in self._fields_)


class DeviceView:
    """Device memory owned by the context (not freed by the caller), e.g. the
    points of a device-side load; valid until the context's next load."""

    def __init__(self, ctx, ptr, nbytes):
        self.ctx = ctx
        self.ptr = ctypes.c_void_p(ptr)
        self.nbytes = int(nbytes)

    def download(self, arr):
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        self.ctx._check(self.ctx.L.gahhg_memcpy_d2h(self.ctx.h, arr.ctypes.data, self.ptr, arr.nbytes))
        return arr


class DeviceBuffer:
    """Device allocation owned by a context (for HBM-resident inputs)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx._check(ctx.L.gahhg_device_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx._check(self.ctx.L.gahhg_memcpy_h2d(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, arr):
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        self.ctx._check(self.ctx.L.gahhg_memcpy_d2h(self.ctx.h, arr.ctypes.data, self.ptr, arr.nbytes))
        return arr

    def free(self):
        if self.ptr is not None and self.ctx.h:
            self.ctx._check(self.ctx.L.gahhg_device_free(self.ctx.h, self.ptr))
        self.ptr = None


class GPISP:
    """One polygon's hierarchy on one GPU (the reference's GPISP object)."""

    def __init__(self, max_depth=15, auto_depth=True, device=0, verbose=False, block_query=256,
                 query_layout=0, query_bins=256, query_group=0, query_sub=8, query_sub_budget=4,
                 query_sort=1, query_sort_bits=14, query_index=1, query_xcd=32, query_table_budget=8,
                 query_graph=0, build_device=1, build_graph=1, query_persistent=0):
        self.L = lib()
        prm = default_params(max_depth=max_depth, auto_depth=1 if auto_depth else 0, device=device,
                             verbose=1 if verbose else 0, block_query=block_query,
                             query_layout=query_layout, query_bins=query_bins, query_group=query_group,
                             query_sub=query_sub, query_sub_budget=query_sub_budget, query_sort=query_sort,
                             query_sort_bits=query_sort_bits, query_index=query_index, query_xcd=query_xcd,
                             query_table_budget=query_table_budget, query_graph=query_graph,
                             build_device=build_device, build_graph=build_graph,
                             query_persistent=query_persistent)
        h = ctypes.c_void_p()
        rc = self.L.gahhg_create(ctypes.byref(h), ctypes.byref(prm))
        if rc:
            raise GahhgError(rc, "gahhg_create (is a GPU visible?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.gahhg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            raise GahhgError(rc, self.L.gahhg_last_error(self.h).decode())

    # -- polygon -----------------------------------------------------------
    def load_polygon_txt(self, path):
        self._check(self.L.gahhg_load_polygon_txt(self.h, str(path).encode()))
        return self

    loadPolyFromTXT = load_polygon_txt

    def load_polygon_lune(self, path):
        """A Lune/Ketzner-format polygon file; returns the point the file
        gives as inside (lon, lat degrees) or None."""
        ref = np.zeros(2)
        has = np.zeros(1, dtype=np.int32)
        self._check(self.L.gahhg_load_polygon_lune(self.h, str(path).encode(), ref, has))
        return (float(ref[0]), float(ref[1])) if has[0] else None

    def load_polygon_txt_device(self, path):
        """load_polygon_txt with the vertex lines parsed on the GPU; the load
        figures are kept in self.load_info."""
        info = LoadInfo()
        self._check(self.L.gahhg_load_polygon_txt_device(self.h, str(path).encode(), ctypes.byref(info)))
        self.load_info = info.as_dict()
        return self

    # -- device-side query loading -------------------------------------------
    def load_query_csv_device(self, path):
        """Query CSV parsed and converted on the GPU.  Returns (DeviceView of
        n x 3 doubles, n); the load figures are kept in self.load_info."""
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        info = LoadInfo()
        self._check(self.L.gahhg_load_query_csv_device(self.h, str(path).encode(), ctypes.byref(p),
                                                       ctypes.byref(n), ctypes.byref(info)))
        self.load_info = info.as_dict()
        return DeviceView(self, p.value, 24 * n.value), n.value

    def parse_query_csv_device(self, text):
        """As load_query_csv_device, from the file's bytes in host memory."""
        text = bytes(text)
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        info = LoadInfo()
        self._check(self.L.gahhg_parse_query_csv_device(self.h, text, len(text), ctypes.byref(p),
                                                        ctypes.byref(n), ctypes.byref(info)))
        self.load_info = info.as_dict()
        return DeviceView(self, p.value, 24 * n.value), n.value

    def set_polygon_lonlat(self, lon, lat, ring_off):
        self._check(self.L.gahhg_set_polygon_lonlat(
            self.h, np.ascontiguousarray(lon, dtype=np.float64), np.ascontiguousarray(lat, dtype=np.float64),
            np.ascontiguousarray(ring_off, dtype=np.int64), len(ring_off) - 1))
        return self

    # -- build -------------------------------------------------------------
    def build(self):
        self._check(self.L.gahhg_build(self.h))
        return self

    # -- query -------------------------------------------------------------
    def query_xyz(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        n = xyz.size // 3
        labels = np.zeros(n, dtype=np.int32)
        self._check(self.L.gahhg_query_xyz(self.h, xyz, n, labels))
        return labels

    def query_lonlat(self, lon, lat):
        lon = np.ascontiguousarray(lon, dtype=np.float64)
        lat = np.ascontiguousarray(lat, dtype=np.float64)
        labels = np.zeros(lon.size, dtype=np.int32)
        self._check(self.L.gahhg_query_lonlat(self.h, lon, lat, lon.size, labels))
        return labels

    def query_device(self, d_xyz, n, d_labels):
        self._check(self.L.gahhg_query_xyz_device(self.h, d_xyz.ptr, int(n), d_labels.ptr))

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    # -- stats / export ----------------------------------------------------
    def stats(self):
        s = Stats()
        self._check(self.L.gahhg_get_stats(self.h, ctypes.byref(s)))
        return dict((f, getattr(s, f)) for f, _ in _STAT_FIELDS)

    def write_report(self, path, poly_name="", query_name=""):
        self._check(self.L.gahhg_write_report(self.h, str(path).encode(), poly_name.encode(),
                                              query_name.encode()))

    # -- batches -----------------------------------------------------------
    def set_polygons_lonlat(self, polys):
        """polys: list of (lon, lat, ring_off) as for set_polygon_lonlat."""
        lon, lat, roff, proff = [], [], [0], [0]
        base = 0
        for plon, plat, pro in polys:
            pro = np.asarray(pro, dtype=np.int64)
            lon.append(np.asarray(plon, dtype=np.float64))
            lat.append(np.asarray(plat, dtype=np.float64))
            roff.extend((pro[1:] + base).tolist())
            base += int(pro[-1])
            proff.append(len(roff) - 1)
        self._check(self.L.gahhg_set_polygons_lonlat(
            self.h, np.ascontiguousarray(np.concatenate(lon)), np.ascontiguousarray(np.concatenate(lat)),
            np.asarray(roff, dtype=np.int64), np.asarray(proff, dtype=np.int32), len(polys)))
        return self

    def load_polygons_txt(self, paths):
        """A batch from polygon text files (polygon k = paths[k])."""
        paths = [str(x).encode() for x in paths]
        arr = (ctypes.c_char_p * len(paths))(*paths)
        self._check(self.L.gahhg_load_polygons_txt(self.h, arr, len(paths)))
        return self

    @property
    def n_polygons(self):
        n = ctypes.c_int32()
        self._check(self.L.gahhg_num_polygons(self.h, ctypes.byref(n)))
        return n.value

    def query_batch_xyz(self, xyz, poly):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        poly = np.ascontiguousarray(poly, dtype=np.int32)
        labels = np.zeros(poly.size, dtype=np.int32)
        self._check(self.L.gahhg_query_batch_xyz(self.h, xyz, poly, poly.size, labels))
        return labels

    def query_pairs_xyz(self, xyz):
        """Every polygon of the batch against the same points: labels of
        shape (polygons, points)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        labels = np.zeros((self.n_polygons, xyz.shape[0]), dtype=np.int32)
        self._check(self.L.gahhg_query_pairs_xyz(self.h, xyz.reshape(-1), xyz.shape[0], labels.reshape(-1)))
        return labels

    def query_batch_device(self, d_xyz, d_poly, n, d_labels):
        self._check(self.L.gahhg_query_batch_xyz_device(self.h, d_xyz.ptr, d_poly.ptr, int(n), d_labels.ptr))

    def query_device_noalloc(self, d_xyz, d_poly, n, d_labels):
        """gahhg_query_batch_xyz_device_noalloc: the same pass with its
        scratch in the library's fixed device array (d_poly None for a
        single polygon; at most NOALLOC_MAX_POINTS points)."""
        self._check(self.L.gahhg_query_batch_xyz_device_noalloc(
            self.h, d_xyz.ptr, None if d_poly is None else d_poly.ptr, int(n), d_labels.ptr))

    def set_polygon_depths(self, depths):
        """Fixed depth per polygon (0 = the context's rule); call after
        setting the polygons."""
        d = np.ascontiguousarray(depths, dtype=np.int32)
        self._check(self.L.gahhg_set_polygon_depths(self.h, d, d.size))
        return self

    def trace_device(self, d_xyz, d_poly, n, d_labels, d_trace):
        """Traced query: labels as query_batch_device plus 6 int32 per point
        (TRACE_FIELDS) in d_trace."""
        self._check(self.L.gahhg_trace_query_device(self.h, d_xyz.ptr, d_poly.ptr if d_poly is not None else None,
                                                    int(n), d_labels.ptr, d_trace.ptr))

    def trace_xyz(self, xyz, poly=None):
        """Host wrapper of trace_device: returns (labels, trace[n, 6])."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        dx, dl, dt = self.alloc(max(xyz.nbytes, 8)), self.alloc(max(4 * n, 4)), self.alloc(max(24 * n, 4))
        dp = None
        try:
            dx.upload(xyz)
            if poly is not None:
                dp = self.alloc(max(4 * n, 4))
                dp.upload(np.ascontiguousarray(poly, dtype=np.int32))
            self.trace_device(dx, dp, n, dl, dt)
            lab = dl.download(np.empty(n, dtype=np.int32))
            tr = dt.download(np.empty((n, 6), dtype=np.int32))
        finally:
            for b in (dx, dl, dt, dp):
                if b is not None:
                    b.free()
        return lab, tr

    def build_trace(self):
        """Build trace of the last build: cells[p, level, top cell, 4] =
        (cells, refined cells, bottom cells with refs, refs) and
        edges[p, top cell, 3] = (edges, sum arc length rad, sum turning rad)."""
        np_ = self.n_polygons
        cells = np.zeros((np_, TRACE_LEVELS, TOP_CELLS, 4), dtype=np.int32)
        edges = np.zeros((np_, TOP_CELLS, 3), dtype=np.uint64)
        self._check(self.L.gahhg_build_trace(self.h, cells.ctypes.data, edges.ctypes.data))
        e = edges.astype(np.float64)
        e[..., 1:] /= FIXED_POINT
        return cells, e

    def write_polygon_report(self, p, path, poly_name, query_name, n_queries, inside, query_ms, pre_ms):
        self._check(self.L.gahhg_write_polygon_report(self.h, int(p), str(path).encode(), str(poly_name).encode(),
                                                      str(query_name).encode(), int(n_queries), int(inside),
                                                      float(query_ms), float(pre_ms)))

    def save_grid(self, path):
        self._check(self.L.gahhg_save_grid(self.h, str(path).encode()))
        return self

    def load_grid(self, path):
        """Loads hierarchies saved by save_grid; the context is then built."""
        self._check(self.L.gahhg_load_grid(self.h, str(path).encode()))
        return self

    def autotune_device(self, d_xyz, d_poly, n, d_labels):
        """Times the query's lane-group width and block size on these points
        and keeps the fastest; returns (group, block, microseconds)."""
        out = np.zeros(3, dtype=np.int32)
        self._check(self.L.gahhg_autotune_query(self.h, d_xyz.ptr, d_poly.ptr if d_poly is not None else None,
                                                int(n), d_labels.ptr, out.ctypes.data))
        return int(out[0]), int(out[1]), int(out[2])

    def autotune_layout_device(self, d_xyz, d_poly, n, d_labels, passes=1.0):
        """Re-packs the query layout for each candidate start-table / sub-
        table setting, keeps the one with the least layout ms / passes +
        query ms; returns a dict of the chosen settings and their times."""
        out = np.zeros(6, dtype=np.float64)
        self._check(self.L.gahhg_autotune_layout(self.h, d_xyz.ptr, d_poly.ptr if d_poly is not None else None,
                                                 int(n), d_labels.ptr, float(passes), out.ctypes.data))
        keys = ("query_bins", "query_sub", "query_sub_budget", "query_table_budget")
        res = {k: int(v) for k, v in zip(keys, out[:4])}
        res.update(layout_ms=float(out[4]), query_ms=float(out[5]))
        return res

    def set_depth_model(self, params, cost):
        """Learned refinement depth (params: the model's 5313 values; cost:
        expected points, query-time scale, ns per edge, ns per next-level
        cell, ns per level).  params None switches it off."""
        if params is None:
            self._check(self.L.gahhg_set_depth_model(self.h, None, None))
            return self
        pa = np.ascontiguousarray(params, dtype=np.float32)
        co = np.ascontiguousarray(cost, dtype=np.float64)
        assert pa.size == 5313 and co.size == 5
        self._check(self.L.gahhg_set_depth_model(self.h, pa.ctypes.data, co.ctypes.data))
        return self

    def set_table_model(self, params, cost):
        """Learned start-table depth (label-neutral): params = the model's
        5313 values, cost = 19 values (expected points without a hint,
        query scale, per candidate the table build ns a_c, b_c per edge of
        one polygon, ns per table bin in a batch).  params None switches it
        off."""
        if params is None:
            self._check(self.L.gahhg_set_table_model(self.h, None, None))
            return self
        pa = np.ascontiguousarray(params, dtype=np.float32)
        co = np.ascontiguousarray(cost, dtype=np.float64)
        assert pa.size == 5313 and co.size == 3 + 2 * len(TABLE_BINS)
        self._check(self.L.gahhg_set_table_model(self.h, pa.ctypes.data, co.ctypes.data))
        return self

    def set_table_choice(self, cand):
        """Fixed start-table candidate per polygon (-1: chosen); None clears."""
        if cand is None:
            self._check(self.L.gahhg_set_table_choice(self.h, None, 0))
            return self
        a = np.ascontiguousarray(cand, dtype=np.int32)
        self._check(self.L.gahhg_set_table_choice(self.h, a.ctypes.data, a.size))
        return self

    def table_layout(self):
        """(bins per face side, sub-table side) per polygon of the last build."""
        out = np.zeros((self.n_polygons, 2), dtype=np.int32)
        self._check(self.L.gahhg_get_table_layout(self.h, out.ctypes.data))
        return out

    def table_features(self):
        x = np.zeros((self.n_polygons, TOP_CELLS, len(TABLE_BINS), 12), dtype=np.float32)
        self._check(self.L.gahhg_table_features(self.h, x.ctypes.data))
        return x

    def set_query_hint(self, xyz, poly=None, weight=1.0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        pp = None if poly is None else np.ascontiguousarray(poly, dtype=np.int32)
        self._check(self.L.gahhg_set_query_hint(self.h, xyz.ctypes.data if len(xyz) else None,
                                                pp.ctypes.data if pp is not None else None, len(xyz),
                                                float(weight)))
        self._hint_keep = (xyz, pp)
        return self

    def depth_features(self, level):
        x = np.zeros((self.n_polygons, TOP_CELLS, 2, 12), dtype=np.float32)
        self._check(self.L.gahhg_depth_features(self.h, int(level), x.ctypes.data))
        return x

    def polygon_stats(self, p):
        s = Stats()
        self._check(self.L.gahhg_get_polygon_stats(self.h, p, ctypes.byref(s)))
        return dict((f, getattr(s, f)) for f, _ in _STAT_FIELDS)

    def polygon_grid(self, p):
        cnt = np.zeros(5, dtype=np.int64)
        self._check(self.L.gahhg_polygon_grid_counts(self.h, p, cnt))
        d, V, B, C, R = (int(x) for x in cnt)
        v = np.zeros(V * 3)
        cells = np.zeros(C * 7, dtype=np.int32)
        borders = np.zeros(B * 6, dtype=np.int32)
        normals = np.zeros(B * 3)
        refs = np.zeros(max(R, 1), dtype=np.int32)
        levels = np.zeros(7 * d, dtype=np.int32)
        self._check(self.L.gahhg_export_polygon_grid(self.h, p, v.ctypes.data, cells.ctypes.data,
                                                     borders.ctypes.data, normals.ctypes.data, refs.ctypes.data,
                                                     levels.ctypes.data))
        return dict(depth=d, v=v.reshape(-1, 3), cells=cells.reshape(-1, 7), borders=borders.reshape(-1, 6),
                    normals=normals.reshape(-1, 3), refs=refs[:R], levels=levels.reshape(-1, 7))

    def grid(self):
        cnt = np.zeros(5, dtype=np.int64)
        self._check(self.L.gahhg_grid_counts(self.h, cnt))
        d, V, B, C, R = (int(x) for x in cnt)
        v = np.zeros(V * 3)
        cells = np.zeros(C * 7, dtype=np.int32)
        borders = np.zeros(B * 6, dtype=np.int32)
        normals = np.zeros(B * 3)
        refs = np.zeros(max(R, 1), dtype=np.int32)
        levels = np.zeros(7 * d, dtype=np.int32)
        self._check(self.L.gahhg_export_grid(self.h, v.ctypes.data, cells.ctypes.data, borders.ctypes.data,
                                             normals.ctypes.data, refs.ctypes.data, levels.ctypes.data))
        return dict(depth=d, v=v.reshape(-1, 3), cells=cells.reshape(-1, 7), borders=borders.reshape(-1, 6),
                    normals=normals.reshape(-1, 3), refs=refs[:R], levels=levels.reshape(-1, 7))

    def sort_pairs(self, keys, vals, end_bit=32, n_bound=None):
        """The grouping sort on the device: (keys, vals) sorted stably by
        key bits [0, end_bit), with launches sized for n_bound points."""
        k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
        v = np.ascontiguousarray(vals, dtype=np.uint32).copy()
        nb = k.size if n_bound is None else int(n_bound)
        self._check(self.L.gahhg_sort_pairs(self.h, k.ctypes.data, v.ctypes.data, k.size, nb, end_bit))
        return k, v

    def kat_cp_prop(self, refs, rs_rn, centers, group):
        """getCPProp's choice per case (see gahhg_kat_cp_prop): int32 (n, 3)
        type / slot / vertex and float64 (n,) squared distances."""
        refs = np.ascontiguousarray(refs, dtype=np.int32)
        rs_rn = np.ascontiguousarray(rs_rn, dtype=np.int32).reshape(-1)
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1)
        n = rs_rn.size // 2
        out = np.zeros(3 * n, dtype=np.int32)
        md = np.zeros(n)
        self._check(self.L.gahhg_kat_cp_prop(self.h, refs, refs.size, rs_rn, centers, n, int(group), out, md))
        return out.reshape(n, 3), md

    def kat(self, op, inputs):
        inputs = np.ascontiguousarray(inputs, dtype=np.float64).reshape(-1, 12)
        out = np.zeros((inputs.shape[0], 3))
        self._check(self.L.gahhg_kat_device(self.h, op, inputs.reshape(-1), inputs.shape[0], out.reshape(-1)))
        return out
