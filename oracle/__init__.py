"""ctypes wrapper of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: the oracle is a CPU restatement of the reference's
CPU path and is used as the checker by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (farm_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")

ORC_OK = 0
ORC_EINVAL = 1
ORC_EREF_ABORT = 2


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_create": (C.c_void_p, [C.c_int64, C.c_void_p, C.c_void_p]),
        "orc_destroy": (None, [C.c_void_p]),
        "orc_size": (C.c_int64, [C.c_void_p]),
        "orc_rasterize": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
        "orc_raster_info": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_raster_copy": (C.c_int, [C.c_void_p, C.c_int64] + [C.c_void_p] * 8),
        "orc_intersect": (
            C.c_int,
            [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p],
        ),
        "orc_candidates": (C.c_int64, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
        "orc_within": (
            C.c_int,
            [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
             C.c_int, C.c_void_p],
        ),
        "orc_candidates_within": (C.c_int64, [C.c_void_p, C.c_double, C.c_int64, C.c_void_p, C.c_void_p]),
        "orc_mbb": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
        "orc_failed": (C.c_int, [C.c_void_p, C.c_int64]),
        "orc_intersection": (
            C.c_int,
            [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int],
        ),
        "orc_layer": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_segment_intersect": (C.c_int, [C.c_void_p]),
        "orc_double_to_int": (C.c_int, [C.c_double]),
        "orc_retrieve": (C.c_int64, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_void_p, C.c_int64]),
        "orc_pixel_box": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p]),
        "orc_point_to_segment_distance": (C.c_double, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class OracleSet:
    """Polygons built from .dt-style rings, rasterized by the restatement."""

    def __init__(self, ring_off: np.ndarray, xy: np.ndarray):
        self.ring_off = np.ascontiguousarray(ring_off, dtype=np.uint64)
        self.xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        self._h = C.c_void_p(lib.orc_create(len(self.ring_off) - 1, _p(self.ring_off), _p(self.xy)))

    def __len__(self) -> int:
        return int(lib.orc_size(self._h))

    def rasterize(self, vpr: int = 10, hierarchy: bool = False, nlow: int = 1, nthreads: int = 8) -> int:
        """Returns -1 on success, else the index of the first polygon on which
        the reference would abort."""
        bad = C.c_int64(-1)
        rc = lib.orc_rasterize(self._h, vpr, int(hierarchy), nlow, nthreads, C.byref(bad))
        if rc == ORC_EINVAL:
            raise ValueError("bad rasterize arguments")
        return bad.value if rc == ORC_EREF_ABORT else -1

    def failed(self) -> np.ndarray:
        """Per polygon: True where the reference raster builder aborts or
        never terminates."""
        return np.array([lib.orc_failed(self._h, i) == 1 for i in range(len(self))], dtype=bool)

    def mbb(self, i: int) -> np.ndarray:
        out = np.zeros(4)
        lib.orc_mbb(self._h, i, _p(out))
        return out

    def raster(self, i: int) -> dict:
        mbr = np.zeros(4)
        dims = np.zeros(2, dtype=np.int32)
        steps = np.zeros(2)
        counts = np.zeros(5, dtype=np.int64)
        if lib.orc_raster_info(self._h, i, _p(mbr), _p(dims), _p(steps), _p(counts)):
            raise IndexError(i)
        npix = int(counts[0])
        r = {
            "mbr": mbr,
            "dimx": int(dims[0]),
            "dimy": int(dims[1]),
            "step": steps,
            "status": np.zeros(npix, np.uint8),
            "pixoff": np.zeros(npix + 1, np.uint32),
            "seqs": np.zeros((int(counts[1]), 2), np.uint32),
            "voff": np.zeros(int(dims[0]) + 2, np.uint32),
            "vnodes": np.zeros(int(counts[2])),
            "hoff": np.zeros(int(dims[1]) + 2, np.uint32),
            "hnodes": np.zeros(int(counts[3])),
            "areas": np.zeros(npix),
            "nverts": int(counts[4]),
        }
        lib.orc_raster_copy(self._h, i, _p(r["status"]), _p(r["pixoff"]), _p(r["seqs"]), _p(r["voff"]),
                            _p(r["vnodes"]), _p(r["hoff"]), _p(r["hnodes"]), _p(r["areas"]))
        return r

    def retrieve(self, i: int, box, unshrunk: bool = False) -> np.ndarray:
        """Pixel ids of polygon i's raster over box (low x, low y, high x,
        high y), in the reference's enumeration order."""
        b = np.ascontiguousarray(box, np.float64)
        n = int(lib.orc_retrieve(self._h, i, _p(b), int(unshrunk), None, 0))
        if n < 0:
            raise IndexError(i)
        ids = np.zeros(max(n, 1), np.int32)
        lib.orc_retrieve(self._h, i, _p(b), int(unshrunk), _p(ids), n)
        return ids[:n]

    def pixel_box(self, i: int, x: int, y: int) -> np.ndarray:
        out = np.zeros(4)
        if lib.orc_pixel_box(self._h, i, x, y, _p(out)):
            raise IndexError(i)
        return out

    def nlayers(self, i: int) -> int:
        low = np.zeros(2)
        step = np.zeros(1)
        dims = np.zeros(2, np.int32)
        return int(lib.orc_layer(self._h, i, -1, _p(low), _p(step), _p(dims), None))

    def layer(self, i: int, k: int) -> dict:
        """Layer k (0 = coarsest) of polygon i after rasterize(hierarchy=True)."""
        low = np.zeros(2)
        step = np.zeros(1)
        dims = np.zeros(2, np.int32)
        nl = lib.orc_layer(self._h, i, k, _p(low), _p(step), _p(dims), None)
        if nl < 0 or k >= nl:
            raise IndexError((i, k))
        st = np.zeros(int(dims[0]) * int(dims[1]), np.uint8)
        lib.orc_layer(self._h, i, k, _p(low), _p(step), _p(dims), _p(st))
        return {"low": low, "step": float(step[0]), "dimx": int(dims[0]), "dimy": int(dims[1]), "status": st}

    def __del__(self):
        if getattr(self, "_h", None):
            lib.orc_destroy(self._h)
            self._h = None


def intersect(a: OracleSet, b: OracleSet, pa: np.ndarray, pb: np.ndarray, mode: int = 0,
              nthreads: int = 8, category_bits: int = 4) -> tuple[np.ndarray, dict]:
    """Ideal::intersect per pair (mode 1: the reference -a heuristic, its fill
    classes at 2^category_bits categories).  out == 2 marks pairs on which
    the reference would abort."""
    pa = np.ascontiguousarray(pa, dtype=np.uint32)
    pb = np.ascontiguousarray(pb, dtype=np.uint32)
    out = np.zeros(len(pa), dtype=np.uint8)
    stats = np.zeros(3, dtype=np.int64)
    m = mode | ((category_bits if category_bits != 4 else 0) << 8)
    rc = lib.orc_intersect(a._h, b._h, len(pa), _p(pa), _p(pb), _p(out), m, nthreads, _p(stats))
    if rc == ORC_EINVAL:
        raise ValueError("pair index out of range")
    return out, {"bb_pixel_pairs": int(stats[0]), "segment_tests": int(stats[1]), "in_hits": int(stats[2])}


def within(a: OracleSet, b: OracleSet, pa: np.ndarray, pb: np.ndarray, d_km: float = 10.0,
           nthreads: int = 8) -> tuple[np.ndarray, np.ndarray, dict]:
    """Ideal::within per pair (sets rasterized with hierarchy=True).
    Returns (bits (2 = reference aborts), running min distance, stats)."""
    pa = np.ascontiguousarray(pa, dtype=np.uint32)
    pb = np.ascontiguousarray(pb, dtype=np.uint32)
    out = np.zeros(len(pa), dtype=np.uint8)
    md = np.zeros(len(pa), dtype=np.float64)
    stats = np.zeros(3, dtype=np.int64)
    rc = lib.orc_within(a._h, b._h, len(pa), _p(pa), _p(pb), d_km, _p(out), _p(md), nthreads, _p(stats))
    if rc == ORC_EINVAL:
        raise ValueError("pair index out of range")
    return out, md, {"candidate_pixel_pairs": int(stats[0]), "distance_evals": int(stats[1])}


def intersection(a: OracleSet, b: OracleSet, pa: np.ndarray, pb: np.ndarray,
                 nthreads: int = 8) -> tuple[np.ndarray, np.ndarray]:
    """Ideal::intersection per pair up to its boundary-piece statuses, and the
    area of A n B.  Returns (counts (n, 5): [intersection points (-1 window
    all OUT, -2 all IN, -3 the reference aborts), pieces, IN, OUT, BORDER],
    areas (n,))."""
    pa = np.ascontiguousarray(pa, dtype=np.uint32)
    pb = np.ascontiguousarray(pb, dtype=np.uint32)
    counts = np.zeros((len(pa), 5), dtype=np.int32)
    area = np.zeros(len(pa), dtype=np.float64)
    rc = lib.orc_intersection(a._h, b._h, len(pa), _p(pa), _p(pb), _p(counts), _p(area), nthreads)
    if rc == ORC_EINVAL:
        raise ValueError("pair index out of range")
    return counts, area


def candidates_within(s: OracleSet, d_km: float = 10.0) -> tuple[np.ndarray, np.ndarray]:
    n = lib.orc_candidates_within(s._h, d_km, 0, None, None)
    pa = np.zeros(max(n, 1), np.uint32)
    pb = np.zeros(max(n, 1), np.uint32)
    lib.orc_candidates_within(s._h, d_km, n, _p(pa), _p(pb))
    return pa[:n], pb[:n]


def candidates(a: OracleSet, b: OracleSet) -> tuple[np.ndarray, np.ndarray]:
    n = lib.orc_candidates(a._h, b._h, 0, None, None)
    pa = np.zeros(max(n, 1), np.uint32)
    pb = np.zeros(max(n, 1), np.uint32)
    lib.orc_candidates(a._h, b._h, n, _p(pa), _p(pb))
    return pa[:n], pb[:n]


def segment_intersect(a, b, c, d) -> bool:
    v = np.array([a[0], a[1], b[0], b[1], c[0], c[1], d[0], d[1]], dtype=np.float32)
    return bool(lib.orc_segment_intersect(_p(v)))


def double_to_int(v: float) -> int:
    return int(lib.orc_double_to_int(float(v)))


def point_to_segment_distance(p, p1, p2, geography: bool = True) -> float:
    a = np.asarray(p, np.float32)
    b = np.asarray(p1, np.float32)
    c = np.asarray(p2, np.float32)
    return float(lib.orc_point_to_segment_distance(_p(a), _p(b), _p(c), int(geography)))
