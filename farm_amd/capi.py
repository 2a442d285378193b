"""ctypes binding of libfarm.so (include/farm_capi.h, include/farm_host.h).

This is the host-side mirror of the reference's query driver: load polygon
sets (.dt / WKT), generate candidate pairs with the host R-tree, rasterize on
the host, upload to the GPU and run the polygon-join kernels.  The library is
required: importing this module raises if libfarm.so is missing, and there is
no CPU fallback for any device call.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfarm.so")

FARM_OK = 0
FARM_EINVAL = 1
FARM_ENOMEM = 2
FARM_EDEVICE = 3
FARM_ECAPACITY = 4
FARM_ESTATE = 5

MODE_EXACT_PLAIN = 0
MODE_EXACT_BATCHED = 1
MODE_APPROX_HEURISTIC = 2
MODE_APPROX_LEARNED = 3

FLAG_NO_PRUNE = 1


class FarmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libfarm error {code}: {msg}")
        self.code = code


class RasterHdr(C.Structure):
    _fields_ = [
        ("mbr", C.c_double * 4),
        ("step", C.c_double * 2),
        ("dimx", C.c_int32),
        ("dimy", C.c_int32),
        ("pix_start", C.c_uint64),
        ("pixoff_start", C.c_uint64),
        ("seq_start", C.c_uint64),
        ("vert_start", C.c_uint64),
        ("nverts", C.c_uint32),
        ("nseqs", C.c_uint32),
        ("vline_start", C.c_uint64),
        ("vnode_start", C.c_uint64),
        ("layer_start", C.c_uint64),
        ("nlayers", C.c_int32),
        ("flags", C.c_int32),
    ]


assert C.sizeof(RasterHdr) == 128

HDR_REJECTED = 1


class Layer(C.Structure):
    _fields_ = [
        ("low", C.c_double * 2),
        ("step", C.c_double),
        ("dimx", C.c_int32),
        ("dimy", C.c_int32),
        ("status_start", C.c_uint64),
    ]


assert C.sizeof(Layer) == 40
LAYER_DTYPE = np.dtype([("low", "<f8", (2,)), ("step", "<f8"), ("dimx", "<i4"), ("dimy", "<i4"),
                        ("status_start", "<u8")])
assert LAYER_DTYPE.itemsize == 40


class RingSet(C.Structure):
    _fields_ = [("npolys", C.c_int64), ("vert_off", C.c_void_p), ("x", C.c_void_p), ("y", C.c_void_p)]


class RasterInfo(C.Structure):
    _fields_ = [("nrejected", C.c_int64), ("first_rejected", C.c_int64), ("nverts", C.c_uint64),
                ("nstatus", C.c_uint64), ("npixoff", C.c_uint64), ("nseqs", C.c_uint64), ("nvoff", C.c_uint64),
                ("nvnodes", C.c_uint64), ("nlayers", C.c_uint64), ("nlstatus", C.c_uint64),
                ("build_ms", C.c_double), ("upload_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class PolySetBuf(C.Structure):
    _fields_ = [("hdr", C.c_void_p), ("status", C.c_void_p), ("pixoff", C.c_void_p), ("seqs", C.c_void_p),
                ("voff", C.c_void_p), ("vnodes", C.c_void_p), ("layers", C.c_void_p), ("lstatus", C.c_void_p),
                ("areas", C.c_void_p), ("fill", C.c_void_p)]

HDR_DTYPE = np.dtype(
    [
        ("mbr", "<f8", (4,)),
        ("step", "<f8", (2,)),
        ("dimx", "<i4"),
        ("dimy", "<i4"),
        ("pix_start", "<u8"),
        ("pixoff_start", "<u8"),
        ("seq_start", "<u8"),
        ("vert_start", "<u8"),
        ("nverts", "<u4"),
        ("nseqs", "<u4"),
        ("vline_start", "<u8"),
        ("vnode_start", "<u8"),
        ("layer_start", "<u8"),
        ("nlayers", "<i4"),
        ("flags", "<i4"),
    ]
)
assert HDR_DTYPE.itemsize == 128


class PolySet(C.Structure):
    _fields_ = [
        ("npolys", C.c_int64),
        ("hdr", C.POINTER(RasterHdr)),
        ("status", C.POINTER(C.c_uint8)),
        ("nstatus", C.c_uint64),
        ("pixoff", C.POINTER(C.c_uint32)),
        ("npixoff", C.c_uint64),
        ("seqs", C.POINTER(C.c_uint32)),
        ("nseqs", C.c_uint64),
        ("x", C.POINTER(C.c_float)),
        ("y", C.POINTER(C.c_float)),
        ("nverts", C.c_uint64),
        ("voff", C.POINTER(C.c_uint32)),
        ("nvoff", C.c_uint64),
        ("vnodes", C.POINTER(C.c_double)),
        ("nvnodes", C.c_uint64),
        ("layers", C.POINTER(Layer)),
        ("nlayers", C.c_uint64),
        ("lstatus", C.POINTER(C.c_uint8)),
        ("nlstatus", C.c_uint64),
        ("areas", C.POINTER(C.c_double)),
        ("fill", C.POINTER(C.c_uint8)),
        ("category_count", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("tile_a", C.c_int32),
        ("tile_b", C.c_int32),
        ("order", C.c_int32),
        ("approx_threshold", C.c_float),
        ("approx_reject", C.c_float),
        ("merge_threshold", C.c_float),
        ("flags", C.c_int32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("tile_a", C.c_int32),
        ("chunk", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("shape_bucket", C.c_int32),
        ("ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Piece(C.Structure):
    _fields_ = [
        ("pair", C.c_uint32),
        ("ring", C.c_int32),
        ("first", C.c_int32),
        ("nvert", C.c_int32),
        ("ax", C.c_float),
        ("ay", C.c_float),
        ("bx", C.c_float),
        ("by", C.c_float),
    ]


assert C.sizeof(Piece) == 32
PIECE_DTYPE = np.dtype([("pair", "<u4"), ("ring", "<i4"), ("first", "<i4"), ("nvert", "<i4"), ("ax", "<f4"),
                        ("ay", "<f4"), ("bx", "<f4"), ("by", "<f4")])

NFEAT = 80
MLP_NPARAMS = 27138


class Stats(C.Structure):
    _fields_ = [
        ("pairs", C.c_uint64),
        ("in_hits", C.c_uint64),
        ("bb_pixel_pairs", C.c_uint64),
        ("refine_items", C.c_uint64),
        ("segment_tests", C.c_uint64),
        ("filter_bytes", C.c_uint64),
        ("refine_bytes", C.c_uint64),
        ("found", C.c_uint64),
        ("filter_ms", C.c_double),
        ("refine_ms", C.c_double),
        ("total_ms", C.c_double),
        ("pruned_tests", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every entry point declared in include/farm_capi.h and include/farm_host.h.
EXPORTS = {
    "farm_version": (C.c_char_p, []),
    "farm_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "farm_close": (C.c_int, [C.c_void_p]),
    "farm_last_error": (C.c_char_p, [C.c_void_p]),
    "farm_upload_polygons": (C.c_int, [C.c_void_p, C.POINTER(PolySet)]),
    "farm_upload_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "farm_rasterize": (C.c_int, [C.c_void_p, C.POINTER(RingSet), C.c_int, C.c_int, C.POINTER(RasterInfo)]),
    "farm_download_polyset": (C.c_int, [C.c_void_p, C.POINTER(PolySetBuf)]),
    "farm_set_categories": (C.c_int, [C.c_void_p, C.c_int]),
    "farm_generate_pairs": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_double, C.c_int, C.c_uint32, C.c_uint32,
         C.POINTER(C.c_uint64)],
    ),
    "farm_download_pairs": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]),
    "farm_intersect": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(Opts), C.c_void_p, C.POINTER(Stats)]),
    "farm_intersect_dev": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(Opts), C.c_void_p, C.POINTER(Stats)],
    ),
    "farm_intersect_noalloc": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(Opts), C.c_void_p, C.POINTER(Stats)],
    ),
    "farm_set_pixel_pair_limit": (C.c_int, [C.c_void_p, C.c_uint64]),
    "farm_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "farm_intersect_costs": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(Stats)],
    ),
    "farm_within": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, C.c_void_p, C.c_void_p, C.POINTER(Stats)],
    ),
    "farm_within_dev": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, C.c_void_p, C.c_void_p, C.POINTER(Stats)],
    ),
    "farm_intersection_area": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(Stats)],
    ),
    "farm_intersection_area_dev": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(Stats)],
    ),
    "farm_intersection_pieces": (
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
         C.POINTER(C.c_uint64), C.POINTER(Stats)],
    ),
    "farm_pieces_wkt": (
        C.c_void_p,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)],
    ),
    "farm_pair_features": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]),
    "farm_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "farm_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "farm_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "farm_mlp_loss_grad": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_float, C.c_void_p, C.c_void_p],
    ),
    "farm_mlp_train": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float,
         C.c_uint64, C.c_void_p],
    ),
    "farm_mlp_predict_pairs": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]),
    "farm_autotune": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_int)]),
    "farm_autotune_config": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(Tuning)]),
    "farm_set_tile": (C.c_int, [C.c_void_p, C.c_int]),
    "farm_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "farm_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "farm_tuning_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "farm_tuning_load": (C.c_int, [C.c_void_p, C.c_char_p]),
    "farm_copy_probe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "farm_hostset_from_rings": (C.c_void_p, [C.c_int64, C.c_void_p, C.c_void_p]),
    "farm_hostset_load_dt": (C.c_void_p, [C.c_char_p]),
    "farm_hostset_load_wkt": (C.c_void_p, [C.c_char_p]),
    "farm_hostset_save_dt": (C.c_int, [C.c_void_p, C.c_char_p]),
    "farm_hostset_concat": (C.c_void_p, [C.c_void_p, C.c_void_p]),
    "farm_hostset_free": (None, [C.c_void_p]),
    "farm_hostset_size": (C.c_int64, [C.c_void_p]),
    "farm_hostset_ring": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    "farm_hostset_mbbs": (C.c_int, [C.c_void_p, C.c_void_p]),
    "farm_hostset_closed_rings": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "farm_hostset_rasterize_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                                C.POINTER(RasterInfo)]),
    "farm_set_staging": (C.c_int, [C.c_void_p, C.c_uint64]),
    "farm_hostset_rasterize": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "farm_hostset_rasterize_h": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "farm_hostset_layer": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.POINTER(Layer), C.POINTER(C.c_void_p)]),
    "farm_hostset_view": (C.c_int, [C.c_void_p, C.POINTER(PolySet)]),
    "farm_hostset_areas": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]),
    "farm_hostset_raster_failures": (C.c_int64, [C.c_void_p]),
    "farm_hostset_set_categories": (C.c_int, [C.c_void_p, C.c_int]),
    "farm_hostset_raster_ok": (C.c_int, [C.c_void_p, C.c_void_p]),
    "farm_candidates": (
        C.c_int64,
        [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
    ),
    "farm_free": (None, [C.c_void_p]),
    "farm_km_table": (C.c_int, [C.c_void_p]),
    "farm_hostset_star_polygons": (
        C.c_void_p,
        [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
         C.c_int],
    ),
    "farm_hostset_star_polygons_ex": (
        C.c_void_p,
        [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
         C.c_int, C.c_double, C.c_int],
    ),
}


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


class HostSet:
    """A polygon set on the host (rings as stored in .dt files)."""

    def __init__(self, handle: int):
        if not handle:
            raise FarmError(FARM_EINVAL, "could not build polygon set")
        self._h = C.c_void_p(handle)

    @classmethod
    def from_rings(cls, ring_off: np.ndarray, xy: np.ndarray) -> "HostSet":
        ring_off = np.ascontiguousarray(ring_off, dtype=np.uint64)
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        return cls(lib.farm_hostset_from_rings(len(ring_off) - 1, _ptr(ring_off), _ptr(xy)))

    @classmethod
    def star_polygons(cls, seed: int, n: int, lo: float, hi: float, nv_lo: int, nv_hi: int, r_mu: float,
                      r_sigma: float, amp: float, nthreads: int | None = None) -> "HostSet":
        """Seeded star polygons generated in parallel by the library
        (farm_hostset_star_polygons)."""
        return cls(lib.farm_hostset_star_polygons(seed, n, lo, hi, nv_lo, nv_hi, r_mu, r_sigma, amp,
                                                  nthreads or default_threads()))

    @classmethod
    def complex_polygons(cls, seed: int, n: int, lo: float, hi: float, nv_cap: int = 200_000,
                         nthreads: int | None = None) -> "HostSet":
        """The complex polygons of the headline workload (datagen.complex_set's
        distribution), generated in parallel by the library."""
        return cls(lib.farm_hostset_star_polygons_ex(seed, n, lo, hi, 1000, nv_cap, 1.2, float(np.log(0.8)), 0.6, 1,
                                                     0.3, nthreads or default_threads()))

    @classmethod
    def lake_polygons(cls, seed: int, n: int, lo: float, hi: float, nthreads: int | None = None) -> "HostSet":
        """Lake-like polygons (datagen.lake_set's distribution), generated in
        parallel by the library."""
        return cls(lib.farm_hostset_star_polygons_ex(seed, n, lo, hi, 20, 2000, 0.0, float(np.log(0.05)), 0.8, 0,
                                                     0.4, nthreads or default_threads()))

    def subset_rings(self, idx: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """(ring_off, xy) of polygons idx as stored, for the oracle."""
        parts = [self.ring(int(i)) for i in idx]
        off = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint64)
        return off, (np.concatenate(parts) if parts else np.zeros((0, 2), np.float32))

    @classmethod
    def load_dt(cls, path: str) -> "HostSet":
        return cls(lib.farm_hostset_load_dt(path.encode()))

    @classmethod
    def load_wkt(cls, path: str) -> "HostSet":
        return cls(lib.farm_hostset_load_wkt(path.encode()))

    @classmethod
    def concat(cls, a: "HostSet", b: "HostSet") -> "HostSet":
        return cls(lib.farm_hostset_concat(a._h, b._h))

    def save_dt(self, path: str) -> None:
        rc = lib.farm_hostset_save_dt(self._h, path.encode())
        if rc:
            raise FarmError(rc, f"cannot write {path}")

    def __len__(self) -> int:
        return int(lib.farm_hostset_size(self._h))

    def ring(self, i: int) -> np.ndarray:
        p = C.c_void_p()
        n = C.c_int64()
        rc = lib.farm_hostset_ring(self._h, i, C.byref(p), C.byref(n))
        if rc:
            raise FarmError(rc, f"no polygon {i}")
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n.value, 2)).copy()

    def mbbs(self) -> np.ndarray:
        out = np.zeros((len(self), 4), dtype=np.float64)
        lib.farm_hostset_mbbs(self._h, _ptr(out))
        return out

    def rasterize(self, vpr: int = 10, nthreads: int | None = None) -> None:
        rc = lib.farm_hostset_rasterize(self._h, vpr, nthreads or default_threads())
        if rc:
            raise FarmError(rc, "rasterize failed")

    def rasterize_h(self, vpr: int = 10, nlow: int = 1, nthreads: int | None = None) -> None:
        """-h rasterization: aligned grid, coarse layers and per-pixel areas
        (needed by Device.within)."""
        rc = lib.farm_hostset_rasterize_h(self._h, vpr, nlow, nthreads or default_threads())
        if rc:
            raise FarmError(rc, "rasterize_h failed")

    def layer(self, i: int, k: int) -> dict:
        """Layer k (0 = coarsest) of polygon i after rasterize_h."""
        L = Layer()
        p = C.c_void_p()
        rc = lib.farm_hostset_layer(self._h, i, k, C.byref(L), C.byref(p))
        if rc:
            raise FarmError(rc, f"no layer {k} of polygon {i}")
        n = L.dimx * L.dimy
        st = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n,)).copy()
        return {"low": np.array(L.low[:]), "step": L.step, "dimx": L.dimx, "dimy": L.dimy, "status": st}

    def set_categories(self, bits: int) -> None:
        """-w bits: fill classes at 2^bits categories for the approximate
        mode and the learned features (after rasterizing)."""
        rc = lib.farm_hostset_set_categories(self._h, bits)
        if rc:
            raise FarmError(rc, f"set_categories({bits}) failed")

    def raster_failures(self) -> int:
        """Index of the first polygon the reference raster builder rejects, or -1."""
        return int(lib.farm_hostset_raster_failures(self._h))

    def raster_ok(self) -> np.ndarray:
        out = np.zeros(len(self), dtype=np.uint8)
        rc = lib.farm_hostset_raster_ok(self._h, _ptr(out))
        if rc:
            raise FarmError(rc, "polygon set is not rasterized")
        return out.astype(bool)

    def view(self) -> PolySet:
        ps = PolySet()
        rc = lib.farm_hostset_view(self._h, C.byref(ps))
        if rc:
            raise FarmError(rc, "polygon set is not rasterized")
        return ps

    def layout(self) -> dict:
        """Numpy views (no copy) of the device layout, valid while self lives."""
        ps = self.view()

        def arr(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,))

        hdr = np.ctypeslib.as_array(C.cast(ps.hdr, C.POINTER(C.c_uint8)), shape=(ps.npolys * 128,)).view(HDR_DTYPE)
        return {
            "hdr": hdr,
            "status": arr(ps.status, ps.nstatus, np.uint8),
            "pixoff": arr(ps.pixoff, ps.npixoff, np.uint32),
            "seqs": arr(ps.seqs, 2 * ps.nseqs, np.uint32).reshape(-1, 2),
            "x": arr(ps.x, ps.nverts, np.float32),
            "y": arr(ps.y, ps.nverts, np.float32),
            "voff": arr(ps.voff, ps.nvoff, np.uint32),
            "vnodes": arr(ps.vnodes, ps.nvnodes, np.float64),
            "fill": arr(ps.fill, ps.nstatus, np.uint8) if ps.fill else None,
            "category_count": int(ps.category_count),
            "layers": (np.ctypeslib.as_array(C.cast(ps.layers, C.POINTER(C.c_uint8)), shape=(ps.nlayers * 40,))
                       .view(LAYER_DTYPE) if ps.nlayers else np.zeros(0, LAYER_DTYPE)),
            "lstatus": arr(ps.lstatus, ps.nlstatus, np.uint8),
            "areas": arr(ps.areas, ps.nstatus, np.float64) if ps.areas else None,
        }

    def closed_rings(self, nthreads: int | None = None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(vert_off, x, y): every ring with its first vertex repeated, the
        input of Device.rasterize."""
        n = len(self)
        off = np.zeros(n + 1, np.uint64)
        if lib.farm_hostset_closed_rings(self._h, _ptr(off), None, None, 1):
            raise FarmError(FARM_EINVAL, "closed_rings failed")
        x = np.empty(max(int(off[-1]), 1), np.float32)
        y = np.empty(max(int(off[-1]), 1), np.float32)
        if lib.farm_hostset_closed_rings(self._h, _ptr(off), _ptr(x), _ptr(y), nthreads or default_threads()):
            raise FarmError(FARM_EINVAL, "closed_rings failed")
        return off, x[:int(off[-1])], y[:int(off[-1])]

    def areas(self, i: int, npix: int) -> np.ndarray:
        p = C.c_void_p()
        rc = lib.farm_hostset_areas(self._h, i, C.byref(p))
        if rc:
            raise FarmError(rc, "no areas")
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(npix,)).copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.farm_hostset_free(h)
            self._h = None


def pieces_wkt(s: HostSet, pa: np.ndarray, pb: np.ndarray, pieces: np.ndarray) -> str:
    """The rings of A n B of every pair with pieces, assembled by the library
    (farm_pieces_wkt): lines "<k>\tMULTIPOLYGON (...)" for window index k;
    pa / pb are the window's pair ids into the uploaded set s."""
    pa = np.ascontiguousarray(pa, np.uint32)
    pb = np.ascontiguousarray(pb, np.uint32)
    pieces = np.ascontiguousarray(pieces, PIECE_DTYPE)
    n = C.c_uint64(0)
    p = lib.farm_pieces_wkt(s._h, _ptr(pa), _ptr(pb), len(pa), _ptr(pieces) if len(pieces) else None, len(pieces),
                            C.byref(n))
    if not p:
        raise FarmError(FARM_EINVAL, "farm_pieces_wkt rejected its input")
    try:
        return C.string_at(p, n.value).decode()
    finally:
        lib.farm_free(p)


def wkt_rings(text: str) -> list[np.ndarray]:
    """The rings of a MULTIPOLYGON (((x y, ...)), ...) line, as (n, 2) f64
    arrays of float32 coordinates: like the .dt loader (std::stof), every
    number is rounded to float first, which gives back the printed floats."""
    body = text[text.index("(") + 1:text.rindex(")")]
    return [np.array([[float(v) for v in q.split()] for q in ring.split("))")[0].split(",")],
                     np.float64).astype(np.float32).astype(np.float64)
            for ring in body.split("((")[1:]]


def wkt_multipolygon_area(text: str) -> float:
    """|sum of the rings' signed shoelace areas| (f64): outer rings and holes
    turn opposite ways."""
    total = 0.0
    for pts in wkt_rings(text):
        x, y = pts[:, 0] - pts[0, 0], pts[:, 1] - pts[0, 1]
        total += 0.5 * float(np.sum(x[:-1] * y[1:] - x[1:] * y[:-1]))
    return abs(total)


def candidates(a: HostSet, b: HostSet, expand_km: float = 0.0, skip_self: bool = False,
               nthreads: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """All (i in a, j in b) with inclusively overlapping MBBs, sorted by (j, i)."""
    pa = C.c_void_p()
    pb = C.c_void_p()
    n = lib.farm_candidates(a._h, b._h, expand_km, int(skip_self), nthreads or default_threads(), C.byref(pa),
                            C.byref(pb))
    if n < 0:
        raise FarmError(FARM_ENOMEM, "candidate generation failed")
    try:
        ia = np.ctypeslib.as_array(C.cast(pa, C.POINTER(C.c_uint32)), shape=(max(n, 1),))[:n].copy()
        ib = np.ctypeslib.as_array(C.cast(pb, C.POINTER(C.c_uint32)), shape=(max(n, 1),))[:n].copy()
    finally:
        lib.farm_free(pa)
        lib.farm_free(pb)
    return ia, ib


@dataclass
class IntersectResult:
    bits: np.ndarray | None
    stats: dict


class Device:
    """One libfarm handle on one GPU (one HIP stream)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib.farm_open(device, C.byref(h))
        if rc:
            raise FarmError(rc, f"farm_open({device}) failed")
        self._h = h
        self.npairs = 0

    def _check(self, rc: int, what: str) -> None:
        if rc:
            raise FarmError(rc, f"{what}: {lib.farm_last_error(self._h).decode()}")

    def upload_polygons(self, s: HostSet) -> None:
        ps = s.view()
        self._check(lib.farm_upload_polygons(self._h, C.byref(ps)), "upload_polygons")

    def rasterize(self, vert_off: np.ndarray, x: np.ndarray, y: np.ndarray, vpr: int, nlow: int = 0) -> dict:
        """Rasterize the rings on the device (farm_rasterize) and keep the
        result as the polygon set; returns the array lengths and rejections."""
        off = np.ascontiguousarray(vert_off, np.uint64)
        xs = np.ascontiguousarray(x, np.float32)
        ys = np.ascontiguousarray(y, np.float32)
        rs = RingSet(len(off) - 1, _ptr(off), _ptr(xs), _ptr(ys))
        info = RasterInfo()
        self._check(lib.farm_rasterize(self._h, C.byref(rs), int(vpr), int(nlow), C.byref(info)), "rasterize")
        self.category_bits = 4
        self.raster_info = info.as_dict()
        self.raster_info["npolys"] = len(off) - 1
        return self.raster_info

    def set_categories(self, bits: int) -> None:
        """-w bits on the device set (farm_set_categories)."""
        self._check(lib.farm_set_categories(self._h, int(bits)), "set_categories")
        self.category_bits = int(bits)

    def rasterize_hostset(self, s: "HostSet", vpr: int, nlow: int = 0, nthreads: int | None = None) -> dict:
        """The host set rasterized on the device, its rings streamed through
        pinned staging buffers (farm_hostset_rasterize_device)."""
        info = RasterInfo()
        self._check(lib.farm_hostset_rasterize_device(self._h, s._h, int(vpr), int(nlow),
                                                      nthreads or default_threads(), C.byref(info)),
                    "rasterize_hostset")
        self.raster_info = info.as_dict()
        self.raster_info["npolys"] = len(s)
        self.category_bits = 4
        return self.raster_info

    def set_staging(self, vertices: int) -> None:
        self._check(lib.farm_set_staging(self._h, int(vertices)), "set_staging")

    def download_polyset(self) -> dict:
        """The device set's arrays (after rasterize), as HostSet.layout names them."""
        i = self.raster_info
        out = {
            "hdr": np.zeros(i["npolys"], HDR_DTYPE),
            "status": np.zeros(i["nstatus"], np.uint8),
            "pixoff": np.zeros(i["npixoff"], np.uint32),
            "seqs": np.zeros((i["nseqs"], 2), np.uint32),
            "voff": np.zeros(i["nvoff"], np.uint32),
            "vnodes": np.zeros(i["nvnodes"], np.float64),
            "layers": np.zeros(i["nlayers"], LAYER_DTYPE),
            "lstatus": np.zeros(i["nlstatus"], np.uint8),
            "areas": np.zeros(i["nstatus"], np.float64),
            "fill": np.zeros(i["nstatus"] if getattr(self, "category_bits", 4) != 4 else 0, np.uint8),
        }
        buf = PolySetBuf(*[(_ptr(out[k]) if out[k].size else None) for k in
                           ("hdr", "status", "pixoff", "seqs", "voff", "vnodes", "layers", "lstatus", "areas",
                            "fill")])
        self._check(lib.farm_download_polyset(self._h, C.byref(buf)), "download_polyset")
        return out

    def upload_pairs(self, a: np.ndarray, b: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        self._check(lib.farm_upload_pairs(self._h, _ptr(a), _ptr(b), len(a)), "upload_pairs")
        self.npairs = len(a)

    def generate_pairs(self, mbb_a: np.ndarray, mbb_b: np.ndarray, expand_km: float = 0.0, skip_self: bool = False,
                       a_base: int = 0, b_base: int = 0) -> int:
        """Candidate pairs built on the device from (n, 4) boxes (low x, low
        y, high x, high y): the same pairs, in the same order, as
        candidates(), kept as the pair list; returns their count."""
        a = np.ascontiguousarray(mbb_a, dtype=np.float64).reshape(-1, 4)
        b = np.ascontiguousarray(mbb_b, dtype=np.float64).reshape(-1, 4)
        n = C.c_uint64(0)
        self._check(lib.farm_generate_pairs(self._h, _ptr(a), len(a), _ptr(b), len(b), float(expand_km),
                                            int(bool(skip_self)), a_base, b_base, C.byref(n)), "generate_pairs")
        self.npairs = int(n.value)
        return self.npairs

    def download_pairs(self, begin: int = 0, end: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        end = self.npairs if end is None else end
        a = np.zeros(max(end - begin, 1), np.uint32)
        b = np.zeros(max(end - begin, 1), np.uint32)
        self._check(lib.farm_download_pairs(self._h, begin, end, _ptr(a), _ptr(b)), "download_pairs")
        return a[:end - begin], b[:end - begin]

    def intersect(self, begin: int = 0, end: int | None = None, mode: int = MODE_EXACT_BATCHED,
                  tile_a: int = 0, want_bits: bool = True, order: int = 0, accept: float = 0.0,
                  reject: float | None = None, flags: int = 0, chunk: int = 0) -> IntersectResult:
        """Exact or approximate intersect of pairs [begin, end).  order=1 orders
        the refinement queue by the MLP cost head; mode MODE_APPROX_LEARNED
        accepts logits > accept, rejects logits < reject and refines the rest."""
        end = self.npairs if end is None else end
        o = Opts(mode=mode, tile_a=tile_a, tile_b=chunk, order=order, approx_threshold=accept,
                 approx_reject=accept if reject is None else reject, flags=flags)
        st = Stats()
        out = np.zeros(max(end - begin, 1), dtype=np.uint8) if want_bits else None
        rc = lib.farm_intersect(self._h, begin, end, C.byref(o), _ptr(out) if want_bits else None, C.byref(st))
        self._check(rc, "intersect")
        return IntersectResult(out[: end - begin] if want_bits else None, st.as_dict())

    def intersect_dev(self, out_dev_ptr: int | None, begin: int = 0, end: int | None = None,
                      mode: int = MODE_EXACT_BATCHED, tile_a: int = 0, order: int = 0, accept: float = 0.0,
                      reject: float | None = None, flags: int = 0, chunk: int = 0) -> dict:
        """intersect() with the 0/1 results left in device memory at
        out_dev_ptr (end - begin bytes; None: counts only)."""
        end = self.npairs if end is None else end
        o = Opts(mode=mode, tile_a=tile_a, tile_b=chunk, order=order, approx_threshold=accept,
                 approx_reject=accept if reject is None else reject, flags=flags)
        st = Stats()
        rc = lib.farm_intersect_dev(self._h, begin, end, C.byref(o), C.c_void_p(out_dev_ptr or 0), C.byref(st))
        self._check(rc, "intersect_dev")
        return st.as_dict()

    def intersect_noalloc(self, out_dev_ptr: int | None, begin: int = 0, end: int | None = None,
                          mode: int = MODE_EXACT_BATCHED, tile_a: int = 0) -> dict:
        """intersect_dev() through farm_intersect_noalloc: no device
        allocation, the library's fixed per-GPU scratch."""
        end = self.npairs if end is None else end
        o = Opts(mode=mode, tile_a=tile_a)
        st = Stats()
        rc = lib.farm_intersect_noalloc(self._h, begin, end, C.byref(o), C.c_void_p(out_dev_ptr or 0), C.byref(st))
        self._check(rc, "intersect_noalloc")
        return st.as_dict()

    def set_pixel_pair_limit(self, limit: int) -> None:
        self._check(lib.farm_set_pixel_pair_limit(self._h, int(limit)), "set_pixel_pair_limit")

    def set_graphs(self, on: bool) -> None:
        self._check(lib.farm_set_graphs(self._h, int(bool(on))), "set_graphs")

    def intersect_costs(self, begin: int = 0, end: int | None = None) -> tuple[np.ndarray, np.ndarray, dict]:
        """Exact bits plus the segment tests executed per pair."""
        end = self.npairs if end is None else end
        n = end - begin
        bits = np.zeros(max(n, 1), np.uint8)
        tests = np.zeros(max(n, 1), np.uint32)
        st = Stats()
        self._check(lib.farm_intersect_costs(self._h, begin, end, _ptr(bits), _ptr(tests), C.byref(st)),
                    "intersect_costs")
        return bits[:n], tests[:n], st.as_dict()

    def within(self, d_km: float = 10.0, begin: int = 0, end: int | None = None,
               want_results: bool = True) -> tuple[np.ndarray | None, np.ndarray | None, dict]:
        """Within-distance decision of pairs [begin, end): bits (1 within, 0
        not, 2 a polygon is rejected by the raster builder), the running
        minimum distance of the refinement, and counters.  want_results=False
        leaves the results on the device (counts only)."""
        end = self.npairs if end is None else end
        n = end - begin
        st = Stats()
        if not want_results:
            self._check(lib.farm_within(self._h, begin, end, d_km, None, None, C.byref(st)), "within")
            return None, None, st.as_dict()
        bits = np.zeros(max(n, 1), np.uint8)
        md = np.zeros(max(n, 1), np.float64)
        self._check(lib.farm_within(self._h, begin, end, d_km, _ptr(bits), _ptr(md), C.byref(st)), "within")
        return bits[:n], md[:n], st.as_dict()

    def intersection_area(self, begin: int = 0, end: int | None = None,
                          want_results: bool = True) -> tuple[np.ndarray | None, np.ndarray | None, dict]:
        """Intersection of pairs [begin, end): counts (n, 5) [crossing points
        (-1 window all OUT, -2 all IN, -3 rejected polygon), pieces, IN, OUT,
        BORDER] and the f64 area of A n B per pair."""
        end = self.npairs if end is None else end
        n = end - begin
        st = Stats()
        if not want_results:
            self._check(lib.farm_intersection_area(self._h, begin, end, None, None, C.byref(st)), "intersection")
            return None, None, st.as_dict()
        counts = np.zeros((max(n, 1), 5), np.int32)
        area = np.zeros(max(n, 1), np.float64)
        self._check(lib.farm_intersection_area(self._h, begin, end, _ptr(counts), _ptr(area), C.byref(st)),
                    "intersection")
        return counts[:n], area[:n], st.as_dict()

    def intersection_pieces(self, begin: int = 0, end: int | None = None,
                            cap: int = 1 << 20) -> tuple[np.ndarray, np.ndarray, np.ndarray, dict]:
        """intersection_area() plus the IN pieces of the boundary of A n B of
        every pair (a PIECE_DTYPE array, pair = index in the window)."""
        end = self.npairs if end is None else end
        n = end - begin
        counts = np.zeros((max(n, 1), 5), np.int32)
        area = np.zeros(max(n, 1), np.float64)
        while True:
            pcs = np.zeros(max(cap, 1), PIECE_DTYPE)
            npc = C.c_uint64(0)
            st = Stats()
            rc = lib.farm_intersection_pieces(self._h, begin, end, _ptr(counts), _ptr(area), _ptr(pcs), cap,
                                              C.byref(npc), C.byref(st))
            if rc == FARM_ECAPACITY and npc.value > cap:
                cap = int(npc.value)
                continue
            self._check(rc, "intersection_pieces")
            return counts[:n], area[:n], pcs[:npc.value], st.as_dict()

    def pair_features(self, begin: int = 0, end: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        end = self.npairs if end is None else end
        n = end - begin
        feat = np.zeros((max(n, 1), NFEAT), np.float32)
        dec = np.zeros(max(n, 1), np.int8)
        self._check(lib.farm_pair_features(self._h, begin, end, _ptr(feat), _ptr(dec)), "pair_features")
        return feat[:n], dec[:n]

    def mlp_set_params(self, params: np.ndarray, norm: np.ndarray | None = None) -> None:
        params = np.ascontiguousarray(params, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._check(lib.farm_mlp_set_params(self._h, _ptr(params), len(params), None if nm is None else _ptr(nm)),
                    "mlp_set_params")

    def mlp_get_params(self) -> np.ndarray:
        out = np.zeros(MLP_NPARAMS, np.float32)
        self._check(lib.farm_mlp_get_params(self._h, _ptr(out), MLP_NPARAMS), "mlp_get_params")
        return out

    def mlp_forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros((len(X), 2), np.float32)
        self._check(lib.farm_mlp_forward(self._h, _ptr(X), len(X), _ptr(out)), "mlp_forward")
        return out

    def mlp_loss_grad(self, X, y, c, lam: float) -> tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        c = np.ascontiguousarray(c, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(MLP_NPARAMS, np.float32)
        self._check(lib.farm_mlp_loss_grad(self._h, _ptr(X), _ptr(y), _ptr(c), len(X), lam, _ptr(loss), _ptr(g)),
                    "mlp_loss_grad")
        return float(loss[0]), g

    def mlp_train(self, X, y, c, epochs: int, batch: int, lr: float, lam: float, seed: int = 0) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        c = np.ascontiguousarray(c, np.float32)
        losses = np.zeros(epochs, np.float32)
        self._check(lib.farm_mlp_train(self._h, _ptr(X), _ptr(y), _ptr(c), len(X), epochs, batch, lr, lam, seed,
                                       _ptr(losses)), "mlp_train")
        return losses

    def mlp_predict_pairs(self, begin: int = 0, end: int | None = None) -> np.ndarray:
        end = self.npairs if end is None else end
        out = np.zeros((max(end - begin, 1), 2), np.float32)
        self._check(lib.farm_mlp_predict_pairs(self._h, begin, end, _ptr(out)), "mlp_predict_pairs")
        return out[: end - begin]

    def autotune(self, begin: int = 0, end: int | None = None, cache: str | None = None) -> dict:
        """Pick the refinement configuration (tile width, chunk size,
        workgroups per CU) on pairs [begin, end); it becomes the default of
        later calls and the entry of the uploaded set's shape bucket.  With a
        cache path (farm_tuning_save's text format), a stored entry for the
        set's bucket is applied instead of re-timing, and a new choice is
        added to the file."""
        if cache and os.path.exists(cache):
            self._check(lib.farm_tuning_load(self._h, cache.encode()), "tuning_load")
            t = self.tuning()
            if t["ms"] > 0:  # the bucket had an entry
                return t
        end = self.npairs if end is None else end
        best = Tuning()
        self._check(lib.farm_autotune_config(self._h, begin, end, C.byref(best)), "autotune_config")
        if cache:
            self._check(lib.farm_tuning_save(self._h, cache.encode()), "tuning_save")
        return best.as_dict()

    def tuning(self) -> dict:
        """The configuration calls with tile_a / tile_b == 0 use."""
        t = Tuning()
        self._check(lib.farm_get_tuning(self._h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, tile_a: int, chunk: int, blocks_per_cu: int = 8) -> None:
        t = Tuning(tile_a=tile_a, chunk=chunk, blocks_per_cu=blocks_per_cu)
        self._check(lib.farm_set_tuning(self._h, C.byref(t)), "set_tuning")

    def tuning_save(self, path: str) -> None:
        self._check(lib.farm_tuning_save(self._h, path.encode()), "tuning_save")

    def tuning_load(self, path: str) -> None:
        self._check(lib.farm_tuning_load(self._h, path.encode()), "tuning_load")

    def copy_probe(self, nbytes: int, iters: int = 10) -> float:
        """GB/s (read + write) of a plain streaming copy of nbytes."""
        g = C.c_double()
        self._check(lib.farm_copy_probe(self._h, nbytes, iters, C.byref(g)), "copy_probe")
        return g.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.farm_close(self._h)
            self._h = None

    def __del__(self):
        self.close()
