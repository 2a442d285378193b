"""GPU parity: within-distance decision (hierarchical box-distance descent and
truncated segment distances) through the C ABI, bit-exact in decisions,
running minimum distances and work counters against the oracle."""
import numpy as np
import pytest

import oracle
from farm_amd import capi, datagen
from tests.helpers import rings_from_list

pytestmark = pytest.mark.gpu


def _run(dev, ro, xy, d_km=10.0, vpr=10, nlow=1, pairs=None, window=None):
    O = oracle.OracleSet(ro, xy)
    O.rasterize(vpr, hierarchy=True, nlow=nlow)
    H = capi.HostSet.from_rings(ro, xy)
    H.rasterize_h(vpr, nlow)
    if pairs is None:
        pa, pb = capi.candidates(H, H, expand_km=d_km, skip_self=True)
    else:
        pa, pb = (np.asarray(p, np.uint32) for p in pairs)
    print(f"  within d={d_km} vpr={vpr} nlow={nlow} polygons={len(O)} pairs={len(pa)}", flush=True)
    want, wmd, wst = oracle.within(O, O, pa, pb, d_km)
    dev.upload_polygons(H)
    dev.upload_pairs(pa, pb)
    lo, hi = (0, len(pa)) if window is None else window
    bits, md, st = dev.within(d_km, lo, hi)
    want, wmd = want[lo:hi], wmd[lo:hi]
    diff = np.nonzero(bits != want)[0]
    assert diff.size == 0, f"{diff.size} decisions differ, first {diff[:10]}: gpu {bits[diff[:10]]} cpu {want[diff[:10]]}"
    mdiff = np.nonzero(md.view(np.uint64) != wmd.view(np.uint64))[0]
    assert mdiff.size == 0, f"{mdiff.size} min distances differ, first {mdiff[:5]}: {md[mdiff[:5]]} vs {wmd[mdiff[:5]]}"
    assert st["found"] == int((want == 1).sum())
    if window is None:
        assert st["bb_pixel_pairs"] == wst["candidate_pixel_pairs"]
        assert st["segment_tests"] == wst["distance_evals"]
    return want, md, st


def test_lakes_10km(device):
    ro, xy = datagen.lake_set(41, 1500, hi=1.4)
    want, md, st = _run(device, ro, xy)
    assert (want == 1).sum() > 1000 and (want == 0).sum() > 1000
    assert (md != 100000.0).sum() > 1000  # the refinement path is exercised


@pytest.mark.parametrize("d_km", [0.0, 2.0, 25.0])
def test_lakes_other_distances(device, d_km):
    ro, xy = datagen.lake_set(5, 500, hi=0.8)
    _run(device, ro, xy, d_km=d_km)


def test_fine_rasters_take_the_global_path(device):
    # vpr 1: finest layers beyond 1024 pixels, so these pairs run with
    # global-memory pixel lists
    ro, xy = datagen.lake_set(9, 250, hi=0.6)
    _run(device, ro, xy, vpr=1)
    _run(device, ro, xy, vpr=3, nlow=20)


def test_mixed_sizes_and_pair_window(device):
    cro, cxy = datagen.complex_set(5, 6, hi=0.8, nv_cap=4000)
    lro, lxy = datagen.lake_set(6, 300, hi=0.8)
    ro = np.concatenate([cro, lro[1:] + cro[-1]])
    xy = np.concatenate([cxy, lxy])
    _run(device, ro, xy)
    _run(device, ro, xy, window=(100, 900))


def _square(x0, y0, s, per_edge=8):
    t = np.arange(per_edge, dtype=np.float64) / per_edge
    pts = []
    for (ax, ay), (bx, by) in [((x0, y0), (x0 + s, y0)), ((x0 + s, y0), (x0 + s, y0 + s)),
                               ((x0 + s, y0 + s), (x0, y0 + s)), ((x0, y0 + s), (x0, y0))]:
        pts.append(np.stack([ax + (bx - ax) * t, ay + (by - ay) * t], 1))
    return np.vstack(pts).astype(np.float32)


def test_hand_built_cases(device):
    # gaps around 10 km, a contained square, grid-aligned edges, identical
    # polygons and a triangle with too few vertices for a second layer
    polys = [
        _square(0.3, 0.5, 0.02, 16),
        _square(0.3 + 0.02 + 0.0898, 0.5, 0.02, 16),   # ~10.0 km east
        _square(0.3 + 0.02 + 0.0944, 0.52, 0.02, 16),  # ~10.5 km
        _square(0.013, 0.017, 1.0, 40),
        _square(0.49, 0.49, 0.04),
        _square(0.063, 0.49, 0.04),
        _square(0.0, 0.0, 1.0, 40),
        _square(0.05, 0.48, 0.04),
        np.array([[0.7, 0.7], [0.71, 0.7], [0.705, 0.71]], np.float32),
        _square(0.3, 0.5, 0.02, 16),
    ]
    ro, xy = rings_from_list(polys)
    n = len(polys)
    pa, pb = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    _run(device, ro, xy, pairs=(pa.ravel(), pb.ravel()))


def test_rejected_polygons_answer_2(device):
    # a one-pixel raster 500 degrees wide: no universal grid step is that
    # large, so MyRaster::grid_align's assertion fires in the reference; every
    # pair with that polygon answers 2
    big = np.array([[0, 0], [500, 0], [500, 50], [0, 50]], np.float32)
    ro, xy = datagen.lake_set(12, 40, hi=0.3)
    rings = [big] + [xy[int(ro[i]):int(ro[i + 1])] for i in range(len(ro) - 1)]
    n = len(rings)
    pa, pb = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    want, _, _ = _run(device, *rings_from_list(rings), pairs=(pa.ravel(), pb.ravel()))
    assert (want == 2).sum() == 2 * n - 1


def test_empty_window(device):
    ro, xy = datagen.lake_set(41, 50, hi=0.3)
    H = capi.HostSet.from_rings(ro, xy)
    H.rasterize_h(10, 1)
    device.upload_polygons(H)
    device.upload_pairs(np.zeros(1, np.uint32), np.ones(1, np.uint32))
    bits, md, st = device.within(10.0, 0, 0)
    assert len(bits) == 0 and st["found"] == 0
