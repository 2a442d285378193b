"""C ABI contract on the device: status codes for misuse, count-only calls,
device-pointer variants and repeated calls on one handle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from farm_amd import capi, datagen

pytestmark = pytest.mark.gpu


def _lakes(n=120, hi=0.5, seed=3):
    ro, xy = datagen.lake_set(seed, n, 0.0, hi)
    return ro, xy


def test_wrong_set_kind_and_bad_windows(device):
    ro, xy = _lakes()
    flat = capi.HostSet.from_rings(ro, xy)
    flat.rasterize(10)
    device.upload_polygons(flat)
    pa, pb = capi.candidates(flat, flat, expand_km=10.0, skip_self=True)
    device.upload_pairs(pa, pb)
    # within needs the -h layers
    with pytest.raises(capi.FarmError) as e:
        device.within(10.0)
    assert e.value.code == capi.FARM_ESTATE
    # pair windows outside the uploaded pairs
    for lo, hi in ((0, len(pa) + 1), (5, 3)):
        with pytest.raises(capi.FarmError) as e:
            device.intersect(lo, hi)
        assert e.value.code == capi.FARM_EINVAL
        with pytest.raises(capi.FarmError) as e:
            device.intersection_area(lo, hi)
        assert e.value.code == capi.FARM_EINVAL
    hier = capi.HostSet.from_rings(ro, xy)
    hier.rasterize_h(10, 1)
    device.upload_polygons(hier)
    device.upload_pairs(pa, pb)
    # intersection needs a set without -h
    with pytest.raises(capi.FarmError) as e:
        device.intersection_area()
    assert e.value.code == capi.FARM_ESTATE
    with pytest.raises(capi.FarmError) as e:
        device.within(-1.0)
    assert e.value.code == capi.FARM_EINVAL
    # pairs referencing polygons outside the set
    with pytest.raises(capi.FarmError) as e:
        device.upload_pairs(np.array([0], np.uint32), np.array([len(hier)], np.uint32))
    assert e.value.code == capi.FARM_EINVAL


def test_count_only_calls_and_repeats(device):
    ro, xy = _lakes(200, 0.6, 5)
    hier = capi.HostSet.from_rings(ro, xy)
    hier.rasterize_h(10, 1)
    pa, pb = capi.candidates(hier, hier, expand_km=10.0, skip_self=True)
    device.upload_polygons(hier)
    device.upload_pairs(pa, pb)
    bits, md, st = device.within(10.0)
    _, _, st2 = device.within(10.0, want_results=False)
    assert st2["found"] == st["found"] == int((bits == 1).sum())
    assert st2["segment_tests"] == st["segment_tests"]
    # windows partition the pairs: concatenated results equal the whole
    cuts = [0, len(pa) // 3, len(pa) // 2, len(pa)]
    parts = [device.within(10.0, cuts[k], cuts[k + 1]) for k in range(3)]
    assert np.array_equal(np.concatenate([p[0] for p in parts]), bits)
    assert np.array_equal(np.concatenate([p[1] for p in parts]).view(np.uint64), md.view(np.uint64))


def test_device_pointer_variants(device):
    (sa, sxy), (ta, txy) = datagen.star_pair_set(8, 200, 200, 0.0, 20.0)
    A, B = capi.HostSet.from_rings(sa, sxy), capi.HostSet.from_rings(ta, txy)
    pa, pb = capi.candidates(A, B)
    both = capi.HostSet.concat(A, B)
    both.rasterize(10)
    device.upload_polygons(both)
    device.upload_pairs(pa, pb + len(A))
    n = len(pa)
    want = device.intersect().bits
    wc, wa, _ = device.intersection_area()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    bufs = [C.c_void_p() for _ in range(3)]
    for b, size in zip(bufs, (n, 20 * n, 8 * n)):
        assert hip.hipMalloc(C.byref(b), max(size, 16)) == 0
    try:
        device.intersect_dev(bufs[0].value)
        st = capi.Stats()
        rc = capi.lib.farm_intersection_area_dev(device._h, 0, n, bufs[1], bufs[2], C.byref(st))
        assert rc == 0
        got = np.zeros(n, np.uint8)
        gc = np.zeros((n, 5), np.int32)
        ga = np.zeros(n, np.float64)
        for host, dev, size in ((got, bufs[0], n), (gc, bufs[1], 20 * n), (ga, bufs[2], 8 * n)):
            assert hip.hipMemcpy(host.ctypes.data, dev, size, 2) == 0  # device to host
    finally:
        for b in bufs:
            hip.hipFree(b)
    assert np.array_equal(got, want)
    assert np.array_equal(gc, wc) and np.array_equal(ga.view(np.uint64), wa.view(np.uint64))


def _star_workload(seed=11, n=300):
    (sa, sxy), (ta, txy) = datagen.star_pair_set(seed, n, n, 0.0, 30.0)
    A, B = capi.HostSet.from_rings(sa, sxy), capi.HostSet.from_rings(ta, txy)
    pa, pb = capi.candidates(A, B)
    both = capi.HostSet.concat(A, B)
    both.rasterize(10)
    OA, OB = oracle.OracleSet(sa, sxy), oracle.OracleSet(ta, txy)
    OA.rasterize(10)
    OB.rasterize(10)
    want, _ = oracle.intersect(OA, OB, pa, pb)
    return both, len(A), pa, pb, want


def test_pixel_pair_limit_fails_before_refinement():
    both, na, pa, pb, want = _star_workload()
    dev = capi.Device(0)
    try:
        dev.upload_polygons(both)
        dev.upload_pairs(pa, pb + na)
        st = dev.intersect().stats
        total = st["bb_pixel_pairs"]
        assert total > 10
        dev.set_pixel_pair_limit(total - 1)
        with pytest.raises(capi.FarmError) as e:
            dev.intersect()
        assert e.value.code == capi.FARM_ECAPACITY
        dev.set_pixel_pair_limit(total)  # exactly at the limit: fine
        assert np.array_equal(dev.intersect().bits, want)
        dev.set_pixel_pair_limit(0)  # default again
        assert np.array_equal(dev.intersect().bits, want)
    finally:
        dev.close()


def test_workspace_grows_on_a_fresh_handle():
    # the first call sizes the per-call buffers for a 1-pair window; the
    # next windows overflow them, run no refinement, and are repeated with
    # grown buffers
    both, na, pa, pb, want = _star_workload(12, 400)
    dev = capi.Device(0)
    try:
        dev.upload_polygons(both)
        dev.upload_pairs(pa, pb + na)
        assert np.array_equal(dev.intersect(0, 1).bits, want[:1])
        n = len(pa)
        assert np.array_equal(dev.intersect(0, n).bits, want)
        assert np.array_equal(dev.intersect(n // 2, n, tile_a=2).bits, want[n // 2:])
    finally:
        dev.close()


def test_graph_replays_match_eager_calls():
    # every replay of a captured launch sequence against the eager result of
    # the same window, with other work and synchronisations in between
    both, na, pa, pb, want = _star_workload(13, 400)
    dev = capi.Device(0)
    try:
        dev.upload_polygons(both)
        dev.upload_pairs(pa, pb + na)
        n = len(pa)
        wins = [(0, n // 2), (n // 2, n), (n // 5, n - 3)]
        dev.set_graphs(False)
        eager = {w: dev.intersect(*w, tile_a=4).bits for w in wins}
        for w in wins:
            assert np.array_equal(eager[w], want[w[0]:w[1]])
        dev.set_graphs(True)
        for rep in range(4):
            for w in wins:
                got = dev.intersect(*w, tile_a=4).bits
                assert np.array_equal(got, eager[w]), f"replay {rep} of window {w}"
                dev.set_graphs(False)  # eager work between replays
                assert np.array_equal(dev.intersect(0, 7, tile_a=2).bits, want[:7])
                dev.set_graphs(True)
                dev.intersect(mode=capi.MODE_APPROX_HEURISTIC)
    finally:
        dev.close()
