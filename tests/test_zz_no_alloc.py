"""farm_intersect_noalloc (no device allocation, the library's fixed per-GPU
scratch) against farm_intersect_dev on the same windows: two calls in a row,
the second on a smaller window, each bit-exact, with the same counters of
the filter and the same count of intersecting pairs."""
import numpy as np
import pytest

from farm_amd import capi, datagen

pytestmark = pytest.mark.gpu


def _workload():
    (sa, sxy), (ta, txy) = datagen.star_pair_set(21, 600, 600, 0.0, 45.0)
    A, B = capi.HostSet.from_rings(sa, sxy), capi.HostSet.from_rings(ta, txy)
    pa, pb = capi.candidates(A, B)
    both = capi.HostSet.concat(A, B)
    both.rasterize(10)
    return both, len(A), pa, pb


def test_noalloc_matches_intersect_dev():
    from farm_amd.hipmem import DeviceBuffer

    both, na, pa, pb = _workload()
    n = len(pa)
    dev = capi.Device(0)
    ref = DeviceBuffer(n)
    got = DeviceBuffer(n)
    try:
        dev.upload_polygons(both)
        dev.upload_pairs(pa, pb + na)
        for (b, e), mode, tile in (((0, n), capi.MODE_EXACT_BATCHED, 8),
                                   ((n // 3, n // 3 + n // 4), capi.MODE_EXACT_BATCHED, 4),
                                   ((5, n // 2), capi.MODE_APPROX_HEURISTIC, 0),
                                   ((7, 200), capi.MODE_EXACT_PLAIN, 0)):
            print(f"window [{b},{e}) mode {mode} tile {tile}", flush=True)
            s_ref = dev.intersect_dev(ref.ptr, b, e, mode=mode, tile_a=tile)
            s_got = dev.intersect_noalloc(got.ptr, b, e, mode=mode, tile_a=tile)
            want = ref.to_host(np.uint8, e - b)
            have = got.to_host(np.uint8, e - b)
            diff = np.nonzero(want != have)[0]
            assert diff.size == 0, f"[{b},{e}) mode {mode}: {diff.size} pairs differ, first {diff[:10]}"
            # counters that do not depend on which wave finds a crossing first
            # (the executed segment tests of intersecting pairs do)
            for k in ("found", "in_hits", "bb_pixel_pairs", "filter_bytes"):
                assert s_ref[k] == s_got[k], (k, s_ref[k], s_got[k])
            assert 0 < s_got["found"] < e - b
        # beyond the scratch's pair capacity: a clear error, nothing launched
        with pytest.raises(capi.FarmError) as exc:
            dev.intersect_noalloc(None, 0, (1 << 18) + 1) if n > (1 << 18) else dev.intersect_noalloc(
                None, 0, n, mode=capi.MODE_APPROX_LEARNED)
        assert exc.value.code in (capi.FARM_ECAPACITY, capi.FARM_EINVAL)
    finally:
        ref.free()
        got.free()
        dev.close()
