"""Candidate pairs built on the device (farm_generate_pairs) against the host
R-tree join (farm_candidates, itself checked against brute force in
test_host_io): the same pairs in the same order, for intersect joins,
within self-joins with the 10 km expansion, skewed box sizes and degenerate
boxes; then a query over device-built pairs against the same query over
uploaded ones."""
import numpy as np
import pytest

from farm_amd import capi, datagen

pytestmark = pytest.mark.gpu


def _upload(device, *sets):
    both = sets[0] if len(sets) == 1 else capi.HostSet.concat(*sets)
    both.rasterize(10)
    device.upload_polygons(both)
    return both


def _same(device, want_a, want_b, a_base=0, b_base=0):
    got_a, got_b = device.download_pairs()
    print(f"  pairs {len(want_a)} (device {len(got_a)})", flush=True)
    assert np.array_equal(got_a, want_a.astype(np.uint32) + np.uint32(a_base))
    assert np.array_equal(got_b, want_b.astype(np.uint32) + np.uint32(b_base))


def test_intersect_join_matches_host(device):
    (sa, sxy), (ta, txy) = datagen.star_pair_set(31, 3000, 3000, 0.0, 60.0)
    A, B = capi.HostSet.from_rings(sa, sxy), capi.HostSet.from_rings(ta, txy)
    _upload(device, A, B)
    pa, pb = capi.candidates(A, B)
    n = device.generate_pairs(A.mbbs(), B.mbbs(), a_base=0, b_base=len(A))
    assert n == len(pa) > 1000
    _same(device, pa, pb, 0, len(A))


def test_within_self_join_matches_host(device):
    ro, xy = datagen.lake_set(32, 4000, 0.0, 3.0)
    H = capi.HostSet.from_rings(ro, xy)
    _upload(device, H)
    for d in (2.0, 10.0, 30.0):
        pa, pb = capi.candidates(H, H, expand_km=d, skip_self=True)
        assert device.generate_pairs(H.mbbs(), H.mbbs(), expand_km=d, skip_self=True) == len(pa)
        _same(device, pa, pb)


def test_skewed_and_degenerate_boxes(device):
    rng = np.random.default_rng(33)
    small = rng.random((3000, 2)) * 50.0
    boxes_a = np.concatenate([
        np.hstack([small, small + rng.random((3000, 2)) * 0.5]),
        [[0.0, 0.0, 50.0, 50.0], [-10.0, 10.0, 60.0, 10.0], [20.0, -5.0, 20.0, 55.0]],  # huge, zero-height, zero-width
        np.repeat([[5.0, 5.0, 6.0, 6.0]], 50, axis=0),                                   # identical boxes
    ])
    boxes_b = np.concatenate([
        np.hstack([small[::-1] - 0.25, small[::-1] + 0.25]),
        [[6.0, 6.0, 7.0, 7.0], [-3.0, -3.0, -1.0, -1.0], [50.0, 50.0, 50.0, 50.0]],      # corner touch, outside, point
    ])
    # a polygon set large enough for the indexes; the boxes are independent of it
    ro, xy = datagen.lake_set(34, len(boxes_a) + len(boxes_b), 0.0, 5.0)
    H = capi.HostSet.from_rings(ro, xy)
    _upload(device, H)
    n = device.generate_pairs(boxes_a, boxes_b)
    # brute force: inclusive overlap, ordered by (j, i)
    a, b = boxes_a[:, None, :], boxes_b[None, :, :]
    ov = ~((a[..., 0] > b[..., 2]) | (b[..., 0] > a[..., 2]) | (a[..., 1] > b[..., 3]) | (b[..., 1] > a[..., 3]))
    jj, ii = np.nonzero(ov.T)
    assert n == len(ii)
    _same(device, ii, jj)
    for aa, bb in ((boxes_a[:0], boxes_b), (boxes_a, boxes_b[:0])):
        assert device.generate_pairs(aa, bb) == 0


def test_errors(device):
    ro, xy = datagen.lake_set(35, 50, 0.0, 1.0)
    H = capi.HostSet.from_rings(ro, xy)
    _upload(device, H)
    m = H.mbbs()
    bad = m.copy()
    bad[3, 2] = np.nan
    inverted = m.copy()
    inverted[7, [0, 2]] = inverted[7, [2, 0]]
    for args, kw in (((m, bad), {}), ((inverted, m), {}), ((m, m), {"b_base": 1}), ((m, m), {"expand_km": -1.0})):
        with pytest.raises(capi.FarmError) as e:
            device.generate_pairs(*args, **kw)
        assert e.value.code == capi.FARM_EINVAL


def test_queries_over_device_pairs(device):
    (sa, sxy), (ta, txy) = datagen.star_pair_set(36, 400, 400, 0.0, 20.0)
    A, B = capi.HostSet.from_rings(sa, sxy), capi.HostSet.from_rings(ta, txy)
    _upload(device, A, B)
    pa, pb = capi.candidates(A, B)
    device.upload_pairs(pa, pb + len(A))
    want = device.intersect().bits
    wc, wa, _ = device.intersection_area()
    device.generate_pairs(A.mbbs(), B.mbbs(), b_base=len(A))
    assert np.array_equal(device.intersect().bits, want)
    c, a, _ = device.intersection_area()
    assert np.array_equal(c, wc) and np.array_equal(a.view(np.uint64), wa.view(np.uint64))
