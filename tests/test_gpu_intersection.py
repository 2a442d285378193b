"""GPU parity: intersection query (crossing points, boundary pieces and
their statuses, and the area of A n B) through the C ABI against the oracle.
Counts are bit-exact; areas agree to 1e-9 relative (the device sums the
piece paths in another order)."""
import numpy as np
import pytest

import oracle
from farm_amd import capi, datagen
from tests.helpers import Workload, check_wkt_areas, rings_from_list

pytestmark = pytest.mark.gpu

AREA_RTOL = 1e-9
AREA_ATOL = 1e-15  # deg^2: slivers whose area is pure rounding


def _check(dev, w: Workload, window=None):
    want_c, want_a = oracle.intersection(w.OA, w.OB, w.pa, w.pb)
    ok = w.both.raster_ok()
    assert np.array_equal(~ok, np.concatenate([w.OA.failed(), w.OB.failed()]))
    w.upload(dev)
    lo, hi = (0, len(w.pa)) if window is None else window
    print(f"  intersection pairs={len(w.pa)} window=({lo},{hi})", flush=True)
    c, a, st = dev.intersection_area(lo, hi)
    want_c, want_a = want_c[lo:hi], want_a[lo:hi]
    rows = np.nonzero((c != want_c).any(1))[0]
    assert rows.size == 0, f"{rows.size} count rows differ, first {rows[:5]}: {c[rows[:5]]} vs {want_c[rows[:5]]}"
    err = np.abs(a - want_a)
    bad = np.nonzero(err > AREA_RTOL * np.abs(want_a) + AREA_ATOL)[0]
    assert bad.size == 0, f"{bad.size} areas differ, first {bad[:5]}: {a[bad[:5]]} vs {want_a[bad[:5]]}"
    assert st["found"] == int((want_a > 0).sum())
    return want_c, want_a, st


def test_complex_x_lakes_v256(device):
    cro, cxy = datagen.complex_set(7, 12, 0.0, 2.0, nv_cap=20000)
    lro, lxy = datagen.lake_set(8, 400, 0.0, 2.0)
    w = Workload((cro, cxy), (lro, lxy), vpr=256)
    c, a, st = _check(device, w)
    assert (c[:, 0] > 0).sum() > 20 and (a > 0).sum() > 50


def test_chunk_skip_with_refinement_boxes_of_another_size(device):
    # after an intersect call with 32-segment chunks the refinement's boxes
    # no longer serve the intersection's 16-segment chunk skip: it builds its
    # own once per set, and the answers stay the oracle's
    cro, cxy = datagen.complex_set(7, 12, 0.0, 2.0, nv_cap=20000)
    lro, lxy = datagen.lake_set(8, 400, 0.0, 2.0)
    w = Workload((cro, cxy), (lro, lxy), vpr=256)
    want_c, want_a = oracle.intersection(w.OA, w.OB, w.pa, w.pb)
    w.upload(device)
    try:
        device.intersect(chunk=32)
        for _ in range(2):  # built on the first call, reused on the second
            c, a, _ = device.intersection_area()
            assert np.array_equal(c, want_c)
            assert (np.abs(a - want_a) <= AREA_RTOL * np.abs(want_a) + AREA_ATOL).all()
    finally:
        device.intersect(chunk=16)


def test_complex_x_lakes_v10_and_window(device):
    cro, cxy = datagen.complex_set(17, 8, 0.0, 1.5, nv_cap=8000)
    lro, lxy = datagen.lake_set(18, 300, 0.0, 1.5)
    w = Workload((cro, cxy), (lro, lxy), vpr=10)
    c, a, st = _check(device, w)
    assert (c[:, 0] > 0).sum() > 20
    _check(device, w, window=(len(w.pa) // 3, len(w.pa) // 3 + 200))


def test_star_pairs(device):
    src, tgt = datagen.star_pair_set(4, 300, 300, 0.0, 20.0)
    w = Workload(src, tgt, vpr=10)
    _check(device, w)


def _rect(x0, y0, x1, y1, n=1):
    t = np.arange(n) / n
    e = [((x0, y0), (x1, y0)), ((x1, y0), (x1, y1)), ((x1, y1), (x0, y1)), ((x0, y1), (x0, y0))]
    return np.vstack([np.stack([a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t], 1) for a, b in e]).astype(
        np.float32)


def test_hand_built_cases(device):
    # crossing bars, containment, disjoint, shared edges, identical squares
    a = [_rect(1.013, 1.401, 1.987, 1.611, 6), _rect(1.013, 1.017, 1.531, 1.329, 7),
         _rect(3.1, 3.1, 3.2, 3.3, 3), _rect(10, 10, 20, 20), _rect(1, 1, 1.5, 1.5)]
    b = [_rect(1.402, 1.017, 1.594, 1.993, 6), _rect(1.1, 1.1, 1.2, 1.2, 3), _rect(1.211, 0.893, 1.377, 1.611, 5),
         _rect(20, 12, 25, 18), _rect(1, 1, 1.5, 1.5)]
    n, m = len(a), len(b)
    pa, pb = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    w = Workload(rings_from_list(a), rings_from_list(b), vpr=10, pairs=(pa.ravel(), pb.ravel()))
    _check(device, w)


def test_rejected_polygons(device):
    # at one vertex per pixel this triangle puts a vertex on a grid corner
    # inside one pixel row, where the reference raster builder never ends:
    # its pairs answer -3
    tri = np.array([[50, 50], [60, 50], [55, 58]], np.float32)
    a = [tri, _rect(52, 49, 58, 52, 4)]
    b = [_rect(51, 51, 59, 53, 4), tri + np.float32(0.5)]
    pa, pb = np.meshgrid(np.arange(2), np.arange(2), indexing="ij")
    w = Workload(rings_from_list(a), rings_from_list(b), vpr=1, pairs=(pa.ravel(), pb.ravel()))
    c, _, _ = _check(device, w)
    assert (c[:, 0] == -3).sum() >= 2


def _diamond(cx, cy, r, n):
    t = np.arange(n) / n
    c = [(cx, cy - r), (cx + r, cy), (cx, cy + r), (cx - r, cy)]
    return np.vstack([np.stack([a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t], 1)
                      for a, b in zip(c, c[1:] + c[:1])]).astype(np.float32)


def test_touching_long_sequences(device):
    # shared vertices (t, u exactly 0 or 1), vertices on the other ring's
    # edges, collinear overlaps, a 1e-7 near miss and a sawtooth whose teeth
    # touch an edge: at -v 256 every ring is one long sequence (the chunked
    # screen), at -v 10 the sequences are short (the phased screen)
    a = [_rect(1, 1, 2, 2, 60), _rect(4, 1, 5, 2, 64)]
    xs = np.linspace(4.0, 5.0, 41)
    ys = np.where(np.arange(41) % 4 == 0, 1.0, np.where(np.arange(41) % 2 == 1, 0.9, 1.1))
    saw = np.vstack([np.stack([xs, ys], 1), [[5.0, 0.5], [4.0, 0.5]]]).astype(np.float32)
    b = [_rect(1.5, 0.5, 2.5, 1.5, 60), _rect(2.0, 1.25, 3.0, 1.75, 60), _rect(1.25, 2.0, 1.75, 3.0, 50),
         _rect(0.5, 1.3, 0.9999999, 1.7, 60), _diamond(4.5, 2.5, 0.5, 40), _diamond(4.0 + 1 / 128, 2.3, 0.3, 40),
         saw]
    n, m = len(a), len(b)
    pa, pb = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    for vpr in (256, 10):
        w = Workload(rings_from_list(a), rings_from_list(b), vpr=vpr, pairs=(pa.ravel(), pb.ravel()))
        c, _, _ = _check(device, w)
        assert (c[:, 0] > 0).sum() >= 4  # crossings were found, not just screened away


def _combs(n):
    # interleaved combs: every tooth of A crosses every tooth of B (n^2 points)
    xs = np.arange(n + 1) / n
    ya = np.where(np.arange(n + 1) % 2 == 0, 0.0, 1.0)
    a = np.vstack([np.stack([xs + 1.0, ya + 1.0], 1), [[2.0, 0.8], [1.0, 0.8]]])
    ys = (np.arange(n + 1) / n) * 0.98 + 0.0103
    xb = np.where(np.arange(n + 1) % 2 == 0, 0.0, 1.0) * 1.02 - 0.0107
    b = np.vstack([np.stack([xb + 1.0, ys + 1.0], 1), [[0.8, 1.9903], [0.8, 1.0103]]])
    return a.astype(np.float32), b.astype(np.float32)


def test_pairs_beyond_the_record_slabs(device):
    # 40,000 crossing records rerun with the 65,536-record slabs; 90,000 go
    # to the third pass, whose slabs are sized to the largest pair
    a2, b2 = _combs(200)
    a3, b3 = _combs(300)
    a = [a2, a3, _rect(1.2, 1.2, 1.4, 1.4, 9)]
    b = [b2, b3, _rect(1.3, 1.3, 1.5, 1.5, 9)]
    pa, pb = np.arange(3), np.arange(3)
    w = Workload(rings_from_list(a), rings_from_list(b), vpr=10, pairs=(pa, pb))
    c, _, st = _check(device, w)
    assert c[0, 0] == 40000 and c[1, 0] == 90000


def _wkt_lines(text):
    out = {}
    for line in text.splitlines():
        k, body = line.split("\t", 1)
        out[int(k)] = body
    return out


def test_intersection_polygons_of_rectangles(device):
    # A n B of axis-aligned rectangles is the rectangle of their overlap
    a = [np.array([[0, 0], [2, 0], [2, 2], [0, 2]], np.float32),
         np.array([[10, 10], [14, 10], [14, 14], [10, 14]], np.float32),
         np.array([[20, 20], [21, 20], [21, 21], [20, 21]], np.float32)]
    b = [np.array([[1, 1], [3, 1], [3, 3], [1, 3]], np.float32),
         np.array([[11, 11], [12, 11], [12, 12], [11, 12]], np.float32),   # inside a[1]
         np.array([[30, 30], [31, 30], [31, 31], [30, 31]], np.float32)]   # apart from a[2]
    w = Workload(rings_from_list(a), rings_from_list(b), vpr=1, pairs=([0, 1, 2], [0, 1, 2]))
    w.upload(device)
    counts, area, pcs, _ = device.intersection_pieces()
    text = capi.pieces_wkt(w.both, w.pa, w.pb + len(w.A), pcs)
    lines = _wkt_lines(text)
    assert sorted(lines) == [0, 1]
    ring0 = capi.wkt_rings(lines[0])
    assert len(ring0) == 1
    corners = {tuple(p) for p in ring0[0].tolist()}
    assert corners == {(1.0, 1.0), (2.0, 1.0), (2.0, 2.0), (1.0, 2.0)}
    assert abs(capi.wkt_multipolygon_area(lines[0]) - 1.0) < 1e-12 and abs(area[0] - 1.0) < 1e-12
    assert abs(capi.wkt_multipolygon_area(lines[1]) - 1.0) < 1e-12  # the contained square, whole
    assert area[2] == 0.0


@pytest.mark.parametrize("seed", [4, 5])
def test_intersection_polygons_areas_match(device, seed):
    # the assembled rings' shoelace area equals the device's Green's-theorem
    # area pair by pair (the same boundary, summed independently)
    src, tgt = datagen.star_pair_set(seed, 300, 300, 0.0, 20.0)
    w = Workload(src, tgt, vpr=10)
    w.upload(device)
    counts, area, pcs, _ = device.intersection_pieces()
    c2, a2, _ = device.intersection_area()
    assert np.array_equal(counts, c2) and np.array_equal(area.view(np.uint64), a2.view(np.uint64))
    lines = _wkt_lines(capi.pieces_wkt(w.both, w.pa, w.pb + len(w.A), pcs))
    nz = np.nonzero(area > 0)[0]
    assert len(nz) > 50
    check_wkt_areas({k: capi.wkt_multipolygon_area(t) for k, t in lines.items()},
                    {int(k): float(area[k]) for k in nz})
    assert set(lines) <= set(nz.tolist()) | set(np.nonzero(counts[:, 2] > 0)[0].tolist())
