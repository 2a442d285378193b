"""Pins the oracle's primitive restatements with closed forms and exact
arithmetic (CPU only)."""
import itertools
from fractions import Fraction

import numpy as np

import oracle


def test_double_to_int_closed_forms():
    # util.h:98-106: truncate, then round up when within 1e-8 below the next integer
    cases = {
        0.0: 0, 0.5: 0, 0.999999995: 1, 0.99999998: 0, 0.9999999: 0, 2.0: 2, 2.999999995: 3, 2.99999998: 2,
        7.000000001: 7, -0.5: 0, -0.999999999: 0, 123456.99999999999: 123457,
    }
    for v, want in cases.items():
        assert oracle.double_to_int(v) == want, v


def _orient(p, q, r):
    return (q[0] - p[0]) * (r[1] - p[1]) - (q[1] - p[1]) * (r[0] - p[0])


def _exact_segments_intersect(a, b, c, d):
    """Closed-segment intersection in exact rational arithmetic."""
    o1, o2, o3, o4 = _orient(a, b, c), _orient(a, b, d), _orient(c, d, a), _orient(c, d, b)

    def on(p, q, r):  # r on segment pq given collinear
        return min(p[0], q[0]) <= r[0] <= max(p[0], q[0]) and min(p[1], q[1]) <= r[1] <= max(p[1], q[1])

    if o1 == 0 and o2 == 0:  # collinear: 1-D overlap on both axes
        return (max(min(a[0], b[0]), min(c[0], d[0])) <= min(max(a[0], b[0]), max(c[0], d[0])) and
                max(min(a[1], b[1]), min(c[1], d[1])) <= min(max(a[1], b[1]), max(c[1], d[1])))
    s = lambda v: (v > 0) - (v < 0)
    if s(o1) != s(o2) and s(o3) != s(o4):
        return True
    return (o1 == 0 and on(a, b, c)) or (o2 == 0 and on(a, b, d)) or (o3 == 0 and on(c, d, a)) or \
        (o4 == 0 and on(c, d, b))


def test_segment_intersect_matches_exact_geometry_on_small_integers():
    # integer coordinates below 2^11: every f32 difference and product is exact,
    # so for segments of nonzero length the reference's binary32 predicate must
    # equal exact closed-segment intersection
    rng = np.random.default_rng(0)
    pts = rng.integers(-40, 41, size=(20000, 4, 2))
    # make a quarter of the cases collinear, another quarter share an endpoint
    pts[::4, 2] = pts[::4, 0] + (pts[::4, 1] - pts[::4, 0]) * rng.integers(-1, 3, size=(len(pts[::4]), 1))
    pts[1::4, 3] = pts[1::4, 1]
    checked = 0
    for a, b, c, d in pts:
        if (a == b).all() or (c == d).all():
            continue
        want = _exact_segments_intersect(*(tuple(int(v) for v in p) for p in (a, b, c, d)))
        assert oracle.segment_intersect(a, b, c, d) == want, (a, b, c, d)
        checked += 1
    assert checked > 19000


def test_segment_intersect_zero_length_edges():
    # a zero-length target edge c == d passes the collinear test for any a, b
    # and is then decided by bounding boxes alone (geometry_computation.h:217-218)
    assert oracle.segment_intersect((0, 0), (2, 2), (2, 0), (2, 0))
    assert not oracle.segment_intersect((0, 0), (2, 2), (3, 0), (3, 0))
    # a zero-length source edge off the target's line: orientations of c, d
    # against it are both zero, so the sign test fails
    assert not oracle.segment_intersect((1, 0), (1, 0), (0, -2), (2, 1))


def test_segment_intersect_hand_cases():
    assert oracle.segment_intersect((0, 0), (2, 2), (0, 2), (2, 0))          # proper crossing
    assert oracle.segment_intersect((0, 0), (1, 0), (1, 0), (2, 5))          # shared endpoint
    assert oracle.segment_intersect((0, 0), (2, 0), (1, 0), (1, 3))          # T junction
    assert oracle.segment_intersect((0, 0), (4, 0), (2, 0), (6, 0))          # collinear overlap
    assert not oracle.segment_intersect((0, 0), (1, 0), (2, 0), (3, 0))      # collinear disjoint
    assert not oracle.segment_intersect((0, 0), (1, 0), (0, 1), (1, 1))      # parallel
    assert not oracle.segment_intersect((0, 0), (1, 1), (2, 0), (3, -1))     # apart


def test_point_to_segment_distance_closed_forms():
    km_per_deg_lat = 40076.0 / 360.0
    # foot inside the segment, purely north-south offset
    d = oracle.point_to_segment_distance((1.0, 0.0), (0.0, 1.0), (2.0, 1.0))
    assert abs(d - km_per_deg_lat) < 1e-9 * km_per_deg_lat
    # projection clamps to the nearer endpoint; longitude degrees are scaled by
    # the table entry of the query point's latitude (1.0 -> index 10 = 0.008984)
    d = oracle.point_to_segment_distance((5.0, 1.0), (0.0, 1.0), (2.0, 1.0))
    assert abs(d - 3.0 / 0.008984) < 1e-9 * d
    # planar form (geography off) is the Euclidean distance
    assert abs(oracle.point_to_segment_distance((3.0, 4.0), (0.0, 0.0), (0.0, 0.0), False) - 5.0) < 1e-12


def test_longitude_km_table_entries():
    # the 900-entry degree-per-km table (util.h:73-74) is rebuilt from its
    # closed form; these entries are the reference's published values
    published = {0: 0.008983, 1: 0.008983, 100: 0.009122, 450: 0.012704, 600: 0.017966, 850: 0.103068,
                 898: 2.573426, 899: 5.146844}
    for k, v in published.items():
        lat = np.float32(k / 10 + 0.05)
        # zero-length segment: param = -1, foot = p1; 1 degree east
        d = oracle.point_to_segment_distance((1.0, lat), (0.0, lat), (0.0, lat))
        assert d == 1.0 / v, (k, d, 1.0 / v)


def test_longitude_km_table_all_entries_golden():
    # every entry of the reference's table (tests/golden/km_table.json, made
    # from the reference's own header by tests/golden/make_km_table.py):
    # the oracle's and the library's rebuilt tables equal it exactly
    import json
    import os

    from farm_amd import capi

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "km_table.json")
    ref = np.array(json.load(open(path))["values"], np.float64)
    assert ref.shape == (900,)
    lib = np.zeros(900, np.float64)
    assert capi.lib.farm_km_table(lib.ctypes.data) == 0
    assert np.array_equal(lib, ref)
    got = np.zeros(900)
    for k in range(900):
        # a query point inside latitude band k, one degree east of a point segment
        lat = np.float32(k / 10 + 0.05)
        got[k] = 1.0 / oracle.point_to_segment_distance((1.0, lat), (0.0, lat), (0.0, lat))
    # the oracle's distance divides by the entry: 1 / (1 / v) is v or its neighbour
    assert np.all(np.abs(got - ref) <= 2 * np.spacing(ref))
    # and exactly: 1 / v as the reference divides
    for k in range(900):
        lat = np.float32(k / 10 + 0.05)
        assert oracle.point_to_segment_distance((1.0, lat), (0.0, lat), (0.0, lat)) == 1.0 / ref[k], k
