"""Pins the oracle's within-distance decision (Ideal::within,
src/geometry/Ideal.cpp:2475-2643) without the reference binary: closed forms
on hand-built squares, and invariants against a brute-force minimum over all
point-segment distances (the reference's own distance definition,
src/test/distance_test.cpp:36-44)."""
import numpy as np
import pytest

import oracle
from farm_amd import datagen
from tests.helpers import rings_from_list

KM_LON_05 = 0.008983  # published table entry for latitudes [0.5, 0.6) (util.h:73-74, index 5)
KM_LAT = 360.0 / 40076.0


def _subdivided_square(x0, y0, side, per_edge):
    t = np.arange(per_edge, dtype=np.float64) / per_edge
    pts = []
    for (ax, ay), (bx, by) in [((x0, y0), (x0 + side, y0)), ((x0 + side, y0), (x0 + side, y0 + side)),
                               ((x0 + side, y0 + side), (x0, y0 + side)), ((x0, y0 + side), (x0, y0))]:
        pts.append(np.stack([ax + (bx - ax) * t, ay + (by - ay) * t], 1))
    return np.vstack(pts).astype(np.float32)


def _within(polys, pairs, d=10.0, vpr=10):
    ro, xy = rings_from_list(polys)
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(vpr, hierarchy=True) == -1
    pa = np.array([p[0] for p in pairs], np.uint32)
    pb = np.array([p[1] for p in pairs], np.uint32)
    return oracle.within(O, O, pa, pb, d)


@pytest.mark.parametrize("gap_km", [0.5, 4.0, 9.0, 10.5, 14.0, 30.0])
def test_square_gap_closed_form(gap_km):
    # two 0.02 deg squares side by side at latitude 0.5; facing edges are
    # vertical, so the closest point pairs have dy = 0 and the distance is
    # the f32 gap over the table entry of the query point's latitude
    side = np.float32(0.02)
    x0 = np.float32(0.3)
    gap = np.float32(gap_km * KM_LON_05)
    x1 = np.float32(x0 + side + gap)
    a = _subdivided_square(x0, 0.5, side, 16)
    b = _subdivided_square(x1, 0.5, side, 16)
    exact = float(np.float64(b[:, 0].min()) - np.float64(a[:, 0].max())) / KM_LON_05
    bits, md, _ = _within([a, b], [(0, 1), (1, 0)])
    for k in range(2):
        assert bits[k] == (1 if exact < 10.0 else 0), (gap_km, bits, md, exact)
        if md[k] != 100000.0:  # the refinement ran: its minimum is the closed form
            assert md[k] == exact, (md[k], exact)
    if exact > 10.0:  # no box pair can be within d, so the refinement decides
        assert md[0] >= exact and md[1] >= exact


def test_polygon_within_itself():
    # identical polygons: either a BORDER box pair is within d by its far
    # corners, or the refinement meets distance 0 on the shared edges
    ro, xy = datagen.lake_set(3, 80, hi=2.0)
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(10, hierarchy=True) == -1
    idx = np.arange(len(O), dtype=np.uint32)
    bits, md, _ = oracle.within(O, O, idx, idx, 10.0)
    assert (bits == 1).all()
    assert np.isin(md, [0.0, 100000.0]).all()


def test_contained_square_far_from_boundary():
    # only BORDER pixels take part: a square deep inside another one is not
    # within 10 km of it (the boundaries are ~50 km apart)
    outer = _subdivided_square(0.013, 0.017, 1.0, 40)
    inner = _subdivided_square(0.49, 0.49, 0.04, 8)
    bits, md, st = _within([outer, inner], [(0, 1), (1, 0)])
    assert bits.tolist() == [0, 0]
    assert md.tolist() == [100000.0, 100000.0]  # pruned before any refinement
    # moved 5.6 km from the boundary, it is within
    near = _subdivided_square(0.063, 0.49, 0.04, 8)
    bits, _, _ = _within([outer, near], [(0, 1), (1, 0)])
    assert bits.tolist() == [1, 1]


def test_grid_aligned_edge_leaves_the_hierarchy():
    # merge_status (Ideal.cpp:1040-1059) classifies a coarse pixel by its
    # summed areas: a pixel the polygon fills completely is IN even when the
    # polygon's edge runs along its side.  With the square's edge x = 0 on a
    # grid line of every layer, no coarse BORDER pixel holds that edge, the
    # descent never reaches it, and the reference answers "not within" for a
    # square 5.6 km away.
    outer = _subdivided_square(0.0, 0.0, 1.0, 40)
    near = _subdivided_square(0.05, 0.48, 0.04, 8)
    bits, md, _ = _within([outer, near], [(0, 1), (1, 0)])
    assert bits.tolist() == [0, 0]
    assert md.tolist() == [100000.0, 100000.0]


def _pt_seg(p, a, b):
    # point_to_segment_distance with geography, vectorised over segments
    # (geometry_computation.h:49-89): f32 differences, f64 arithmetic
    A = (p[0] - a[:, 0]).astype(np.float64)
    B = (p[1] - a[:, 1]).astype(np.float64)
    C = (b[:, 0] - a[:, 0]).astype(np.float64)
    D = (b[:, 1] - a[:, 1]).astype(np.float64)
    dot = A * C + B * D
    ls = C * C + D * D
    with np.errstate(divide="ignore", invalid="ignore"):
        param = np.where(ls != 0, dot / np.where(ls != 0, ls, 1.0), -1.0)
    xx = np.where(param < 0, a[:, 0], np.where(param > 1, b[:, 0], a[:, 0] + param * C))
    yy = np.where(param < 0, a[:, 1], np.where(param > 1, b[:, 1], a[:, 1] + param * D))
    lat = abs(float(p[1]))
    lon = _KM_LON[int(lat * 10)]
    dx = (np.float64(p[0]) - xx) / lon
    dy = (np.float64(p[1]) - yy) / KM_LAT
    return np.sqrt(dx * dx + dy * dy)


def _km_lon_table():
    # 360 / (sin((90 - k/10) * PI / 180) * 40076), PI = 3.14159265, printed
    # with six decimals (util.h:41,76-85)
    k = np.arange(900)
    v = 360.0 / (np.sin((90 - k / 10.0) * 3.14159265 / 180) * 40076)
    return np.array([float(f"{x:.6f}") for x in v])


_KM_LON = _km_lon_table()


def _all_distances(P, Q):
    """Every point-segment distance between the rings P and Q, both ways."""
    out = []
    for p in P:
        out.append(_pt_seg(p, Q[:-1], Q[1:]))
    for q in Q:
        out.append(_pt_seg(q, P[:-1], P[1:]))
    return np.concatenate(out)


def test_decisions_bound_brute_force_minimum():
    ro, xy = datagen.lake_set(11, 160, hi=0.5)
    # keep the brute force small: lakes of at most 300 vertices
    sizes = np.diff(ro.astype(np.int64))
    keep = np.nonzero(sizes <= 300)[0][:60]
    ro, xy = datagen.subset(ro, xy, keep)
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(10, hierarchy=True) == -1
    pa, pb = oracle.candidates_within(O, 10.0)
    assert len(pa) > 200
    bits, md, _ = oracle.within(O, O, pa, pb, 10.0)
    rings = [np.vstack([xy[ro[i]:ro[i + 1]], xy[ro[i]:ro[i] + 1]]) for i in range(len(ro) - 1)]
    refined = 0
    for k in range(len(pa)):
        dist = _all_distances(rings[pa[k]], rings[pb[k]])
        bf = dist.min()
        if bits[k] == 1:
            assert bf <= 10.0 + 1e-3, (k, bf)
        if md[k] != 100000.0:
            refined += 1
            # the truncated sequences only drop evaluations, and the running
            # minimum is one of the point-segment distances
            assert md[k] >= bf
            assert np.isin(md[k], dist), (k, md[k])
        else:
            assert bits[k] == 1 or bf > 0  # pruned pairs never touch
    assert refined > 50
