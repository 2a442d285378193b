"""The refinement's exact prune of (source edge, target chunk) pairs.

k_refine (farm_amd/csrc/hip/intersect_kernels.hip) skips a source edge (a, b)
against a chunk of up to 16 target segments when the chunk's box lies on one
side of line ab by more than M = 2^-18 q |b - a|_1 (plus 2^-17 q^2 when the
chunk box overlaps the box of the lane's edge vertices), q >= the L1 reach
from those vertices to the chunk box.  The skip is only
allowed if the reference predicate (segment_intersect,
src/include/geometry_computation.h:215-221) is false for every segment of
the chunk.  These tests restate the box (k_chunk_box) and the test in f32
numpy and check that property against the oracle's segment_intersect on
random, collinear, nearly collinear, zero-length and mixed-scale inputs.
"""
import numpy as np

import oracle

f32 = np.float32
BIG = f32(2.0 ** 40)


def chunk_box(vx: np.ndarray, vy: np.ndarray):
    """Centre and rounded-up half extents of one chunk's vertices, infinite
    when the chunk holds a zero-length segment or a non-finite vertex."""
    vx = vx.astype(f32)
    vy = vy.astype(f32)
    open_ = bool(np.isfinite(vx).all() and np.isfinite(vy).all())
    same = (vx[1:] == vx[:-1]) & (vy[1:] == vy[:-1])
    open_ = open_ and not same.any()
    x0, x1, y0, y1 = vx.min(), vx.max(), vy.min(), vy.max()
    cx = f32(f32(f32(0.5) * x0) + f32(f32(0.5) * x1))
    cy = f32(f32(f32(0.5) * y0) + f32(f32(0.5) * y1))
    hx = f32(f32(max(f32(x1 - cx), f32(cx - x0)) * f32(1.0000003)) + f32(1.1754944e-38))
    hy = f32(f32(max(f32(y1 - cy), f32(cy - y0)) * f32(1.0000003)) + f32(1.1754944e-38))
    if not (open_ and f32(abs(cx) + hx) < BIG and f32(abs(cy) + hy) < BIG):
        hx = hy = f32(np.inf)
    return cx, cy, hx, hy


def fmaf(a, b, c) -> np.float32:
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def pruned_at(a, b, box, k: int = 18, tile=None) -> bool:
    """The kernel's per-edge prune decision, in f32, with its margin scaled to
    2^-k (the kernel uses k = 18).  The reach q is taken from the box of the
    vertices of the lane's edges (default: the edge's own two vertices, the
    smallest such box and so the most aggressive prune); the margin is
    2^-k q |b - a|_1 when that box and the chunk box are apart on an axis
    (the reference's collinear branch needs their exact overlap), else
    2^-k q (|b - a|_1 + 2q).  Cross products are taken relative to the
    lane's first vertex o (tile[0]): C = s.cross(a - o) per edge, then
    s.cross(centre - o) - C."""
    cx, cy, hx, hy = box
    ax, ay = f32(a[0]), f32(a[1])
    bx_, by_ = f32(b[0]), f32(b[1])
    pts = np.array([[ax, ay], [bx_, by_]] if tile is None else tile, f32)
    ox, oy = pts[0]
    lxl, lxh = pts[:, 0].min(), pts[:, 0].max()
    lyl, lyh = pts[:, 1].min(), pts[:, 1].max()
    bax, bay = f32(bx_ - ax), f32(by_ - ay)
    with np.errstate(all="ignore"):
        axo, ayo = f32(ax - ox), f32(ay - oy)
        cc = fmaf(bax, ayo, -f32(bay * axo))
        ex0, ex1 = f32(cx - lxl), f32(cx - lxh)
        ey0, ey1 = f32(cy - lyl), f32(cy - lyh)
        q = f32(f32(f32(max(abs(ex0), abs(ex1)) + max(abs(ey0), abs(ey1))) + f32(hx + hy)) * f32(1.000001))
        hx1, hy1 = f32(hx * f32(1.000001)), f32(hy * f32(1.000001))
        apart = (ex1 > hx1) or (-ex0 > hx1) or (ey1 > hy1) or (-ey0 > hy1)
        k1 = f32(q * f32(2.0 ** -k))
        k2 = f32(2.0 ** -80) if apart else fmaf(q, f32(q * f32(2.0 ** (1 - k))), 2.0 ** -80)
        hxk, hyk = f32(hx + k1), f32(hy + k1)
        cxo, cyo = f32(cx - ox), f32(cy - oy)
        oc = fmaf(bax, cyo, -fmaf(bay, cxo, cc))
        thr = fmaf(abs(bay), hxk, fmaf(abs(bax), hyk, k2))
        return bool(abs(oc) > thr)


def pruned(a, b, box) -> bool:
    return pruned_at(a, b, box)


def check(a, b, chunk: np.ndarray) -> bool:
    """True when the prune fires; asserts the reference finds no crossing then."""
    box = chunk_box(chunk[:, 0], chunk[:, 1])
    if not pruned(a, b, box):
        return False
    for k in range(len(chunk) - 1):
        assert not oracle.segment_intersect(a, b, chunk[k], chunk[k + 1]), (a, b, chunk[k], chunk[k + 1])
    return True


def polyline(rng, n, start, step):
    pts = np.cumsum(rng.normal(0, step, size=(n, 2)), axis=0) + start
    return pts.astype(f32)


def test_random_chunks_prune_only_where_reference_is_false():
    rng = np.random.default_rng(11)
    fired = 0
    for _ in range(3000):
        chunk = polyline(rng, 17, rng.uniform(0, 40, 2), 0.02)
        a = rng.uniform(0, 40, 2).astype(f32)
        if rng.random() < 0.5:  # a source edge near the chunk
            a = (chunk[rng.integers(17)] + rng.normal(0, 0.05, 2)).astype(f32)
        b = (a + rng.normal(0, 0.03, 2)).astype(f32)
        fired += check(a, b, chunk)
    assert fired > 1000  # the prune is not vacuous


def test_crossing_chunks_are_never_pruned():
    rng = np.random.default_rng(12)
    for _ in range(2000):
        chunk = polyline(rng, 17, rng.uniform(0, 40, 2), 0.02)
        k = rng.integers(16)
        mid = (chunk[k] + chunk[k + 1]) / 2
        d = rng.normal(0, 0.01, 2)
        a, b = (mid - d).astype(f32), (mid + d).astype(f32)
        box = chunk_box(chunk[:, 0], chunk[:, 1])
        if oracle.segment_intersect(a, b, chunk[k], chunk[k + 1]):
            assert not pruned(a, b, box)


def test_collinear_and_nearly_collinear():
    # target vertices on (or one ulp off) the line through a far source edge,
    # where f32 orientation signs are noise and the collinear branch applies
    rng = np.random.default_rng(13)
    for _ in range(1500):
        o = rng.uniform(-50, 50, 2)
        u = rng.normal(0, 1, 2)
        u /= np.linalg.norm(u)
        t = np.sort(rng.uniform(-5, 5, 19))
        pts = (o + t[:, None] * u).astype(f32)
        jig = rng.integers(-1, 2, size=pts.shape)
        away = np.nextafter(pts, np.where(jig > 0, np.inf, -np.inf).astype(f32))
        pts = np.where(jig != 0, away, pts).astype(f32)
        a, b = pts[0], pts[1]
        check(a, b, pts[2:])
        check(b, a, pts[2:])


def test_chunks_a_few_ulps_beside_the_source_line():
    # target chunks beyond the source edge along its line, every vertex
    # strictly on one side by about an ulp (exact f64 orientations): at
    # coordinates and lengths near 1 the reference's f32 orientation signs
    # are noise there and it reports crossings that a side test without the
    # margin would skip
    rng = np.random.default_rng(16)
    ref_true = tried = 0
    while tried < 3000:
        a = rng.uniform(-1, 1, 2).astype(f32)
        u = rng.normal(0, 1, 2)
        u /= np.linalg.norm(u)
        b = (a + rng.uniform(0.3, 1) * u).astype(f32)
        d = b.astype(np.float64) - a
        nrm = np.array([-d[1], d[0]]) / np.linalg.norm(d)
        t = np.sort(rng.uniform(1.2, 3, 17))
        off = rng.uniform(0.02, 1.0) * float(np.spacing(f32(4)))
        chunk = (a + t[:, None] * d + off * nrm).astype(f32)
        e = d[0] * (chunk[:, 1] - np.float64(a[1])) - d[1] * (chunk[:, 0] - np.float64(a[0]))
        if not (e > 0).all():
            continue
        tried += 1
        check(a, b, chunk)
        ref_true += any(oracle.segment_intersect(a, b, chunk[k], chunk[k + 1]) for k in range(16))
    assert ref_true > 20  # the reference does cross here


def test_zero_length_target_segments_are_never_pruned():
    # a repeated vertex inside the source edge's box but off its line: the
    # reference reports a crossing (collinear branch, box overlap only)
    a = np.array([0.0, 0.0], f32)
    b = np.array([1.0, 1.0], f32)
    c = np.array([0.9, 0.1], f32)
    chunk = np.array([[0.8, 0.05], c, c, [0.95, 0.12]], f32)
    assert oracle.segment_intersect(a, b, c, c)
    assert not pruned(a, b, chunk_box(chunk[:, 0], chunk[:, 1]))
    rng = np.random.default_rng(14)
    for _ in range(1000):
        chunk = polyline(rng, 17, rng.uniform(0, 10, 2), 0.05)
        k = rng.integers(1, 17)
        chunk[k] = chunk[k - 1]
        a = rng.uniform(0, 10, 2).astype(f32)
        b = (a + rng.normal(0, 0.3, 2)).astype(f32)
        assert np.isinf(chunk_box(chunk[:, 0], chunk[:, 1])[2])
        check(a, b, chunk)


def test_scales_and_zero_length_source_edges():
    rng = np.random.default_rng(15)
    for scale in (1e-6, 1e-3, 1.0, 1e3, 1e6):
        for _ in range(300):
            chunk = polyline(rng, 17, rng.uniform(-1, 1, 2) * scale, 0.01 * scale)
            a = (chunk[rng.integers(17)] + rng.normal(0, 0.02 * scale, 2)).astype(f32)
            b = a.copy() if rng.random() < 0.2 else (a + rng.normal(0, 0.02 * scale, 2)).astype(f32)
            check(a, b, chunk)
            if (a == b).all():  # a point source edge is never pruned
                assert not pruned(a, b, chunk_box(chunk[:, 0], chunk[:, 1]))


def test_non_finite_and_huge_boxes_are_never_pruned():
    chunk = np.array([[0, 0], [1, np.nan], [2, 0]], f32)
    assert np.isinf(chunk_box(chunk[:, 0], chunk[:, 1])[2])
    chunk = np.array([[3e38, 0], [3.1e38, 1]], f32)
    assert np.isinf(chunk_box(chunk[:, 0], chunk[:, 1])[2])
    box = chunk_box(chunk[:, 0], chunk[:, 1])
    assert not pruned(np.array([0, 5], f32), np.array([1, 5], f32), box)



def test_constructed_rounding_crossings_need_the_margin():
    # Constructed, not sampled: a target segment (c, d) that lies strictly on
    # one side of the source line ab in exact arithmetic, c within a few
    # rounding units of the line beside the start of ab, d straight out on
    # the same side.  Where the reference's binary32 a.cross(b, c) rounds to
    # 0 or to the wrong sign it reports a crossing (a and b are on opposite
    # sides of line cd), so the prune must keep the chunk.  The kernel's
    # margin (2^-18) keeps every such chunk, even with q taken from a alone;
    # a margin of 2^-27 already drops some, so this test fails for a margin
    # shrunk that far.
    rng = np.random.default_rng(21)
    ref_true, dropped_small = 0, 0
    for _ in range(60000):
        a = rng.uniform(-1, 1, 2).astype(f32)
        u = rng.normal(0, 1, 2)
        u /= np.linalg.norm(u)
        b = (a + rng.uniform(0.5, 1.5) * u).astype(f32)
        dab = b.astype(np.float64) - a
        nrm = np.array([-dab[1], dab[0]]) / np.linalg.norm(dab)
        t = 10 ** rng.uniform(-4, -1)
        c = (a + t * dab + rng.uniform(-3, 3) * 4e-8 * t * nrm).astype(f32)
        d = (c.astype(np.float64) + t * rng.uniform(0.01, 0.3) * nrm).astype(f32)

        def side(p):  # exact in f64 for these f32 inputs up to a few ulps; sign only
            return dab[0] * (np.float64(p[1]) - np.float64(a[1])) - dab[1] * (np.float64(p[0]) - np.float64(a[0]))

        if not (side(c) > 0 and side(d) > 0):
            continue
        if not oracle.segment_intersect(a, b, c, d):
            continue
        ref_true += 1
        box = chunk_box(np.array([c[0], d[0]]), np.array([c[1], d[1]]))
        # q from a alone: the smallest reach any tile holding this edge can
        # give, so the most aggressive prune the kernel could make
        point = np.array([a, a], f32)
        assert not pruned_at(a, b, box, 18), (a, b, c, d)
        assert not pruned_at(a, b, box, 18, tile=point), (a, b, c, d)
        dropped_small += pruned_at(a, b, box, 27, tile=point)
    assert ref_true > 200  # the construction reaches the rounding regime
    assert dropped_small > 0  # and a too-small margin would drop reported crossings


def test_lane_tiles_with_the_origin_away_from_the_edge():
    # the kernel takes cross products relative to the lane's first vertex:
    # every edge of an 8-edge lane tile (not only the first) against chunks
    # near and on its line, at coordinates up to 1e3 from the origin of the
    # plane; a prune must never hide a reference crossing
    rng = np.random.default_rng(31)
    fired = 0
    for _ in range(1500):
        centre = rng.uniform(-1e3, 1e3, 2)
        lane = polyline(rng, 9, centre, 0.01)
        e = rng.integers(8)
        a, b = lane[e], lane[e + 1]
        if rng.random() < 0.5:  # chunk along the edge's line, a few ulps off
            d = b.astype(np.float64) - a
            t = np.sort(rng.uniform(-3, 3, 17))
            chunk = (a + t[:, None] * d).astype(f32)
            jig = rng.integers(-1, 2, size=chunk.shape)
            chunk = np.where(jig != 0, np.nextafter(chunk, np.where(jig > 0, np.inf, -np.inf).astype(f32)),
                             chunk).astype(f32)
        else:
            chunk = polyline(rng, 17, a + rng.normal(0, 0.03, 2), 0.005)
        box = chunk_box(chunk[:, 0], chunk[:, 1])
        if pruned_at(a, b, box, tile=lane):
            fired += 1
            for k in range(16):
                assert not oracle.segment_intersect(a, b, chunk[k], chunk[k + 1]), (a, b, chunk[k], chunk[k + 1])
    assert fired > 200
