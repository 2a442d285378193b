"""Pins the oracle's intersection query (Ideal::intersection,
src/geometry/Ideal.cpp:1440-1683, up to its boundary-piece statuses) and the
area of A n B built on it, without the reference binary: closed forms for
rectangles, Sutherland-Hodgman clipping for convex pairs (an independent
algorithm), Monte Carlo estimates for star-shaped pairs, symmetry,
containment and disjoint cases."""
import numpy as np
import pytest

import oracle
from farm_amd import datagen
from tests.helpers import rings_from_list


def _rect(x0, y0, x1, y1, n=1):
    t = np.arange(n) / n
    e = [((x0, y0), (x1, y0)), ((x1, y0), (x1, y1)), ((x1, y1), (x0, y1)), ((x0, y1), (x0, y0))]
    return np.vstack([np.stack([a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t], 1) for a, b in e]).astype(
        np.float32)


def _run(polys, pairs, vpr=10):
    O = oracle.OracleSet(*rings_from_list(polys))
    assert O.rasterize(vpr) == -1
    pa = np.array([p[0] for p in pairs], np.uint32)
    pb = np.array([p[1] for p in pairs], np.uint32)
    return oracle.intersection(O, O, pa, pb)


def _shoelace(P):
    P = np.asarray(P, np.float64)
    Q = np.roll(P, -1, axis=0)
    return 0.5 * abs(np.sum(P[:, 0] * Q[:, 1] - Q[:, 0] * P[:, 1]))


def test_rectangles_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(40):
        x0, y0 = rng.uniform(1, 2, 2)
        w0, h0 = rng.uniform(0.1, 0.6, 2)
        x1, y1 = x0 + rng.uniform(-0.3, 0.3), y0 + rng.uniform(-0.3, 0.3)
        w1, h1 = rng.uniform(0.1, 0.6, 2)
        A = _rect(x0, y0, x0 + w0, y0 + h0, int(rng.integers(1, 9)))
        B = _rect(x1, y1, x1 + w1, y1 + h1, int(rng.integers(1, 9)))
        a32, b32 = A.astype(np.float64), B.astype(np.float64)
        ox = max(0.0, min(a32[:, 0].max(), b32[:, 0].max()) - max(a32[:, 0].min(), b32[:, 0].min()))
        oy = max(0.0, min(a32[:, 1].max(), b32[:, 1].max()) - max(a32[:, 1].min(), b32[:, 1].min()))
        exact = ox * oy
        counts, area = _run([A, B], [(0, 1), (1, 0)])
        for k in range(2):
            assert abs(area[k] - exact) <= 2e-6 * max(exact, 1e-3), (area[k], exact)


def test_plus_sign_counts():
    # two crossing bars: the 4 corners of the central square are the
    # boundary crossings; each ring splits into 4 pieces, alternately inside
    # and outside the other bar
    A = _rect(1.013, 1.401, 1.987, 1.611, 6)
    B = _rect(1.402, 1.017, 1.594, 1.993, 6)
    counts, area = _run([A, B], [(0, 1), (1, 0)])
    for k in range(2):
        assert counts[k].tolist() == [4, 8, 4, 4, 0]
    exact = (1.594 - 1.402) * (1.611 - 1.401)
    assert np.allclose(area, exact, rtol=2e-6)


def _regular(cx, cy, r, n, rot):
    a = rot + 2 * np.pi * np.arange(n) / n
    return np.stack([cx + r * np.cos(a), cy + r * np.sin(a)], 1).astype(np.float32)


def _sutherland_hodgman(subject, clip):
    out = [tuple(p) for p in np.asarray(subject, np.float64)]
    C = np.asarray(clip, np.float64)
    for i in range(len(C)):
        a, b = C[i], C[(i + 1) % len(C)]
        inp, out = out, []
        if not inp:
            break

        def inside(p):
            return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]) >= 0

        def cut(p, q):
            x1, y1, x2, y2 = p[0], p[1], q[0], q[1]
            x3, y3, x4, y4 = a[0], a[1], b[0], b[1]
            den = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4)
            t = ((x1 - x3) * (y3 - y4) - (y1 - y3) * (x3 - x4)) / den
            return (x1 + t * (x2 - x1), y1 + t * (y2 - y1))

        for j in range(len(inp)):
            p, q = inp[j], inp[(j + 1) % len(inp)]
            if inside(q):
                if not inside(p):
                    out.append(cut(p, q))
                out.append(q)
            elif inside(p):
                out.append(cut(p, q))
    return _shoelace(out) if len(out) >= 3 else 0.0


def test_convex_pairs_match_sutherland_hodgman():
    rng = np.random.default_rng(5)
    polys, pairs, want = [], [], []
    for k in range(60):
        A = _regular(*rng.uniform(1, 2, 2), rng.uniform(0.1, 0.4), int(rng.integers(5, 40)), rng.uniform(0, 6.3))
        B = _regular(*rng.uniform(1, 2, 2), rng.uniform(0.1, 0.4), int(rng.integers(5, 40)), rng.uniform(0, 6.3))
        polys += [A, B]
        pairs += [(2 * k, 2 * k + 1), (2 * k + 1, 2 * k)]
        w = _sutherland_hodgman(A, B)
        want += [w, w]
    counts, area = _run(polys, pairs)
    want = np.array(want)
    scale = np.array([min(_shoelace(polys[p[0]]), _shoelace(polys[p[1]])) for p in pairs])
    assert (np.abs(area - want) <= 1e-5 * scale).all(), np.max(np.abs(area - want) / scale)
    assert (want > 0).sum() > 40 and (want == 0).sum() > 5


def _inside(P, x, y):
    # even-odd crossing number, vectorised over query points
    P = np.asarray(P, np.float64)
    Q = np.roll(P, -1, axis=0)
    ins = np.zeros(len(x), bool)
    for (x1, y1), (x2, y2) in zip(P, Q):
        c = (y1 > y) != (y2 > y)
        xi = x1 + (y - y1) * (x2 - x1) / np.where(y2 != y1, y2 - y1, 1.0)
        ins ^= c & (x < xi)
    return ins


def test_star_pairs_against_monte_carlo_and_symmetry():
    ro, xy = datagen.star_polygons(np.random.default_rng(9), 40, 1.0, 3.0, np.full(40, 120), np.log(0.5), 0.3)
    rings = [xy[int(ro[i]):int(ro[i + 1]) - 1] for i in range(len(ro) - 1)]
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(10) == -1
    pa, pb = oracle.candidates(O, O)
    keep = pa < pb
    pa, pb = pa[keep], pb[keep]
    assert len(pa) > 20
    c1, a1 = oracle.intersection(O, O, pa, pb)
    c2, a2 = oracle.intersection(O, O, pb, pa)
    sizes = np.array([min(_shoelace(rings[i]), _shoelace(rings[j])) for i, j in zip(pa, pb)])
    assert (np.abs(a1 - a2) <= 1e-6 * sizes).all()
    assert (a1 <= sizes * (1 + 1e-6)).all()
    rng = np.random.default_rng(1)
    for k in range(len(pa)):
        P, Q = rings[pa[k]], rings[pb[k]]
        lo = np.maximum(P.min(0), Q.min(0)).astype(np.float64)
        hi = np.minimum(P.max(0), Q.max(0)).astype(np.float64)
        if (hi <= lo).any():
            assert a1[k] == 0
            continue
        n = 20000
        x = rng.uniform(lo[0], hi[0], n)
        y = rng.uniform(lo[1], hi[1], n)
        frac = (_inside(P, x, y) & _inside(Q, x, y)).mean()
        box = (hi - lo).prod()
        est, sd = frac * box, np.sqrt(max(frac * (1 - frac), 1.0 / n) / n) * box
        assert abs(a1[k] - est) <= 5 * sd + 1e-9, (k, a1[k], est, sd)


def test_containment_and_disjoint():
    big = _rect(1.013, 1.017, 1.531, 1.329, 7)
    small = _rect(1.1, 1.1, 1.2, 1.2, 3)
    far = _rect(3.1, 3.1, 3.2, 3.3, 3)
    counts, area = _run([big, small, far], [(0, 1), (1, 0), (0, 2), (2, 0)])
    assert np.allclose(area[:2], _shoelace(small.astype(np.float64)), rtol=1e-12)
    assert area[2] == 0 and area[3] == 0


def _inside_f64(px, py, x, y):
    # even-odd ray casting in f64
    c = np.zeros(np.shape(px), bool)
    for i in range(len(x) - 1):
        x0, y0, x1, y1 = x[i], y[i], x[i + 1], y[i + 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = x0 + (py - y0) * (x1 - x0) / (y1 - y0)
        c ^= ((y0 > py) != (y1 > py)) & (px < xi)
    return c


def _green_inside(x, y, X, Y, ox, oy):
    # sum over ring (x, y) of the cross products of its portions inside ring
    # (X, Y): every edge split at its f64 crossings with the other ring, each
    # piece tested at its midpoint
    x0, y0, x1, y1 = x[:-1, None], y[:-1, None], x[1:, None], y[1:, None]
    X0, Y0, X1, Y1 = X[None, :-1], Y[None, :-1], X[None, 1:], Y[None, 1:]
    d = (x1 - x0) * (Y1 - Y0) - (y1 - y0) * (X1 - X0)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = ((X0 - x0) * (Y1 - Y0) - (Y0 - y0) * (X1 - X0)) / d
        u = ((X0 - x0) * (y1 - y0) - (Y0 - y0) * (x1 - x0)) / d
    hit = (d != 0) & (t >= 0) & (t <= 1) & (u >= 0) & (u <= 1)
    mx, my, cr = [], [], []
    for e in range(len(x) - 1):
        ts = np.sort(np.concatenate([[0.0, 1.0], t[e][hit[e]]]))
        px, py = x[e] + ts * (x[e + 1] - x[e]), y[e] + ts * (y[e + 1] - y[e])
        mx.append((px[:-1] + px[1:]) / 2)
        my.append((py[:-1] + py[1:]) / 2)
        cr.append((px[:-1] - ox) * (py[1:] - oy) - (px[1:] - ox) * (py[:-1] - oy))
    mx, my, cr = np.concatenate(mx), np.concatenate(my), np.concatenate(cr)
    return cr[_inside_f64(mx, my, X, Y)].sum()


def test_area_against_an_independent_f64_green_sum():
    # Star pairs in a [0,20]^2 box (many crossings per pair), the oracle's
    # area against Green's theorem with f64 crossings and midpoint tests.
    # The oracle's crossing points are f32 (the reference's records), so
    # the two agree to about 1e-3 of the smallest areas; a piece classified
    # by a test point on the other ring's boundary is off by far more.  Pair
    # 788 of seed 5 has a piece whose middle sub-segment is 6e-7 long next
    # to a crossing: classified by that segment's midpoint its area was
    # 0.0378 instead of 0.0150.
    (sa, sxy), (ta, txy) = datagen.star_pair_set(5, 300, 300, 0.0, 20.0)
    OA, OB = oracle.OracleSet(sa, sxy), oracle.OracleSet(ta, txy)
    OA.rasterize(10)
    OB.rasterize(10)
    pa, pb = oracle.candidates(OA, OB)
    _, area = oracle.intersection(OA, OB, pa, pb)
    nz = np.nonzero(area > 0)[0]
    rng = np.random.default_rng(3)
    pick = np.unique(np.concatenate([[788], rng.choice(nz, size=120, replace=False)]))
    assert area[788] > 0
    for k in pick:
        i, j = int(pa[k]), int(pb[k])
        a = sxy[int(sa[i]):int(sa[i + 1])].astype(np.float64)
        b = txy[int(ta[j]):int(ta[j + 1])].astype(np.float64)
        ox, oy = max(a[:, 0].min(), b[:, 0].min()), max(a[:, 1].min(), b[:, 1].min())

        def orient(r):
            return np.sign(np.sum((r[:-1, 0] - ox) * (r[1:, 1] - oy) - (r[1:, 0] - ox) * (r[:-1, 1] - oy)))

        ref = abs(orient(a) * _green_inside(a[:, 0], a[:, 1], b[:, 0], b[:, 1], ox, oy) +
                  orient(b) * _green_inside(b[:, 0], b[:, 1], a[:, 0], a[:, 1], ox, oy)) / 2
        assert abs(area[k] - ref) <= 1e-3 * ref + 1e-9, (int(k), float(area[k]), float(ref))
