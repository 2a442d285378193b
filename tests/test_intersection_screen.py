"""The line screen of the intersection kernel (k_intersection) passes every
edge pair that the reference's crossing test accepts.

The kernel skips an edge pair (a + s*d1, b_j .. b_j+1) when both
o_j = (b_j - a) x d1 and o_j+1 exceed m = 2^-19 * |d1|_1 * (R_j + R_j+1)
(R = |b - a|_1), or both fall below -m; for long sequences it uses one
margin per source edge, M = 2^-18 * |d1|_1 * Rmax >= m, with Rmax bounding
R over the target sequence's box, and o computed with one fused
multiply-add (one rounding fewer than the reference's c1).  The crossing test it guards
(segment_intersect_batch with intersection records,
geometry_computation.h:239-267) keeps the pair when |dn| >= 1e-12 and
t = c2 / dn, u = c1 / dn both lie in [-1e-9, 1 + 1e-9], where c1 = o_j and
dn = d1 x d2, every f32 product and difference rounded once.  Here both
are evaluated in numpy f32 / f64 (same roundings) on near-degenerate edge
pairs: shared end points, end points on the other edge, collinear and
nearly parallel edges, coordinates far from the origin and tiny edges."""
import numpy as np

F = np.float32


def _accepts(ax, ay, a1x, a1y, bx, by, b1x, b1y):
    d1x, d1y = a1x - ax, a1y - ay
    d2x, d2y = b1x - bx, b1y - by
    rx, ry = bx - ax, by - ay
    dn = d1x * d2y - d1y * d2x
    c2 = rx * d2y - ry * d2x
    c1 = rx * d1y - ry * d1x
    with np.errstate(divide="ignore", invalid="ignore"):
        den = dn.astype(np.float64)
        t = c2.astype(np.float64) / den
        u = c1.astype(np.float64) / den
    ok = np.abs(den) >= 1e-12
    return ok & (t >= -1e-9) & (t <= 1 + 1e-9) & (u >= -1e-9) & (u <= 1 + 1e-9)


def _screen(ax, ay, a1x, a1y, bx, by, b1x, b1y):
    d1x, d1y = a1x - ax, a1y - ay
    k = F(2.0 ** -19) * (np.abs(d1x) + np.abs(d1y))

    def probe(qx, qy):
        rx, ry = qx - ax, qy - ay
        return rx * d1y - ry * d1x, np.abs(rx) + np.abs(ry)

    o0, r0 = probe(bx, by)
    o1, r1 = probe(b1x, b1y)
    m = k * (r0 + r1)
    return ((o0 <= m) | (o1 <= m)) & ((o0 >= -m) | (o1 >= -m))


def _screen_box(ax, ay, a1x, a1y, bx, by, b1x, b1y):
    # the kernel's form for long sequences: one margin per source edge from
    # the target sequence's box (here the smallest possible box, the edge's
    # own two vertices), each vertex classified against it once
    d1x, d1y = a1x - ax, a1y - ay
    bx0, bx1 = np.minimum(bx, b1x), np.maximum(bx, b1x)
    by0, by1 = np.minimum(by, b1y), np.maximum(by, b1y)
    rmax = np.maximum(np.abs(bx0 - ax), np.abs(bx1 - ax)) + np.maximum(np.abs(by0 - ay), np.abs(by1 - ay))
    M = F(2.0 ** -18) * (np.abs(d1x) + np.abs(d1y)) * rmax

    def fused(qx, qy):
        # fmaf(qx - ax, d1y, -((qy - ay) * d1x)): one rounding of the exact
        # f32 x f32 product minus the rounded one (f64 holds the product)
        p = ((qy - ay) * d1x).astype(np.float64)
        return ((qx - ax).astype(np.float64) * d1y.astype(np.float64) - p).astype(F)

    o0, o1 = fused(bx, by), fused(b1x, b1y)
    return ~(((o0 > M) & (o1 > M)) | ((o0 < -M) & (o1 < -M)))


def _near_degenerate(rng, n, scale, offset):
    # end points on a coarse lattice (exact touches and collinear runs are
    # common), half of them nudged by a few ulps
    def pt():
        v = (rng.integers(0, 5, size=(n, 2)) * scale / 4 + offset).astype(F)
        nudge = (rng.integers(-3, 4, size=(n, 2)) * (rng.random((n, 2)) < 0.5)).astype(F)
        return (v + nudge * np.spacing(v)).astype(F)

    return [c.astype(F) for c in np.hstack([pt(), pt(), pt(), pt()]).T]


def test_screen_never_rejects_an_accepted_pair():
    rng = np.random.default_rng(11)
    accepted = 0
    for scale, offset in ((1.0, 0.0), (1e-3, 178.0), (0.5, -60.0), (1e-5, 37.5), (20.0, 1.0)):
        for _ in range(4):
            args = _near_degenerate(rng, 250_000, scale, offset)
            acc = _accepts(*args)
            for screen in (_screen, _screen_box):
                scr = screen(*args)
                assert not (acc & ~scr).any(), (screen.__name__, scale, offset, np.nonzero(acc & ~scr)[0][:5])
            accepted += int(acc.sum())
    assert accepted > 100_000  # the cases do reach the accept branch


def test_screen_rejects_most_pairs_of_a_chain():
    # the screen is worth running: against a wiggly chain most edge pairs fail it
    rng = np.random.default_rng(12)
    t = np.linspace(0, 8 * np.pi, 2001)
    cx = (t / 10).astype(F)
    cy = (0.3 * np.sin(t) + 0.05 * rng.standard_normal(t.size)).astype(F)
    ax = rng.random(500).astype(F) * F(2.5)
    ay = (rng.random(500).astype(F) - F(0.5)) * F(0.8)
    a1x, a1y = (ax + F(0.01)).astype(F), (ay + F(0.004)).astype(F)
    A = [np.repeat(v, 2000) for v in (ax, ay, a1x, a1y)]
    B = [np.tile(v, 500) for v in (cx[:-1], cy[:-1], cx[1:], cy[1:])]
    acc = _accepts(*A, *B)
    for screen in (_screen, _screen_box):
        scr = screen(*A, *B)
        assert not (acc & ~scr).any()
        assert scr.mean() < 0.05
