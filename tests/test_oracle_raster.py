"""Pins the oracle's raster restatement (init_raster + rasterization) with
closed forms, exact-rational clipping and invariants, then checks the
product's host raster builder against it bit for bit (CPU only)."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from farm_amd import capi, datagen
from tests.helpers import rings_from_list


def _grid_closed_form(mbb, npix):
    """MyRaster::init_raster(int) (MyRaster.cpp:8-33) as a closed form."""
    w, h = mbb[2] - mbb[0], mbb[3] - mbb[1]
    multi = abs(h / w)
    dimx = int((npix / multi) ** 0.5)
    dimy = int(dimx * multi)
    dimx, dimy = max(dimx, 1), max(dimy, 1)
    sx, sy = w / dimx, h / dimy
    if sx < 1e-5:
        sx = 1e-5
        dimx = int(w / sx + 1)
    if sy < 1e-5:
        sy = 1e-5
        dimy = int(h / sy + 1)
    return dimx, dimy, sx, sy


def _subdivided_rect(x0, y0, w, h, k):
    t = np.arange(k) / k
    bottom = np.stack([x0 + w * t, np.full(k, y0)], 1)
    right = np.stack([np.full(k, x0 + w), y0 + h * t], 1)
    top = np.stack([x0 + w - w * t, np.full(k, y0 + h)], 1)
    left = np.stack([np.full(k, x0), y0 + h - h * t], 1)
    return np.vstack([bottom, right, top, left]).astype(np.float32)


@pytest.mark.parametrize("x0,y0,w,h,k,vpr", [(10.3, 20.7, 4.0, 3.0, 40, 1), (1.25, 3.5, 0.5, 2.0, 60, 2),
                                             (50.0, 50.0, 10.0, 10.0, 100, 4)])
def test_rectangle_closed_form(x0, y0, w, h, k, vpr):
    ring = _subdivided_rect(x0, y0, w, h, k)
    ro, xy = rings_from_list([ring])
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(vpr) == -1
    r = O.raster(0)
    mbb = O.mbb(0)
    bias = float(np.float32(1e-9))
    assert mbb[0] == float(ring[:, 0].min()) - bias and mbb[3] == float(ring[:, 1].max()) + bias
    nv = len(ring) + 2  # closing vertex + the copy of p[0] the decoder appends
    dimx, dimy, sx, sy = _grid_closed_form(mbb, nv // vpr)
    assert (r["dimx"], r["dimy"]) == (dimx, dimy) and r["step"][0] == sx and r["step"][1] == sy
    st = r["status"].reshape(dimy, dimx)
    want = np.full((dimy, dimx), 15, np.uint8)  # interior: IN
    want[0, :] = want[-1, :] = 14               # boundary ring: BORDER, nearly full
    want[:, 0] = want[:, -1] = 14
    assert np.array_equal(st, want)
    # areas: IN pixels full, boundary pixels full but for the MBB bias strip
    # (clip points are rounded to f32, hence the loose tolerance)
    areas = r["areas"].reshape(dimy, dimx)
    assert np.allclose(areas, sx * sy, rtol=1e-4)


def _clip(poly, lo, hi, axis, keep_low):
    """Sutherland-Hodgman against one axis-aligned half plane, exact."""
    out = []
    n = len(poly)
    inside = (lambda p: p[axis] <= hi) if keep_low else (lambda p: p[axis] >= lo)
    bound = hi if keep_low else lo
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        ip, iq = inside(p), inside(q)
        if ip:
            out.append(p)
        if ip != iq:
            t = (bound - p[axis]) / (q[axis] - p[axis])
            out.append(tuple(p[j] + t * (q[j] - p[j]) for j in range(2)))
    return out


def _exact_area_in_box(poly, bx):
    p = poly
    for axis, lo, hi in ((0, bx[0], bx[2]), (1, bx[1], bx[3])):
        p = _clip(p, lo, hi, axis, False)
        if not p:
            return Fraction(0)
        p = _clip(p, lo, hi, axis, True)
        if not p:
            return Fraction(0)
    a = sum(p[i][0] * p[(i + 1) % len(p)][1] - p[(i + 1) % len(p)][0] * p[i][1] for i in range(len(p)))
    return abs(a) / 2


def test_triangle_fill_classes_exact():
    # right triangle, legs subdivided so the grid has many pixels; the class of
    # every pixel follows from the exact clipped area (ceil(14 * ratio),
    # clamped to [1, 14] for BORDER pixels, 15 for IN, 0 for OUT)
    # (the hypotenuse is one edge; only the legs are subdivided)
    k = 90
    x0, y0, w, h = 3.3125, 7.40625, 6.1, 4.3
    legs = [(x0 + w * i / k, y0) for i in range(k)] + [(x0 + w, y0)] + [(x0, y0 + h)] + \
           [(x0, y0 + h - h * i / k) for i in range(1, k)]
    ring = np.array(legs, np.float32)
    ro, xy = rings_from_list([ring])
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(1) == -1
    r = O.raster(0)
    dimx, dimy = r["dimx"], r["dimy"]
    sx, sy = r["step"]
    lx, ly = r["mbr"][0], r["mbr"][1]
    tri = [(Fraction(float(a)), Fraction(float(b))) for a, b in ((x0, y0), (x0 + w, y0), (x0, y0 + h))]
    st = r["status"].reshape(dimy, dimx)
    checked = 0
    for yy in range(dimy):
        for xx in range(dimx):
            box = [Fraction(lx + xx * sx), Fraction(ly + yy * sy), Fraction(lx + (xx + 1) * sx),
                   Fraction(ly + (yy + 1) * sy)]
            ratio = _exact_area_in_box(tri, box) / (Fraction(sx) * Fraction(sy))
            s = int(st[yy, xx])
            if ratio == 0:
                assert s in (0, 1), (xx, yy, s)  # OUT, or a BORDER pixel the edge only grazes
                continue
            if ratio == 1:
                assert s in (14, 15), (xx, yy, s)
                continue
            q = float(ratio) * 14
            if abs(q - round(q)) < 1e-5:
                continue  # the f32 clip points may round across the class boundary
            assert s == min(14, max(1, int(np.ceil(q)))), (xx, yy, s, float(ratio))
            checked += 1
    assert checked > 20


def _pip(poly, x, y):
    inside = False
    n = len(poly)
    for i in range(n - 1):
        (x1, y1), (x2, y2) = poly[i], poly[i + 1]
        if (y1 > y) != (y2 > y):
            xi = x1 + (y - y1) * (x2 - x1) / (y2 - y1)
            if x < xi:
                inside = not inside
    return inside


@pytest.mark.parametrize("seed", [3, 4])
def test_raster_invariants_on_star_polygons(seed):
    (ro, xy), _ = datagen.star_pair_set(seed, 150, 1)
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(10) == -1
    for i in range(len(O)):
        r = O.raster(i)
        ring = xy[int(ro[i]):int(ro[i + 1])].astype(np.float64)
        dimx, dimy = r["dimx"], r["dimy"]
        sx, sy = r["step"]
        lx, ly = r["mbr"][:2]
        st = r["status"]
        # IN / OUT pixels: centre inside / outside the polygon
        for pid in np.nonzero((st == 0) | (st == 15))[0]:
            cx, cy = lx + (pid % dimx + 0.5) * sx, ly + (pid // dimx + 0.5) * sy
            assert _pip(ring, cx, cy) == (st[pid] == 15), (i, pid)
        # sum of clipped pixel areas equals the polygon area (clip points are
        # rounded to f32: ~1e-6 absolute per point at these coordinates)
        area = 0.5 * abs(np.dot(ring[:-1, 0], ring[1:, 1]) - np.dot(ring[1:, 0], ring[:-1, 1]))
        assert abs(r["areas"].sum() - area) < 1e-4 * area
        # every edge of nonzero length lies in some BORDER pixel's sequences
        nv = r["nverts"]
        covered = np.zeros(nv - 1, bool)
        po, sq = r["pixoff"], r["seqs"]
        for pid in np.nonzero((st > 0) & (st < 15))[0]:
            for s in range(po[pid], po[pid + 1]):
                covered[sq[s, 0]:sq[s, 0] + sq[s, 1]] = True
        ext = np.vstack([ring, ring[:1]])
        length = np.abs(np.diff(ext, axis=0)).sum(1)
        assert covered[length > 0].all(), i


def _compare_rasters(ro, xy, vpr):
    O = oracle.OracleSet(ro, xy)
    O.rasterize(vpr)
    H = capi.HostSet.from_rings(ro, xy)
    H.rasterize(vpr, 8)
    ok = H.raster_ok()
    assert np.array_equal(~ok, O.failed())
    L = H.layout()
    for i in range(len(H)):
        if not ok[i]:
            continue
        r, h = O.raster(i), L["hdr"][i]
        npix = int(h["dimx"]) * int(h["dimy"])
        assert (r["dimx"], r["dimy"]) == (int(h["dimx"]), int(h["dimy"])), i
        assert np.array_equal(r["mbr"], h["mbr"]) and np.array_equal(r["step"], h["step"]), i
        ps, po, sq = int(h["pix_start"]), int(h["pixoff_start"]), int(h["seq_start"])
        assert np.array_equal(L["status"][ps:ps + npix], r["status"]), i
        assert np.array_equal(L["pixoff"][po:po + npix + 1], r["pixoff"]), i
        assert np.array_equal(L["seqs"][sq:sq + int(h["nseqs"])], r["seqs"]), i
        vl, vn = int(h["vline_start"]), int(h["vnode_start"])
        assert np.array_equal(L["voff"][vl:vl + int(h["dimx"]) + 2], r["voff"]), i
        assert np.array_equal(L["vnodes"][vn:vn + len(r["vnodes"])], r["vnodes"]), i
        assert np.array_equal(H.areas(i, npix), r["areas"]), i
    return len(H)


@pytest.mark.parametrize("vpr", [1, 4, 10, 64])
def test_product_raster_matches_oracle_star(vpr):
    (ro, xy), _ = datagen.star_pair_set(11, 600, 1)
    assert _compare_rasters(ro, xy, vpr) == 600


def test_product_raster_matches_oracle_lakes_and_complex():
    ro, xy = datagen.lake_set(12, 400, hi=5.0)
    _compare_rasters(ro, xy, 256)
    _compare_rasters(ro, xy, 16)
    ro, xy = datagen.complex_set(13, 12, hi=5.0, nv_cap=8000)
    _compare_rasters(ro, xy, 256)
    _compare_rasters(ro, xy, 64)


def test_product_raster_matches_oracle_near_origin():
    # coordinates below 2^-6 take the tolerant (linear) dedup path
    (ro, xy), _ = datagen.star_pair_set(14, 100, 1, lo=0.0, hi=0.02)
    xy = (xy * np.float32(0.01)).astype(np.float32)
    _compare_rasters(ro, xy, 10)


def test_product_raster_matches_oracle_negative_coordinates():
    # western/southern hemisphere: the exact (hashed) dedup holds on both
    # sides of zero
    ro, xy = datagen.lake_set(15, 150, lo=-123.0, hi=-120.0)
    xy = (xy + np.float32([0.0, 80.0])).astype(np.float32)
    _compare_rasters(ro, xy, 10)
    _compare_rasters(ro, xy, 1)


def test_product_raster_matches_oracle_fine_pixels_far_from_origin():
    # pixels a few float spacings wide (x near 178, where one float step is
    # 2^-16): the edge buckets must still hold every edge a rounded cell bound
    # reaches
    ro, xy = datagen.lake_set(16, 60, lo=0.0, hi=1.0)
    xy = (xy.astype(np.float64) * 0.01 + [178.0, -60.0]).astype(np.float32)
    _compare_rasters(ro, xy, 1)
    _compare_rasters(ro, xy, 4)


@pytest.mark.parametrize("bits", [2, 3, 8])
def test_fill_classes_at_other_category_counts(bits):
    # -w b: every pixel's class at cc = 2^b from the same clipped area; a
    # subdivided rectangle's boundary ring is nearly full (class cc - 2), its
    # interior IN (cc - 1), as at 16 (14 and 15); OUT / BORDER / IN unchanged
    ring = _subdivided_rect(10.3, 20.7, 4.0, 3.0, 40)
    H = capi.HostSet.from_rings(*rings_from_list([ring]))
    H.rasterize(1)
    L = H.layout()
    assert L["fill"] is None and L["category_count"] == 16
    st = L["status"].copy()
    H.set_categories(bits)
    L = H.layout()
    cc = 1 << bits
    assert L["category_count"] == cc
    if cc == 16:
        assert L["fill"] is None
        return
    fill = L["fill"]
    assert np.array_equal(fill == 0, st == 0) and np.array_equal(fill == cc - 1, st == 15)
    assert np.array_equal(fill[(st > 0) & (st < 15)], np.full(((st > 0) & (st < 15)).sum(), cc - 2))
    # a right triangle: BORDER classes follow ceil((cc - 2) * ratio), clamped to [1, cc - 2]
    tri = np.array([[3.3125, 7.40625], [9.4125, 7.40625], [3.3125, 11.70625]], np.float32)
    T = capi.HostSet.from_rings(*rings_from_list([tri]))
    T.rasterize(1)
    T.set_categories(bits)
    Lt = T.layout()
    h = Lt["hdr"][0]
    npix = int(h["dimx"] * h["dimy"])
    areas = T.areas(0, npix)
    ratio = areas / (h["step"][0] * h["step"][1])
    stt = Lt["status"][:npix]
    border = (stt > 0) & (stt < 15)
    want = np.clip(np.ceil(ratio[border] * (cc - 2)), 1, cc - 2).astype(np.uint8)
    assert np.array_equal(Lt["fill"][:npix][border], want)
